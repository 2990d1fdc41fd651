// Renderer kernels, part 1: tile binning. Together with composite.hip they
// replace the reference's draw_indirect of 6 vertices per instance through
// the fixed-function rasterizer (src/renderer.rs:163-195, 284-308).
//
//   bin       one thread per drawn instance, in sorted (back-to-front) order:
//             count the screen tiles its footprint touches, exclusive-scan the
//             counts across the grid (decoupled look-back) and emit
//             (tile id, Gaussian index) pairs. Pairs come out in draw order, so
//             a stable sort by tile id keeps every tile's list in draw order.
//   ranges    start/end of every tile's run in the tile-sorted pairs.
#include <hip/hip_runtime.h>

#include "device_util.hpp"
#include "kernels.hpp"
#include "types.hpp"

namespace gs3d {

constexpr int kBinThreads = 256;
// instances per thread of a binning tile
#ifndef GS3D_BIN_ITEMS
#define GS3D_BIN_ITEMS 8
#endif
constexpr int kBinItems = GS3D_BIN_ITEMS;
constexpr uint32_t kBinTile = kBinThreads * kBinItems;

uint32_t bin_max_tiles(uint32_t max_instances) { return (max_instances + kBinTile - 1) / kBinTile; }

__global__ __launch_bounds__(256) void k_bin_reset(uint32_t* ranges, uint32_t range_words, uint32_t* status,
                                                   uint32_t max_bin_tiles, uint32_t* ticket, uint32_t* pair_count,
                                                   uint32_t* sort_hist, uint32_t hist_words) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t j = gid; j < range_words; j += stride) ranges[j] = 0u;
  for (uint32_t j = gid; j < max_bin_tiles; j += stride) status[j] = 0u;
  for (uint32_t j = gid; j < hist_words; j += stride) sort_hist[j] = 0u;
  if (gid == 0) {
    *ticket = 0u;
    *pair_count = 0u;
  }
}

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
  const uint32_t lane = threadIdx.x % kWave;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, kWave);
    if (lane >= uint32_t(off)) x += y;
  }
  return x;
}

__global__ __launch_bounds__(kBinThreads) void k_bin(BinArgs a) {
  __shared__ uint32_t s_wtot[kBinItems][kBinThreads / kWave];  // each wave's count per item
  __shared__ uint32_t s_total;
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_base;
  __shared__ uint32_t s_hist[256];  // digit 0 of this block's stored tile ids
  if (a.sort_hist) s_hist[threadIdx.x] = 0u;

  const uint32_t t = take_ticket(a.ticket, &s_tile);
  const uint32_t v = a.indirect_args[1];
  const uint32_t num_tiles = (v + kBinTile - 1) / kBinTile;
  if (t >= num_tiles) return;
  const uint32_t tid = threadIdx.x;
  const uint32_t base = t * kBinTile;

  uint32_t cnt[kBinItems], excl[kBinItems], gidx[kBinItems];
  uint2 rr[kBinItems];  // packed rectangles (tile_rect)
  // every item's index, then every item's rectangle: unconditional loads
  // (an item past the end reads instance v - 1 and is dropped below), so the
  // wave waits once for each set instead of twice per item
  const uint32_t last = v - 1u;  // v > 0: this tile holds an instance
#pragma unroll
  for (int k = 0; k < kBinItems; ++k) {
    const uint32_t s = base + uint32_t(k) * kBinThreads + tid;
    gidx[k] = a.indices[s < v ? s : last];
  }
  // an index outside the model (a sort that gave up waiting leaves such
  // entries) is reported, and neither read through nor listed
  uint32_t bad = 0;
  if (a.rects) {
#pragma unroll
    for (int k = 0; k < kBinItems; ++k) {  // 8 bytes per instance written by the vertex work
      const bool b = gidx[k] >= a.num_gaussians;
      bad |= uint32_t(b) << k;
      rr[k] = a.rects[b ? 0u : gidx[k]];
    }
  } else {
    const float sd = a.fc->max_std_dev;
#pragma unroll
    for (int k = 0; k < kBinItems; ++k) {
      const bool b = gidx[k] >= a.num_gaussians;
      bad |= uint32_t(b) << k;
      rr[k] = tile_rect(a.splats[b ? 0u : gidx[k]], sd * sd, a.grid);
    }
  }
#pragma unroll
  for (int k = 0; k < kBinItems; ++k) {
    const uint32_t s = base + uint32_t(k) * kBinThreads + tid;
    const bool listed = s < v && ((bad >> k) & 1u) == 0u;
    if (s < v && !listed) atomicOr(a.err, kErrSplatIndex);
    cnt[k] = listed ? rect_tiles(unpack_rect(rr[k])) : 0u;
  }
  // exclusive scan in (item, thread) order == sorted instance order: a wave
  // scan per item, then a table of the waves' totals that one wave turns into
  // offsets (two barriers; each thread then reads its waves' 8 offsets
  // instead of summing the whole table, which held 32 more registers)
  const uint32_t w = tid / kWave;
#pragma unroll
  for (int k = 0; k < kBinItems; ++k) {
    const uint32_t inc = wave_inclusive_scan(cnt[k]);
    excl[k] = inc - cnt[k];
    if (tid % kWave == kWave - 1) s_wtot[k][w] = inc;
  }
  __syncthreads();
  // the first wave turns the table into offsets: (item, wave) order is the
  // instances' order
  constexpr uint32_t kCells = kBinItems * (kBinThreads / kWave);
  static_assert(kCells <= kWave, "one lane per table cell");
  if (tid < kWave) {
    const uint32_t c = tid < kCells ? (&s_wtot[0][0])[tid] : 0u;
    const uint32_t inc = wave_inclusive_scan(c);
    if (tid < kCells) (&s_wtot[0][0])[tid] = inc - c;
    if (tid == kCells - 1) s_total = inc;
  }
  __syncthreads();
  const uint32_t run = s_total;
#pragma unroll
  for (int k = 0; k < kBinItems; ++k) excl[k] += s_wtot[k][w];
  if (tid < kWave) {
    // the first wave looks back (run is the same in every thread)
    uint32_t e = 0;
    if (t == 0) {
      if (tid == 0) status_store(a.status, kFlagInclusive | run);
    } else {
      if (tid == 0) status_store(a.status + t, kFlagAggregate | run);
      e = lookback_exclusive_wave(a.status, t, a.err);
      if (tid == 0) status_store(a.status + t, kFlagInclusive | (e + run));
    }
    if (tid == 0) {
      s_base = e;
      if (t == num_tiles - 1) {
        *a.pair_count = e + run;
        if (e + run > a.capacity) atomicOr(a.err, kErrTileOverflow);
      }
    }
  }
  __syncthreads();
  const uint32_t off0 = s_base;
#pragma unroll
  for (int k = 0; k < kBinItems; ++k) {
    if (cnt[k] == 0u) continue;
    uint32_t p = off0 + excl[k];
    const TileRect r = unpack_rect(rr[k]);
    for (uint32_t ty = r.ty0; ty <= r.ty1; ++ty)
      for (uint32_t tx = r.tx0; tx <= r.tx1; ++tx) {
        if (r.small && !((r.mask >> (4u * (ty - r.ty0) + (tx - r.tx0))) & 1u)) continue;
        if (p < a.capacity) {
          const uint32_t key = ty * a.grid.tiles_x + tx;
          a.pair_keys[p] = key;
          a.pair_vals[p] = gidx[k];
          if (a.sort_hist) atomicAdd(&s_hist[key & 0xFFu], 1u);
        }
        ++p;
      }
  }
  if (a.sort_hist) {
    __syncthreads();
    if (s_hist[tid]) atomicAdd(&a.sort_hist[tid], s_hist[tid]);
  }
}

__global__ __launch_bounds__(256) void k_ranges(const uint32_t* __restrict__ pair_count, uint32_t capacity,
                                                const uint32_t* __restrict__ keys, uint32_t* __restrict__ ranges,
                                                uint32_t num_tiles, uint32_t* pair_max, uint32_t* host_total,
                                                uint32_t* err) {
  const uint32_t total = *pair_count;
  // the largest pair total since the host last looked, for it to check the
  // lists' capacity without a copy
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const uint32_t m = atomicMax(pair_max, total);
    if (host_total) *host_total = m > total ? m : total;
  }
  const uint32_t p = total < capacity ? total : capacity;
  // four consecutive keys per thread and step (one 16-byte load, plus the
  // neighbours either side), so few dependent steps cover the pairs
  constexpr uint32_t kNone = 0xFFFFFFFFu;  // before the first / after the last key
  const uint32_t quads = (p + 3u) / 4u;
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < quads; q += gridDim.x * blockDim.x) {
    const uint32_t j0 = q * 4u;
    uint32_t k[4];
    if (j0 + 4u <= p) {
      const uint4 v = *reinterpret_cast<const uint4*>(keys + j0);
      k[0] = v.x;
      k[1] = v.y;
      k[2] = v.z;
      k[3] = v.w;
    } else {
#pragma unroll
      for (uint32_t t = 0; t < 4u; ++t) k[t] = j0 + t < p ? keys[j0 + t] : kNone;
    }
    const uint32_t prev = j0 > 0u ? keys[j0 - 1u] : kNone;
    const uint32_t next = j0 + 4u < p ? keys[j0 + 4u] : kNone;
#pragma unroll
    for (uint32_t t = 0; t < 4u; ++t) {
      const uint32_t j = j0 + t;
      if (j >= p) break;
      if (k[t] >= num_tiles) {  // never written by k_bin: report, do not index
        atomicOr(err, kErrTileKey);
        continue;
      }
      const uint32_t before = t == 0u ? prev : k[t - 1u];
      const uint32_t after = j + 1u < p ? (t < 3u ? k[t + 1u] : next) : kNone;
      if (before != k[t]) ranges[2u * k[t]] = j;
      if (after != k[t]) ranges[2u * k[t] + 1u] = j + 1u;
    }
  }
}

void launch_bin_reset(hipStream_t st, uint32_t* ranges, uint32_t num_tiles, uint32_t* status, uint32_t max_bin_tiles,
                      uint32_t* ticket, uint32_t* pair_count, uint32_t* sort_hist, uint32_t hist_words) {
  uint32_t words = num_tiles * 2u;
  uint32_t m = words > max_bin_tiles ? words : max_bin_tiles;
  uint32_t blocks = (m + 255u) / 256u;
  if (blocks > 1024u) blocks = 1024u;
  if (blocks < 1u) blocks = 1u;
  hipLaunchKernelGGL(k_bin_reset, dim3(blocks), dim3(256), 0, st, ranges, words, status, max_bin_tiles, ticket,
                     pair_count, sort_hist, hist_words);
}

void launch_bin(hipStream_t st, const BinArgs& a) {
  if (a.max_tiles == 0) return;
  hipLaunchKernelGGL(k_bin, dim3(a.max_tiles), dim3(kBinThreads), 0, st, a);
}

void launch_ranges(hipStream_t st, const uint32_t* pair_count, uint32_t capacity, const uint32_t* keys,
                   uint32_t* ranges, uint32_t num_tiles, int num_cus, uint32_t* pair_max, uint32_t* host_total,
                   uint32_t* err) {
  uint32_t blocks = (capacity / 4u + 255u) / 256u;
  const uint32_t cap = uint32_t(num_cus) * 8u;
  if (blocks > cap) blocks = cap;
  if (blocks < 1u) blocks = 1u;
  hipLaunchKernelGGL(k_ranges, dim3(blocks), dim3(256), 0, st, pair_count, capacity, keys, ranges, num_tiles,
                     pair_max, host_total, err);
}

}  // namespace gs3d
