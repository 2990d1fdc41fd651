// Stable LSD radix sort of (u32 key, u32 value) pairs, 8-bit digits, one
// onesweep pass per digit (Adinets & Merrill 2022) with decoupled look-back
// (Merrill & Garland 2016). Replaces the reference's vendored wgpu_sort
// (src/radix_sorter.rs:98-940, src/shader/radix_sort.wgsl): same result (a
// stable ascending sort), different machinery:
//   * the first pass's 256-bin histogram comes from one upsweep over the
//     keys (per-wave LDS atomics) or from the kernel that wrote the keys;
//     every pass counts the next pass's digits of the keys it holds;
//   * each pass is a persistent kernel: a workgroup takes tiles of
//     256 x ITEMS keys (ITEMS = 16 unless the caller picks 8 or 12) by
//     ticket, ranks them with wave64 ballot matching (8 ballots per key give
//     the lanes sharing its digit), publishes per-digit counts, looks back
//     over its predecessors for the global offset, reorders the tile in LDS
//     and writes runs of equal digits contiguously.
// Count comes from device memory (GPU-driven, like the reference's indirect
// dispatch): n = min(*count * count_mul, count_max).
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <mutex>
#include <unordered_map>

#include "gs3d.h"

#include "device_util.hpp"
#include "kernels.hpp"
#include "types.hpp"

namespace gs3d {

constexpr int kRsThreads = 256;
constexpr int kRsWaves = kRsThreads / kWave;
// keys per thread of a onesweep tile: 20 by default (5120-key tiles); 8, 12,
// 16 and 24 are compiled too for callers that time them (the batch
// autotuner); the look-back scratch is sized for the smallest tile
constexpr int kRsItems = 20;
constexpr int kRsMinItems = 8;
constexpr uint32_t kRsTile = kRsThreads * kRsItems;
constexpr uint32_t kRsMinTile = kRsThreads * kRsMinItems;
// predecessors' look-back words requested per step of a digit's look-back
#ifndef GS3D_SORT_LOOK_WINDOW
#define GS3D_SORT_LOOK_WINDOW 4
#endif
constexpr int kRsLookWindow = GS3D_SORT_LOOK_WINDOW;

__device__ __forceinline__ uint32_t sort_n(const SortCount& c) {
  if (c.count == nullptr) return c.count_max;  // a count fixed at launch
  const uint64_t v = uint64_t(*c.count) * uint64_t(c.count_mul);
  return v < uint64_t(c.count_max) ? uint32_t(v) : c.count_max;
}

// Inclusive scan of one value per thread across a 256-thread workgroup.
// `s_wsum` holds kRsWaves words of LDS. Contains __syncthreads.
__device__ __forceinline__ uint32_t block_inclusive_scan_256(uint32_t x, uint32_t* s_wsum) {
  const uint32_t lane = threadIdx.x % kWave;
  const uint32_t w = threadIdx.x / kWave;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, kWave);
    if (lane >= uint32_t(off)) x += y;
  }
  if (lane == kWave - 1) s_wsum[w] = x;
  __syncthreads();
  uint32_t pre = 0;
  for (uint32_t q = 0; q < w; ++q) pre += s_wsum[q];
  __syncthreads();
  return x + pre;
}

__global__ __launch_bounds__(256) void k_rs_zero(SortCount cnt, SortScratch sc, int passes, uint32_t tile_keys,
                                                 uint32_t* pad_keys, uint32_t pad_begin, uint32_t pad_end,
                                                 int zero_hist) {
  const uint32_t n = sort_n(cnt);
  const uint32_t tiles = (n + tile_keys - 1) / tile_keys;
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t stride = gridDim.x * blockDim.x;
  if (zero_hist)
    for (uint32_t j = gid; j < uint32_t(passes) * 256u; j += stride) sc.hist[j] = 0u;
  if (gid < uint32_t(passes)) sc.tickets[gid] = 0u;
  const uint64_t per_pass = uint64_t(tiles) * 256u;
  for (uint64_t j = gid; j < per_pass * uint64_t(passes); j += stride) {
    const uint64_t p = j / per_pass;
    const uint64_t r = j - p * per_pass;
    sc.lookback[p * uint64_t(sc.max_tiles) * 256u + r] = 0u;
  }
  // radix_sort.wgsl:92-110: keys past the sorted range are 0xFFFFFFFF
  if (pad_keys)
    for (uint32_t j = pad_begin + gid; j < pad_end; j += stride) pad_keys[j] = 0xFFFFFFFFu;
}

// Histogram of the first pass's digit: one read of the keys, each wave
// counting into its own LDS histogram (no atomic contention between waves).
__global__ __launch_bounds__(256) void k_rs_upsweep(SortCount cnt, const uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ hist) {
  __shared__ uint32_t s_h[kRsWaves][256];
  const uint32_t n = sort_n(cnt);
  const uint32_t w = threadIdx.x / kWave;
  for (int q = 0; q < kRsWaves; ++q) s_h[q][threadIdx.x] = 0u;
  __syncthreads();
  const uint32_t stride = gridDim.x * blockDim.x * 4u;
  for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) * 4u; i < n; i += stride) {
    uint32_t k4[4] = {0u, 0u, 0u, 0u};
    if (i + 3u < n) {
      const uint4 v = *reinterpret_cast<const uint4*>(keys + i);
      k4[0] = v.x;
      k4[1] = v.y;
      k4[2] = v.z;
      k4[3] = v.w;
    } else {
      for (int q = 0; q < 4; ++q) k4[q] = (i + q < n) ? keys[i + q] : 0u;
    }
    const int m = (i + 3u < n) ? 4 : int(n - i);
    for (int q = 0; q < m; ++q) atomicAdd(&s_h[w][k4[q] & 0xFFu], 1u);
  }
  __syncthreads();
  uint32_t c = 0;
  for (int q = 0; q < kRsWaves; ++q) c += s_h[q][threadIdx.x];
  if (c) atomicAdd(&hist[threadIdx.x], c);
}

// ---- one tile of a pass: load, rank, reorder in LDS ------------------------------

// LDS of a pass over one tile of 256 * ITEMS keys.
template <int ITEMS>
struct TileLds {
  uint32_t keys[kRsThreads * ITEMS];
  uint32_t vals[kRsThreads * ITEMS];
  uint32_t wcnt[kRsWaves][256];  // per wave and digit: count, then the wave's first slot
  uint32_t gbase[256];           // global position of slot 0 of the digit's run
  uint32_t wsum[kRsWaves];
};

// Loads tile keys/values [base, base + valid) (key 0xFFFFFFFF past the end:
// digit 255, ranked after every valid key) in the (wave, item, lane) order
// the ranking assumes.
template <int ITEMS>
__device__ __forceinline__ void load_tile(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                          uint32_t base, uint32_t valid, uint32_t (&key)[ITEMS],
                                          uint32_t (&val)[ITEMS]) {
  const uint32_t lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  const uint32_t first = w * (kWave * ITEMS) + lane;  // item 0's slot
  const uint32_t* kp = kin + base + first;
  const uint32_t* vp = vin + base + first;
  if (valid >= uint32_t(kRsThreads * ITEMS)) {
    // a full tile (every one but the last): no per-item test, one address
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      key[k] = kp[k * kWave];
      val[k] = vp[k * kWave];
    }
  } else {
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const bool ok = first + uint32_t(k) * kWave < valid;
      key[k] = ok ? kp[k * kWave] : 0xFFFFFFFFu;
      val[k] = ok ? vp[k * kWave] : 0u;
    }
  }
}

// Stable rank of every key within its wave by wave64 match-by-ballot (8
// ballots give the lanes sharing the key's digit); lds.wcnt[w][d] ends as the
// wave's count of digit d. lds.wcnt must be zero on entry.
template <int ITEMS>
__device__ __forceinline__ void rank_tile(TileLds<ITEMS>& lds, const uint32_t (&key)[ITEMS], int shift,
                                          uint32_t (&lrank)[ITEMS]) {
  const uint32_t w = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const uint32_t d = (key[k] >> shift) & 0xFFu;
    // lanes whose digit differs in some bit: the OR over the digit's bits of
    // (the lanes with the bit set) XOR (all lanes if this lane's bit is set)
    uint32_t ml = 0u, mh = 0u;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int32_t sb = int32_t(d << (31 - b)) >> 31;  // 0 or -1
      const uint64_t bb = __builtin_amdgcn_ballot_w64(sb != 0);
      ml |= uint32_t(bb) ^ uint32_t(sb);
      mh |= uint32_t(bb >> 32) ^ uint32_t(sb);
    }
    const uint64_t m = (uint64_t(~mh) << 32) | ~ml;  // the lanes sharing this lane's digit
    const uint32_t r = mask_rank(m);
    const uint32_t prev = lds.wcnt[w][d];
    if (r == 0u) lds.wcnt[w][d] = prev + uint32_t(__popcll(m));
    lrank[k] = prev + r;
  }
}

// Thread d: turns the wave counts of digit d into wave offsets and returns
// the tile's count of digit d. Call after a barrier that follows rank_tile.
template <int ITEMS>
__device__ __forceinline__ uint32_t digit_wave_offsets(TileLds<ITEMS>& lds) {
  const uint32_t d = threadIdx.x;
  uint32_t tot = 0;
#pragma unroll
  for (int q = 0; q < kRsWaves; ++q) {
    const uint32_t c = lds.wcnt[q][d];
    lds.wcnt[q][d] = tot;
    tot += c;
  }
  return tot;
}

// Thread d: adds the digit's first slot in the tile to its wave offsets, so
// that wcnt[w][d] is wave w's first slot for digit d.
template <int ITEMS>
__device__ __forceinline__ void set_wave_slots(TileLds<ITEMS>& lds, uint32_t dbase) {
#pragma unroll
  for (int q = 0; q < kRsWaves; ++q) lds.wcnt[q][threadIdx.x] += dbase;
}

// Writes every key and value to its slot in the tile sorted by the digit
// (stable: ranks follow input order). lds.wcnt[w][d] must hold wave w's
// first slot for digit d (set_wave_slots).
template <int ITEMS>
__device__ __forceinline__ void reorder_tile(TileLds<ITEMS>& lds, const uint32_t (&key)[ITEMS],
                                             const uint32_t (&val)[ITEMS], const uint32_t (&lrank)[ITEMS],
                                             int shift) {
  const uint32_t w = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const uint32_t dk = (key[k] >> shift) & 0xFFu;
    const uint32_t slot = lds.wcnt[w][dk] + lrank[k];
    lds.keys[slot] = key[k];
    lds.vals[slot] = val[k];
  }
}

// Writes the tile's valid slots to gbase[digit] + slot: runs of equal digits
// land contiguously. lds.gbase must be set and visible.
template <int ITEMS>
__device__ __forceinline__ void scatter_tile(const TileLds<ITEMS>& lds, uint32_t valid, uint32_t n, int shift,
                                             uint32_t* __restrict__ kout, uint32_t* __restrict__ vout) {
  for (uint32_t j = threadIdx.x; j < valid; j += kRsThreads) {
    const uint32_t k = lds.keys[j];
    const uint32_t pos = lds.gbase[(k >> shift) & 0xFFu] + j;
    if (pos < n) {  // always true for consistent counts; guards the write
      kout[pos] = k;
      vout[pos] = lds.vals[j];
    }
  }
}

template <int ITEMS>
__global__ __launch_bounds__(kRsThreads) void k_rs_onesweep(SortCount cnt, const uint32_t* __restrict__ kin,
                                                           const uint32_t* __restrict__ vin,
                                                           uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                           const uint32_t* __restrict__ hist, uint32_t* lookback,
                                                           uint32_t* ticket, int shift, uint32_t* err,
                                                           uint32_t* hist_next) {
  constexpr uint32_t kTile = kRsThreads * ITEMS;
  __shared__ TileLds<ITEMS> lds;
  __shared__ uint32_t s_next[256];  // next pass's digit counts of the tile
  __shared__ uint32_t s_hscan[256];
  __shared__ uint32_t s_tile;

  const uint32_t n = sort_n(cnt);
  const uint32_t num_tiles = (n + kTile - 1) / kTile;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid % kWave;
  const uint32_t w = tid / kWave;

  {
    const uint32_t h = hist[tid];
    s_hscan[tid] = block_inclusive_scan_256(h, lds.wsum) - h;
  }

  for (;;) {
    const uint32_t t = take_ticket(ticket, &s_tile);
    if (t >= num_tiles) break;
    const uint32_t base = t * kTile;
    const uint32_t valid = (n - base) < kTile ? (n - base) : kTile;

#pragma unroll
    for (int q = 0; q < kRsWaves; ++q) lds.wcnt[q][tid] = 0u;
    s_next[tid] = 0u;
    __syncthreads();

    uint32_t key[ITEMS], val[ITEMS], lrank[ITEMS];
    load_tile<ITEMS>(kin, vin, base, valid, key, val);
    if (hist_next) {
      // the next pass's histogram: the same keys, whatever their order.
      if (shift + 8 == 24) {
        // The top byte (sign and high exponent bits of float keys such as
        // depths) takes a few values, on which the wave's LDS atomics would
        // serialize: keys whose top byte is the wave's first key's are
        // counted in a register and added once per thread.
        const uint32_t dom = __builtin_amdgcn_readfirstlane(key[0] >> 24);
        uint32_t own = 0u;
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
          const uint32_t li = w * (kWave * ITEMS) + uint32_t(k) * kWave + lane;
          const uint32_t nd = key[k] >> 24;
          if (li < valid) {
            if (nd == dom)
              ++own;
            else
              atomicAdd(&s_next[nd], 1u);
          }
        }
        if (own) atomicAdd(&s_next[dom], own);
      } else {
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
          const uint32_t li = w * (kWave * ITEMS) + uint32_t(k) * kWave + lane;
          if (li < valid) atomicAdd(&s_next[(key[k] >> (shift + 8)) & 0xFFu], 1u);
        }
      }
    }
    rank_tile(lds, key, shift, lrank);
    __syncthreads();

    // per digit (thread = digit): wave offsets, tile total, publish
    const uint32_t d = tid;
    const uint32_t tot = digit_wave_offsets(lds);
    const uint32_t pub = tot - ((d == 255u) ? (kTile - valid) : 0u);
    status_store(lookback + size_t(t) * 256u + d, (t == 0 ? kFlagInclusive : kFlagAggregate) | pub);
    const uint32_t dbase = block_inclusive_scan_256(tot, lds.wsum) - tot;
    set_wave_slots(lds, dbase);
    __syncthreads();

    reorder_tile(lds, key, val, lrank, shift);

    // global offset of this tile's run of digit d
    // (each step requests the words of the next kRsLookWindow predecessors
    // at once and consumes them in order up to an inclusive prefix or a
    // predecessor that has not published yet)
    uint32_t excl = 0;
    if (t > 0) {
      int64_t j = int64_t(t) - 1;
      uint32_t spins = 0;
      bool done = false;
      while (!done && j >= 0) {
        uint32_t sw[kRsLookWindow];
#pragma unroll
        for (int q = 0; q < kRsLookWindow; ++q)
          sw[q] = j - q >= 0 ? status_load(lookback + size_t(j - q) * 256u + d) : kFlagInclusive;
        int adv = 0;
#pragma unroll
        for (int q = 0; q < kRsLookWindow; ++q) {
          if (done || adv < q) continue;  // stopped at an earlier word
          const uint32_t flag = sw[q] & kFlagMask;
          if (flag == 0u) continue;  // not published: resume here
          excl += sw[q] & kValueMask;
          ++adv;
          done = flag == kFlagInclusive;
        }
        j -= adv;
        if (adv == 0) {
          __builtin_amdgcn_s_sleep(1);
          if (++spins > kSpinLimit) {
            atomicOr(err, kErrLookbackSpin);
            break;
          }
        }
      }
      status_store(lookback + size_t(t) * 256u + d, kFlagInclusive | (excl + pub));
    }
    lds.gbase[d] = s_hscan[d] + excl - dbase;
    __syncthreads();
    if (hist_next && s_next[tid]) atomicAdd(&hist_next[tid], s_next[tid]);

    scatter_tile(lds, valid, n, shift, kout, vout);
    __syncthreads();
  }
}

// ---- reduce-then-scan passes -------------------------------------------------------
//
// The same pass without any wait between workgroups, in three launches:
//   count    per tile, the count of each digit: counts[d][t] (digit-major, so
//            that each digit's column over the tiles is contiguous)
//   colscan  per digit (one workgroup each), the exclusive prefix of its
//            column, in place, and the digit's total
//   scatter  per tile: the digit's global base (exclusive scan of the
//            totals) plus the tile's prefix, then rank, reorder and scatter
//            as above.
// It reads the keys once more per pass (4 bytes per key), but no tile waits
// on another: a onesweep tile's look-back round trips to its predecessors'
// status words grow with the number of tiles that publish at once.

__global__ __launch_bounds__(kRsThreads) void k_rs_count(SortCount cnt, const uint32_t* __restrict__ keys,
                                                        int shift, uint32_t* __restrict__ counts) {
  __shared__ uint32_t s_h[kRsWaves][256];
  const uint32_t n = sort_n(cnt);
  const uint32_t num_tiles = (n + kRsTile - 1) / kRsTile;
  const uint32_t t = blockIdx.x;
  if (t >= num_tiles) return;
  const uint32_t tid = threadIdx.x, lane = tid % kWave, w = tid / kWave;
#pragma unroll
  for (int q = 0; q < kRsWaves; ++q) s_h[q][tid] = 0u;
  __syncthreads();
  const uint32_t base = t * kRsTile;
  const uint32_t valid = (n - base) < kRsTile ? (n - base) : kRsTile;
  uint32_t key[kRsItems];
#pragma unroll
  for (int k = 0; k < kRsItems; ++k) {
    const uint32_t li = w * (kWave * kRsItems) + uint32_t(k) * kWave + lane;
    key[k] = li < valid ? keys[base + li] : 0u;
  }
#pragma unroll
  for (int k = 0; k < kRsItems; ++k) {
    const uint32_t li = w * (kWave * kRsItems) + uint32_t(k) * kWave + lane;
    if (li < valid) atomicAdd(&s_h[w][(key[k] >> shift) & 0xFFu], 1u);
  }
  __syncthreads();
  uint32_t c = 0;
#pragma unroll
  for (int q = 0; q < kRsWaves; ++q) c += s_h[q][tid];
  counts[size_t(tid) * num_tiles + t] = c;
}

__global__ __launch_bounds__(kRsThreads) void k_rs_colscan(SortCount cnt, uint32_t* __restrict__ counts,
                                                          uint32_t* __restrict__ totals) {
  __shared__ uint32_t s_wsum[kRsWaves];
  __shared__ uint32_t s_last;
  const uint32_t n = sort_n(cnt);
  const uint32_t num_tiles = (n + kRsTile - 1) / kRsTile;
  const uint32_t d = blockIdx.x, tid = threadIdx.x;
  uint32_t* col = counts + size_t(d) * num_tiles;
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < num_tiles; c0 += kRsThreads) {
    const uint32_t i = c0 + tid;
    const uint32_t x = i < num_tiles ? col[i] : 0u;
    const uint32_t inc = block_inclusive_scan_256(x, s_wsum);
    if (i < num_tiles) col[i] = carry + inc - x;
    if (tid == kRsThreads - 1) s_last = inc;
    __syncthreads();
    carry += s_last;
    __syncthreads();
  }
  if (tid == 0) totals[d] = carry;
}

__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(SortCount cnt, const uint32_t* __restrict__ kin,
                                                          const uint32_t* __restrict__ vin,
                                                          uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                          const uint32_t* __restrict__ counts,
                                                          const uint32_t* __restrict__ totals, int shift) {
  __shared__ TileLds<kRsItems> lds;
  const uint32_t n = sort_n(cnt);
  const uint32_t num_tiles = (n + kRsTile - 1) / kRsTile;
  const uint32_t t = blockIdx.x;
  if (t >= num_tiles) return;
  const uint32_t tid = threadIdx.x;
  const uint32_t base = t * kRsTile;
  const uint32_t valid = (n - base) < kRsTile ? (n - base) : kRsTile;

  // requested first: the loads overlap the scans below
  uint32_t key[kRsItems], val[kRsItems], lrank[kRsItems];
  load_tile<kRsItems>(kin, vin, base, valid, key, val);
  const uint32_t tot_all = totals[tid];
  const uint32_t col = counts[size_t(tid) * num_tiles + t];
#pragma unroll
  for (int q = 0; q < kRsWaves; ++q) lds.wcnt[q][tid] = 0u;
  // global base of digit tid's run of this tile
  const uint32_t gb = block_inclusive_scan_256(tot_all, lds.wsum) - tot_all + col;  // has barriers

  rank_tile(lds, key, shift, lrank);
  __syncthreads();
  const uint32_t tot = digit_wave_offsets(lds);
  const uint32_t dbase = block_inclusive_scan_256(tot, lds.wsum) - tot;
  set_wave_slots(lds, dbase);
  lds.gbase[tid] = gb - dbase;
  __syncthreads();
  reorder_tile(lds, key, val, lrank, shift);
  __syncthreads();
  scatter_tile(lds, valid, n, shift, kout, vout);
}

bool sort_items_supported(uint32_t items) {
  return items == 8 || items == 12 || items == 16 || items == 20 || items == 24;
}

uint32_t sort_tile_items() { return kRsTile; }

uint32_t sort_max_tiles(uint32_t count_max) { return (count_max + kRsMinTile - 1) / kRsMinTile; }

size_t sort_lookback_words(uint32_t count_max, int passes) {
  return size_t(sort_max_tiles(count_max)) * 256u * size_t(passes);
}

static void launch_zero(hipStream_t st, int num_cus, SortCount cnt, const SortScratch& sc, int passes,
                        uint32_t tile_keys, uint32_t* pad_keys, uint32_t pad_begin, uint32_t pad_end,
                        bool zero_hist) {
  uint32_t blocks = (sc.max_tiles * 256u * uint32_t(passes) + 255u) / 256u;
  const uint32_t cap = uint32_t(num_cus) * 4u;
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_rs_zero, dim3(blocks), dim3(256), 0, st, cnt, sc, passes, tile_keys, pad_keys, pad_begin,
                     pad_end, zero_hist ? 1 : 0);
}

// Resident onesweep workgroups per CU for ITEMS keys per thread (the
// persistent grid): what the kernel's registers and LDS allow, as the runtime
// computes it; more would only queue behind the resident ones.
template <int ITEMS>
static uint32_t onesweep_wg_per_cu() {
  static const uint32_t w = [] {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_rs_onesweep<ITEMS>, kRsThreads, 0) != hipSuccess || b < 1)
      b = 1;
    return uint32_t(b);
  }();
  return w;
}

template <int ITEMS>
static void launch_onesweep_passes(hipStream_t st, int num_cus, SortCount cnt, const uint32_t* ks0,
                                   const uint32_t* vs0, uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1,
                                   int passes, const SortScratch& sc) {
  const uint32_t tiles = (sc.max_tiles * kRsMinTile + kRsThreads * ITEMS - 1) / (kRsThreads * ITEMS);
  // 4096-key tiles: 3 workgroups per CU, although 4 fit (126 VGPRs, 40 KB of
  // LDS): the fourth is 5% slower on sorts of ~5M keys (faster only from
  // ~20M keys on)
  uint32_t wpc = onesweep_wg_per_cu<ITEMS>();
  if (ITEMS == 16 && wpc > 3u) wpc = 3u;
  uint32_t grid = uint32_t(num_cus) * wpc;
  if (grid > tiles) grid = tiles;
  if (grid < 1) grid = 1;
  for (int p = 0; p < passes; ++p) {
    const uint32_t* ks = p == 0 ? ks0 : ((p & 1) ? k1 : k0);
    const uint32_t* vs = p == 0 ? vs0 : ((p & 1) ? v1 : v0);
    uint32_t* kd = (p & 1) ? k0 : k1;
    uint32_t* vd = (p & 1) ? v0 : v1;
    uint32_t* next = (p + 1 < passes) ? sc.hist + (p + 1) * 256 : nullptr;
    hipLaunchKernelGGL(k_rs_onesweep<ITEMS>, dim3(grid), dim3(kRsThreads), 0, st, cnt, ks, vs, kd, vd,
                       sc.hist + p * 256, sc.lookback + size_t(p) * sc.max_tiles * 256u, sc.tickets + p, 8 * p,
                       sc.err, next);
  }
}

static int sort_algo() {
  static const int a = [] {
    const char* e = std::getenv("GS3D_SORT_ALGO");
    return (e && e[0] == 'r') ? 1 : 0;
  }();
  return a;
}

int launch_radix_sort(hipStream_t st, int num_cus, SortCount cnt, uint32_t* k0, uint32_t* v0, uint32_t* k1,
                      uint32_t* v1, int passes, const SortScratch& sc, uint32_t* pad_keys, uint32_t pad_begin,
                      uint32_t pad_end, bool hist_ready, const uint32_t* k_first, const uint32_t* v_first) {
  if (sc.max_tiles == 0 || passes <= 0) return 0;
  // pass 0 reads (k_first, v_first) when given, (k0, v0) otherwise
  const uint32_t* ks0 = k_first ? k_first : k0;
  const uint32_t* vs0 = k_first ? v_first : v0;
  if (sort_algo() == 1) {
    // reduce-then-scan: counts [256][tiles] in the look-back words, digit
    // totals in the histogram words; nothing to clear but the pad keys
    if (pad_keys) launch_zero(st, num_cus, cnt, sc, 0, kRsTile, pad_keys, pad_begin, pad_end, false);
    for (int p = 0; p < passes; ++p) {
      const uint32_t* ks = p == 0 ? ks0 : ((p & 1) ? k1 : k0);
      const uint32_t* vs = p == 0 ? vs0 : ((p & 1) ? v1 : v0);
      uint32_t* kd = (p & 1) ? k0 : k1;
      uint32_t* vd = (p & 1) ? v0 : v1;
      hipLaunchKernelGGL(k_rs_count, dim3(sc.max_tiles), dim3(kRsThreads), 0, st, cnt, ks, 8 * p, sc.lookback);
      hipLaunchKernelGGL(k_rs_colscan, dim3(256), dim3(kRsThreads), 0, st, cnt, sc.lookback, sc.hist);
      hipLaunchKernelGGL(k_rs_scatter, dim3(sc.max_tiles), dim3(kRsThreads), 0, st, cnt, ks, vs, kd, vd,
                         sc.lookback, sc.hist, 8 * p);
    }
    return passes & 1;
  }
  const uint32_t items = sc.items ? sc.items : uint32_t(kRsItems);
  launch_zero(st, num_cus, cnt, sc, passes, kRsThreads * items, pad_keys, pad_begin, pad_end, !hist_ready);
  if (!hist_ready) {
    uint32_t up_blocks = (sc.max_tiles * kRsMinTile + 4095u) / 4096u;
    if (up_blocks > uint32_t(num_cus) * 4u) up_blocks = uint32_t(num_cus) * 4u;
    if (up_blocks < 1) up_blocks = 1;
    hipLaunchKernelGGL(k_rs_upsweep, dim3(up_blocks), dim3(256), 0, st, cnt, ks0, sc.hist);
  }
  if (items == 8)
    launch_onesweep_passes<8>(st, num_cus, cnt, ks0, vs0, k0, v0, k1, v1, passes, sc);
  else if (items == 12)
    launch_onesweep_passes<12>(st, num_cus, cnt, ks0, vs0, k0, v0, k1, v1, passes, sc);
  else if (items == 16)
    launch_onesweep_passes<16>(st, num_cus, cnt, ks0, vs0, k0, v0, k1, v1, passes, sc);
  else if (items == 24)
    launch_onesweep_passes<24>(st, num_cus, cnt, ks0, vs0, k0, v0, k1, v1, passes, sc);
  else
    launch_onesweep_passes<20>(st, num_cus, cnt, ks0, vs0, k0, v0, k1, v1, passes, sc);
  return passes & 1;
}


// ---- the allocation-free sort ----------------------------------------------------
//
// One fixed array per GPU holds everything a sort of at most
// GS3D_RADIX_SORT_NOALLOC_MAX_KEYS pairs needs beyond the caller's input and
// output: the ping-pong keys and values, the look-back words, the four
// histograms and the tickets. It is zero-filled when the code object loads and
// lives until the process ends. Host code never passes the array's name to a
// kernel (on the host it is only a stand-in for the symbol): its device address
// comes from hipGetSymbolAddress, once per device, at device set-up.
constexpr uint32_t kNoAllocMaxKeys = GS3D_RADIX_SORT_NOALLOC_MAX_KEYS;
constexpr uint32_t kNoAllocMaxTiles = (kNoAllocMaxKeys + kRsMinTile - 1) / kRsMinTile;
constexpr int kNoAllocPasses = 4;
constexpr size_t kNoAllocKeysOff = 0;
constexpr size_t kNoAllocValsOff = kNoAllocKeysOff + kNoAllocMaxKeys;
constexpr size_t kNoAllocLookOff = kNoAllocValsOff + kNoAllocMaxKeys;
constexpr size_t kNoAllocHistOff = kNoAllocLookOff + size_t(kNoAllocMaxTiles) * 256u * kNoAllocPasses;
constexpr size_t kNoAllocTicketOff = kNoAllocHistOff + 256u * kNoAllocPasses;
constexpr size_t kNoAllocWords = kNoAllocTicketOff + 64u;  // tickets, padded to 256 bytes

static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

namespace {
std::mutex g_noalloc_mu;
std::unordered_map<int, uint32_t*> g_noalloc_ptr;  // device ordinal -> address of `scratch`
}  // namespace

hipError_t noalloc_sort_init(int ordinal) {
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_ptr.count(ordinal)) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (p == nullptr) return hipErrorInvalidSymbol;
  g_noalloc_ptr[ordinal] = static_cast<uint32_t*>(p);
  return hipSuccess;
}

size_t noalloc_sort_scratch_bytes() { return kNoAllocWords * sizeof(uint32_t); }

int launch_radix_sort_noalloc(hipStream_t st, int num_cus, int ordinal, const uint32_t* keys_in,
                              const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out, uint32_t n,
                              uint32_t* err) {
  if (n > kNoAllocMaxKeys) return 1;
  uint32_t* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    const auto it = g_noalloc_ptr.find(ordinal);
    if (it == g_noalloc_ptr.end()) return 2;
    base = it->second;
  }
  if (n == 0) return 0;
  SortScratch sc{};
  sc.hist = base + kNoAllocHistOff;
  sc.lookback = base + kNoAllocLookOff;
  sc.tickets = base + kNoAllocTicketOff;
  sc.max_tiles = kNoAllocMaxTiles;
  sc.err = err;
  const SortCount cnt{nullptr, 1u, n};
  // pass 0 reads the caller's input into the fixed pair, the passes then
  // alternate between the caller's output and the fixed pair; four passes end
  // in the output
  launch_radix_sort(st, num_cus, cnt, keys_out, vals_out, base + kNoAllocKeysOff, base + kNoAllocValsOff,
                    kNoAllocPasses, sc, nullptr, 0, 0, false, keys_in, vals_in);
  return 0;
}

}  // namespace gs3d
