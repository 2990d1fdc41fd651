// Renderer kernels: tile binning and tile compositing. They replace the
// reference's draw_indirect of 6 vertices per instance through the
// fixed-function rasterizer and ALPHA_BLENDING (src/renderer.rs:163-195,
// 284-308; render.wesl:143-181).
//
//   bin       one thread per drawn instance, in sorted (back-to-front) order:
//             count the screen tiles its footprint touches, exclusive-scan the
//             counts across the grid (decoupled look-back) and emit
//             (tile id, Gaussian index) pairs. Pairs come out in draw order, so
//             a stable sort by tile id keeps every tile's list in draw order.
//   ranges    start/end of every tile's run in the tile-sorted pairs.
//   composite one workgroup per tile, one lane per pixel; splat records are
//             staged through LDS in batches of one per lane and evaluated by
//             every pixel. Fast mode walks front to back with transmittance
//             and stops once every pixel is saturated; exact mode walks back to
//             front and stores unorm8 after every splat, as blending into an
//             Rgba8Unorm target does.
#include <hip/hip_runtime.h>

#include "device_util.hpp"
#include "kernels.hpp"
#include "types.hpp"

namespace gs3d {

constexpr int kBinThreads = 256;
constexpr int kBinItems = 4;
constexpr uint32_t kBinTile = kBinThreads * kBinItems;

// Early-termination threshold on transmittance for the fast mode: what is
// left behind contributes at most 1e-4 of a splat colour.
constexpr float kMinTransmittance = 1e-4f;

uint32_t bin_max_tiles(uint32_t max_instances) { return (max_instances + kBinTile - 1) / kBinTile; }

__global__ __launch_bounds__(256) void k_bin_reset(uint32_t* ranges, uint32_t range_words, uint32_t* status,
                                                   uint32_t max_bin_tiles, uint32_t* ticket, uint32_t* pair_count) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t j = gid; j < range_words; j += stride) ranges[j] = 0u;
  for (uint32_t j = gid; j < max_bin_tiles; j += stride) status[j] = 0u;
  if (gid == 0) {
    *ticket = 0u;
    *pair_count = 0u;
  }
}

// Inclusive tile rectangle a splat can touch; false when it touches none.
// Pixel x is covered only if |x + 0.5 - cx| <= hx, so [cx - hx - 1,
// cx + hx + 1] is a conservative pixel range.
__device__ __forceinline__ bool splat_tiles(const Splat2D& s, const TileGrid& g, uint32_t& tx0, uint32_t& ty0,
                                            uint32_t& tx1, uint32_t& ty1) {
  if (!s.drawable) return false;
  const float fx0 = floorf(s.cx - s.hx - 1.0f), fx1 = floorf(s.cx + s.hx + 1.0f);
  const float fy0 = floorf(s.cy - s.hy - 1.0f), fy1 = floorf(s.cy + s.hy + 1.0f);
  const float wmax = float(g.width - 1), hmax = float(g.height - 1);
  if (!(fx1 >= 0.0f) || !(fx0 <= wmax) || !(fy1 >= 0.0f) || !(fy0 <= hmax)) return false;
  const uint32_t x0 = uint32_t(fmaxf(fx0, 0.0f)), x1 = uint32_t(fminf(fx1, wmax));
  const uint32_t y0 = uint32_t(fmaxf(fy0, 0.0f)), y1 = uint32_t(fminf(fy1, hmax));
  tx0 = x0 / g.tile_w;
  tx1 = x1 / g.tile_w;
  ty0 = y0 / g.tile_h;
  ty1 = y1 / g.tile_h;
  return true;
}

__device__ __forceinline__ uint32_t scan256_inclusive(uint32_t x, uint32_t* s_wsum) {
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

__global__ __launch_bounds__(kBinThreads) void k_bin(BinArgs a) {
  __shared__ uint32_t s_wsum[kBinThreads / kWave];
  __shared__ uint32_t s_ktot[kBinItems];
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_base;

  const uint32_t t = take_ticket(a.ticket, &s_tile);
  const uint32_t v = a.indirect_args[1];
  const uint32_t num_tiles = (v + kBinTile - 1) / kBinTile;
  if (t >= num_tiles) return;
  const uint32_t tid = threadIdx.x;
  const uint32_t base = t * kBinTile;

  uint32_t cnt[kBinItems], excl[kBinItems], gidx[kBinItems];
  uint32_t r0[kBinItems], r1[kBinItems];  // packed (tx0 | ty0 << 16), (tx1 | ty1 << 16)
#pragma unroll
  for (int k = 0; k < kBinItems; ++k) {
    const uint32_t s = base + uint32_t(k) * kBinThreads + tid;
    cnt[k] = 0u;
    gidx[k] = 0u;
    r0[k] = r1[k] = 0u;
    if (s < v) {
      const uint32_t i = a.indices[s];
      gidx[k] = i;
      const Splat2D sp = a.splats[i];
      uint32_t tx0, ty0, tx1, ty1;
      if (splat_tiles(sp, a.grid, tx0, ty0, tx1, ty1)) {
        cnt[k] = (tx1 - tx0 + 1u) * (ty1 - ty0 + 1u);
        r0[k] = tx0 | (ty0 << 16);
        r1[k] = tx1 | (ty1 << 16);
      }
    }
  }
  // exclusive scan in (item, thread) order == sorted instance order
  uint32_t run = 0;
#pragma unroll
  for (int k = 0; k < kBinItems; ++k) {
    const uint32_t inc = scan256_inclusive(cnt[k], s_wsum);
    excl[k] = run + inc - cnt[k];
    if (tid == kBinThreads - 1) s_ktot[k] = inc;
    __syncthreads();
    run += s_ktot[k];
  }
  if (tid == 0) {
    uint32_t e = 0;
    if (t == 0) {
      status_store(a.status, kFlagInclusive | run);
    } else {
      status_store(a.status + t, kFlagAggregate | run);
      e = lookback_exclusive(a.status, t, a.err);
      status_store(a.status + t, kFlagInclusive | (e + run));
    }
    s_base = e;
    if (t == num_tiles - 1) {
      *a.pair_count = e + run;
      if (e + run > a.capacity) atomicOr(a.err, kErrTileOverflow);
    }
  }
  __syncthreads();
  const uint32_t off0 = s_base;
#pragma unroll
  for (int k = 0; k < kBinItems; ++k) {
    if (cnt[k] == 0u) continue;
    uint32_t p = off0 + excl[k];
    const uint32_t tx0 = r0[k] & 0xFFFFu, ty0 = r0[k] >> 16, tx1 = r1[k] & 0xFFFFu, ty1 = r1[k] >> 16;
    for (uint32_t ty = ty0; ty <= ty1; ++ty)
      for (uint32_t tx = tx0; tx <= tx1; ++tx, ++p)
        if (p < a.capacity) {
          a.pair_keys[p] = ty * a.grid.tiles_x + tx;
          a.pair_vals[p] = gidx[k];
        }
  }
}

__global__ __launch_bounds__(256) void k_ranges(const uint32_t* __restrict__ pair_count, uint32_t capacity,
                                                const uint32_t* __restrict__ keys, uint32_t* __restrict__ ranges) {
  const uint32_t total = *pair_count;
  const uint32_t p = total < capacity ? total : capacity;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < p; j += gridDim.x * blockDim.x) {
    const uint32_t k = keys[j];
    if (j == 0 || keys[j - 1] != k) ranges[2 * k] = j;
    if (j == p - 1 || keys[j + 1] != k) ranges[2 * k + 1] = j + 1;
  }
}

__device__ __forceinline__ uint32_t unorm8(float x) {
  x = fminf(fmaxf(x, 0.0f), 1.0f);
  return uint32_t(int(x * 255.0f + 0.5f));
}

__device__ __forceinline__ float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

// Fragment stage (render.wesl:143-181) for one staged splat at pixel centre
// (px, py): the fragment alpha, or -1 when discarded. r0 = (cx, cy, ux, uy),
// r1 = (vx, vy, r, g), r2 = (b, a, hx, hy).
__device__ __forceinline__ float frag_alpha(int mode, float sd2, float ell2, const float4& r0, const float4& r1,
                                            const float4& r2, float px, float py) {
  const float dx = px - r0.x;
  const float dy = py - r0.y;
  if (mode == kModePoint) return (dx >= -r2.z && dx < r2.z && dy >= -r2.w && dy < r2.w) ? 1.0f : -1.0f;
  const float qx = dx * r0.z + dy * r0.w;
  const float qy = dx * r1.x + dy * r1.y;
  const float r2q = qx * qx + qy * qy;
  if (!(r2q <= sd2)) return -1.0f;
  if (mode == kModeEllipse) return r2.y + (1.0f - r2.y) * ((r2q > ell2) ? 1.0f : 0.0f);
  return r2.y * expf(-r2q);
}

constexpr int kMaxCompositeThreads = 256;

__global__ __launch_bounds__(kMaxCompositeThreads) void k_composite(CompositeArgs a) {
  __shared__ float4 s_r0[kMaxCompositeThreads];
  __shared__ float4 s_r1[kMaxCompositeThreads];
  __shared__ float4 s_r2[kMaxCompositeThreads];

  const uint32_t nt = blockDim.x;
  const uint32_t tid = threadIdx.x;
  const uint32_t tile = blockIdx.x;
  const uint32_t tx = tile % a.grid.tiles_x, ty = tile / a.grid.tiles_x;
  const uint32_t px = tx * a.grid.tile_w + tid % a.grid.tile_w;
  const uint32_t py = ty * a.grid.tile_h + tid / a.grid.tile_w;
  const bool inside = px < a.grid.width && py < a.grid.height;
  const float fpx = float(px) + 0.5f, fpy = float(py) + 0.5f;
  const uint32_t start = a.ranges[2 * tile], end = a.ranges[2 * tile + 1];
  const FrameConsts& fc = *a.fc;
  const int mode = fc.display_mode;
  const float sd = fc.max_std_dev;
  const float sd2 = sd * sd;
  const float e = sd - 0.1f;
  const float ell2 = e * e;
  const bool clamp_src = a.format == kFormatRgba8Unorm;

  if (a.blend == kBlendFast) {
    float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
    bool done = !inside;
    for (uint32_t bend = end; bend > start;) {
      const uint32_t cnt = (bend - start) < nt ? (bend - start) : nt;
      __syncthreads();
      if (tid < cnt) {
        const uint32_t k = a.pair_vals[bend - 1 - tid];  // nearest first
        const float4* rp = reinterpret_cast<const float4*>(a.splats + k);
        s_r0[tid] = rp[0];
        s_r1[tid] = rp[1];
        s_r2[tid] = rp[2];
      }
      __syncthreads();
      if (!done) {
        for (uint32_t q = 0; q < cnt; ++q) {
          const float4 r0 = s_r0[q], r1 = s_r1[q], r2 = s_r2[q];
          float al = frag_alpha(mode, sd2, ell2, r0, r1, r2, fpx, fpy);
          if (al < 0.0f) continue;
          float sr = r1.z, sg = r1.w, sb = r2.x;
          if (clamp_src) {
            sr = clamp01(sr);
            sg = clamp01(sg);
            sb = clamp01(sb);
            al = clamp01(al);
          }
          const float wgt = T * al;
          cr = cr + wgt * sr;
          cg = cg + wgt * sg;
          cb = cb + wgt * sb;
          T = T * (1.0f - al);
          if (T < kMinTransmittance) {
            done = true;
            break;
          }
        }
      }
      bend -= cnt;
      if (__syncthreads_and(done ? 1 : 0)) break;
    }
    if (inside) {
      const size_t o = size_t(py) * a.grid.width + px;
      if (a.format == kFormatRgba8Unorm) {
        const uint32_t pk = unorm8(cr) | (unorm8(cg) << 8) | (unorm8(cb) << 16) | (255u << 24);
        reinterpret_cast<uint32_t*>(a.out)[o] = pk;
      } else {
        reinterpret_cast<float4*>(a.out)[o] = make_float4(cr, cg, cb, 1.0f);
      }
    }
  } else {
    // exact: back to front, destination rounded to the target after each splat
    float dr = 0.0f, dg = 0.0f, db = 0.0f, da = 1.0f;
    const bool q8 = a.format == kFormatRgba8Unorm;
    for (uint32_t b0 = start; b0 < end; b0 += nt) {
      const uint32_t cnt = (end - b0) < nt ? (end - b0) : nt;
      __syncthreads();
      if (tid < cnt) {
        const uint32_t k = a.pair_vals[b0 + tid];  // farthest first
        const float4* rp = reinterpret_cast<const float4*>(a.splats + k);
        s_r0[tid] = rp[0];
        s_r1[tid] = rp[1];
        s_r2[tid] = rp[2];
      }
      __syncthreads();
      if (inside) {
        for (uint32_t q = 0; q < cnt; ++q) {
          const float4 r0 = s_r0[q], r1 = s_r1[q], r2 = s_r2[q];
          float al = frag_alpha(mode, sd2, ell2, r0, r1, r2, fpx, fpy);
          if (al < 0.0f) continue;
          float sr = r1.z, sg = r1.w, sb = r2.x;
          if (q8) {
            sr = clamp01(sr);
            sg = clamp01(sg);
            sb = clamp01(sb);
            al = clamp01(al);
          }
          const float om = 1.0f - al;
          dr = sr * al + dr * om;
          dg = sg * al + dg * om;
          db = sb * al + db * om;
          da = al + da * om;
          if (q8) {
            dr = float(unorm8(dr)) / 255.0f;
            dg = float(unorm8(dg)) / 255.0f;
            db = float(unorm8(db)) / 255.0f;
            da = float(unorm8(da)) / 255.0f;
          }
        }
      }
    }
    if (inside) {
      const size_t o = size_t(py) * a.grid.width + px;
      if (q8) {
        const uint32_t pk = unorm8(dr) | (unorm8(dg) << 8) | (unorm8(db) << 16) | (unorm8(da) << 24);
        reinterpret_cast<uint32_t*>(a.out)[o] = pk;
      } else {
        reinterpret_cast<float4*>(a.out)[o] = make_float4(dr, dg, db, da);
      }
    }
  }
}

void launch_bin_reset(hipStream_t st, uint32_t* ranges, uint32_t num_tiles, uint32_t* status, uint32_t max_bin_tiles,
                      uint32_t* ticket, uint32_t* pair_count) {
  uint32_t words = num_tiles * 2u;
  uint32_t m = words > max_bin_tiles ? words : max_bin_tiles;
  uint32_t blocks = (m + 255u) / 256u;
  if (blocks > 1024u) blocks = 1024u;
  if (blocks < 1u) blocks = 1u;
  hipLaunchKernelGGL(k_bin_reset, dim3(blocks), dim3(256), 0, st, ranges, words, status, max_bin_tiles, ticket,
                     pair_count);
}

void launch_bin(hipStream_t st, const BinArgs& a) {
  if (a.max_tiles == 0) return;
  hipLaunchKernelGGL(k_bin, dim3(a.max_tiles), dim3(kBinThreads), 0, st, a);
}

void launch_ranges(hipStream_t st, const uint32_t* pair_count, uint32_t capacity, const uint32_t* keys,
                   uint32_t* ranges, int num_cus) {
  uint32_t blocks = (capacity + 255u) / 256u;
  const uint32_t cap = uint32_t(num_cus) * 8u;
  if (blocks > cap) blocks = cap;
  if (blocks < 1u) blocks = 1u;
  hipLaunchKernelGGL(k_ranges, dim3(blocks), dim3(256), 0, st, pair_count, capacity, keys, ranges);
}

void launch_composite(hipStream_t st, const CompositeArgs& a) {
  const uint32_t tiles = a.grid.tiles_x * a.grid.tiles_y;
  const uint32_t threads = a.grid.tile_w * a.grid.tile_h;
  if (tiles == 0) return;
  hipLaunchKernelGGL(k_composite, dim3(tiles), dim3(threads), 0, st, a);
}

}  // namespace gs3d
