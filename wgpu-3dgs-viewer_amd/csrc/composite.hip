// Tile compositing: replaces the reference's fixed-function rasterizer and
// ALPHA_BLENDING over draw_indirect quads (src/renderer.rs:163-195, 284-308;
// fragment stage render.wesl:143-181).
//
// One workgroup per work item, one lane per pixel of a screen tile (8x8,
// 16x16 or 32x8 pixels); each wave owns an 8x8 block of the tile. The tile's
// splats (its run of the tile-sorted pairs, in draw order) are staged through
// LDS in batches of one splat per lane; before evaluating a staged splat a
// wave tests the splat's pixel bound against its 8x8 block and skips it as a
// whole when they do not meet (most splats of dense scenes cover a few
// pixels).
//
// fast  : front to back in f32 with transmittance. The over operator is
//         associative, so a tile whose list is longer than one segment is
//         split into segments composited in parallel, each to a partial
//         (colour, transmittance) per pixel, and merged in order afterwards:
//         the few very dense tiles of a clustered scene no longer set the
//         kernel's length. Multi-segment tiles are scheduled first. A lane
//         stops when its transmittance falls below 1e-4, a workgroup when
//         all its lanes have.
// exact : back to front, one workgroup per tile; the destination is rounded
//         to the target format after every splat, as the blender does for an
//         Rgba8Unorm target (rounding is not associative: no segments).
#include <hip/hip_runtime.h>

#include "device_util.hpp"
#include "kernels.hpp"
#include "types.hpp"

namespace gs3d {

constexpr int kCompMaxThreads = 256;
constexpr int kPlanThreads = 1024;

// Lanes stop once what is left behind contributes at most 1e-4 of a colour.
constexpr float kMinTransmittance = 1e-4f;

__device__ __forceinline__ uint32_t unorm8(float x) {
  x = fminf(fmaxf(x, 0.0f), 1.0f);
  return uint32_t(int(x * 255.0f + 0.5f));
}

__device__ __forceinline__ float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

struct StagedSplats {
  float4 geo[kCompMaxThreads];    // cx, cy, ux, uy
  float4 geo2[kCompMaxThreads];   // vx, vy, opacity, point half-extent x
  float4 col[kCompMaxThreads];    // r, g, b, point half-extent y
  float4 bound[kCompMaxThreads];  // cx - hx, cx + hx, cy - hy, cy + hy
};

__device__ __forceinline__ void stage(StagedSplats& s, uint32_t slot, const Splat2D* splats, uint32_t k,
                                      bool clamp_src) {
  const float4* rp = reinterpret_cast<const float4*>(splats + k);
  const float4 a = rp[0];  // cx cy ux uy
  const float4 b = rp[1];  // vx vy r g
  const float4 c = rp[2];  // b a hx hy
  float cr = b.z, cg = b.w, cb = c.x;
  if (clamp_src) {
    cr = clamp01(cr);
    cg = clamp01(cg);
    cb = clamp01(cb);
  }
  s.geo[slot] = a;
  s.geo2[slot] = make_float4(b.x, b.y, c.y, c.z);
  s.col[slot] = make_float4(cr, cg, cb, c.w);
  s.bound[slot] = make_float4(a.x - c.z, a.x + c.z, a.y - c.w, a.y + c.w);
}

// Fragment alpha (render.wesl:143-166) of a staged splat at pixel centre
// (px, py); negative when the fragment is discarded.
__device__ __forceinline__ float fragment(int mode, float sd2, float ell2, const float4& g, const float4& g2,
                                          const float4& col, float px, float py, bool fast_exp) {
  const float dx = px - g.x;
  const float dy = py - g.y;
  if (mode == kModePoint) return (dx >= -g2.w && dx < g2.w && dy >= -col.w && dy < col.w) ? 1.0f : -1.0f;
  const float qx = dx * g.z + dy * g.w;
  const float qy = dx * g2.x + dy * g2.y;
  const float r2 = qx * qx + qy * qy;
  if (!(r2 <= sd2)) return -1.0f;
  if (mode == kModeEllipse) return g2.z + (1.0f - g2.z) * ((r2 > ell2) ? 1.0f : 0.0f);
  return g2.z * (fast_exp ? __expf(-r2) : expf(-r2));
}

// true when the splat's bound misses the wave's 8x8 block of pixel centres
// [x0 + 0.5, x0 + 7.5] x [y0 + 0.5, y0 + 7.5] (with half a pixel of margin)
__device__ __forceinline__ bool misses_block(const float4& b, float x0, float y0) {
  const bool miss = (b.y < x0) || (b.x > x0 + 8.0f) || (b.w < y0) || (b.z > y0 + 8.0f);
  return __builtin_amdgcn_readfirstlane(int(miss)) != 0;
}

struct PixelSetup {
  uint32_t px, py;
  bool inside;
  float fpx, fpy, bx0, by0;
};

__device__ __forceinline__ PixelSetup setup_pixel(const TileGrid& g, uint32_t tile) {
  PixelSetup p;
  const uint32_t tx = tile % g.tiles_x, ty = tile / g.tiles_x;
  const uint32_t w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const uint32_t blocks_x = g.tile_w / 8;
  const uint32_t bx = tx * g.tile_w + (w % blocks_x) * 8;
  const uint32_t by = ty * g.tile_h + (w / blocks_x) * 8;
  p.px = bx + (lane & 7u);
  p.py = by + (lane >> 3);
  p.inside = p.px < g.width && p.py < g.height;
  p.fpx = float(p.px) + 0.5f;
  p.fpy = float(p.py) + 0.5f;
  p.bx0 = float(bx);
  p.by0 = float(by);
  return p;
}

__device__ __forceinline__ void store_pixel(const CompositeArgs& a, const PixelSetup& p, float r, float g, float b) {
  const size_t o = size_t(p.py) * a.grid.width + p.px;
  if (a.format == kFormatRgba8Unorm)
    reinterpret_cast<uint32_t*>(a.out)[o] = unorm8(r) | (unorm8(g) << 8) | (unorm8(b) << 16) | (255u << 24);
  else
    reinterpret_cast<float4*>(a.out)[o] = make_float4(r, g, b, 1.0f);
}

__device__ __forceinline__ uint32_t segments_of(uint32_t len, uint32_t seg) {
  return len <= seg ? 1u : (len + seg - 1u) / seg;
}

// Work list: one item per (tile, segment); items of multi-segment tiles come
// first. One workgroup scans all tiles (14,400 at 2560x1440 in 16x16).
__global__ __launch_bounds__(kPlanThreads) void k_plan(CompositeArgs a) {
  __shared__ uint32_t s_sum[kPlanThreads / kWave];
  __shared__ uint32_t s_heavy_total;
  const uint32_t nt = a.grid.tiles_x * a.grid.tiles_y;
  const uint32_t per = (nt + kPlanThreads - 1) / kPlanThreads;
  const uint32_t t0 = threadIdx.x * per;
  const uint32_t t1 = (t0 + per < nt) ? t0 + per : nt;
  const uint32_t lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  for (int pass = 0; pass < 2; ++pass) {
    // pass 0: multi-segment tiles (their segments); pass 1: the other tiles
    uint32_t cnt = 0;
    for (uint32_t t = t0; t < t1; ++t) {
      const uint32_t ns = segments_of(a.ranges[2 * t + 1] - a.ranges[2 * t], a.seg_len);
      cnt += (pass == 0) ? (ns > 1 ? ns : 0u) : (ns > 1 ? 0u : 1u);
    }
    uint32_t x = cnt;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t y = __shfl_up(x, off, kWave);
      if (lane >= uint32_t(off)) x += y;
    }
    if (lane == kWave - 1) s_sum[w] = x;
    __syncthreads();
    uint32_t pre = 0;
    for (uint32_t q = 0; q < w; ++q) pre += s_sum[q];
    uint32_t total = 0;
    for (uint32_t q = 0; q < kPlanThreads / kWave; ++q) total += s_sum[q];
    __syncthreads();
    uint32_t item = (pass == 0 ? 0u : s_heavy_total) + pre + x - cnt;
    const uint32_t max_items = nt + a.capacity / a.seg_len + 1u;  // composite_max_items
    for (uint32_t t = t0; t < t1; ++t) {
      const uint32_t ns = segments_of(a.ranges[2 * t + 1] - a.ranges[2 * t], a.seg_len);
      if (pass == 0 && ns > 1) {
        a.seg_base[t] = item;
        for (uint32_t s = 0; s < ns; ++s, ++item) {
          if (item < max_items) {
            a.item_tile[item] = t;
            a.item_seg[item] = s;
          }
        }
      } else if (pass == 1 && ns <= 1) {
        if (item < max_items) {
          a.item_tile[item] = t;
          a.item_seg[item] = 0u;
        }
        ++item;
      }
    }
    if (threadIdx.x == 0) {
      if (pass == 0) {
        s_heavy_total = total;
        a.plan[1] = total;  // multi-segment items
      } else {
        // all items (never more than the list holds; the bound is exact
        // for lists built from at most `capacity` pairs)
        const uint32_t max_items = nt + a.capacity / a.seg_len + 1u;
        const uint32_t all = s_heavy_total + total;
        a.plan[0] = all < max_items ? all : max_items;
        a.plan[2] = 0u;  // work ticket
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kCompMaxThreads) void k_composite_fast(CompositeArgs a) {
#pragma clang fp contract(fast)
  __shared__ StagedSplats s;
  __shared__ uint32_t s_item;
  const FrameConsts& fc = *a.fc;
  const int mode = fc.display_mode;
  const float sd2 = fc.max_std_dev * fc.max_std_dev;
  const float e = fc.max_std_dev - 0.1f;
  const float ell2 = e * e;
  const bool q8 = a.format == kFormatRgba8Unorm;
  const uint32_t nt = blockDim.x, tid = threadIdx.x;
  const uint32_t items = a.plan[0], heavy = a.plan[1];

  for (;;) {
    const uint32_t item = take_ticket(a.plan + 2, &s_item);
    if (item >= items) break;
    const uint32_t tile = a.item_tile[item], seg = a.item_seg[item];
    const PixelSetup p = setup_pixel(a.grid, tile);
    const uint32_t tstart = a.ranges[2 * tile], tend = a.ranges[2 * tile + 1];
    // segment [start, end) counted from the front (nearest) of the list
    const uint32_t front = tend - seg * a.seg_len;
    const uint32_t back = (item < heavy && front - tstart > a.seg_len) ? front - a.seg_len : tstart;

    float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
    bool done = !p.inside;
    for (uint32_t bend = front; bend > back;) {
      const uint32_t cnt = (bend - back) < nt ? (bend - back) : nt;
      __syncthreads();
      if (tid < cnt) stage(s, tid, a.splats, a.pair_vals[bend - 1 - tid], q8);  // nearest first
      __syncthreads();
      if (!__all(done)) {
        for (uint32_t q = 0; q < cnt; ++q) {
          if (misses_block(s.bound[q], p.bx0, p.by0)) continue;
          if (done) continue;
          const float4 g = s.geo[q], g2 = s.geo2[q], col = s.col[q];
          const float al = fragment(mode, sd2, ell2, g, g2, col, p.fpx, p.fpy, true);
          if (al < 0.0f) continue;
          const float wgt = T * al;
          cr += wgt * col.x;
          cg += wgt * col.y;
          cb += wgt * col.z;
          T *= 1.0f - al;
          done = T < kMinTransmittance;
        }
      }
      bend -= cnt;
      if (__syncthreads_and(done ? 1 : 0)) break;
    }
    if (p.inside) {
      if (item < heavy)
        a.partial[size_t(item) * nt + tid] = make_float4(cr, cg, cb, T);
      else
        store_pixel(a, p, cr, cg, cb);
    }
  }
}

// Front-to-back merge of the segments of multi-segment tiles.
__global__ __launch_bounds__(kCompMaxThreads) void k_merge(CompositeArgs a) {
#pragma clang fp contract(fast)
  const uint32_t tile = blockIdx.x;
  const uint32_t len = a.ranges[2 * tile + 1] - a.ranges[2 * tile];
  const uint32_t ns = segments_of(len, a.seg_len);
  if (ns <= 1) return;
  const PixelSetup p = setup_pixel(a.grid, tile);
  if (!p.inside) return;
  const uint32_t base = a.seg_base[tile];
  float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
  for (uint32_t s = 0; s < ns && T >= kMinTransmittance; ++s) {
    const float4 v = a.partial[size_t(base + s) * blockDim.x + threadIdx.x];
    cr += T * v.x;
    cg += T * v.y;
    cb += T * v.z;
    T *= v.w;
  }
  store_pixel(a, p, cr, cg, cb);
}

__global__ __launch_bounds__(kCompMaxThreads) void k_composite_exact(CompositeArgs a) {
  __shared__ StagedSplats s;
  const uint32_t tile = blockIdx.x;
  const PixelSetup p = setup_pixel(a.grid, tile);
  const uint32_t start = a.ranges[2 * tile], end = a.ranges[2 * tile + 1];
  const FrameConsts& fc = *a.fc;
  const int mode = fc.display_mode;
  const float sd2 = fc.max_std_dev * fc.max_std_dev;
  const float e = fc.max_std_dev - 0.1f;
  const float ell2 = e * e;
  const bool q8 = a.format == kFormatRgba8Unorm;
  const uint32_t nt = blockDim.x, tid = threadIdx.x;

  float dr = 0.0f, dg = 0.0f, db = 0.0f, da = 1.0f;
  for (uint32_t b0 = start; b0 < end; b0 += nt) {
    const uint32_t cnt = (end - b0) < nt ? (end - b0) : nt;
    __syncthreads();
    if (tid < cnt) stage(s, tid, a.splats, a.pair_vals[b0 + tid], q8);  // farthest first
    __syncthreads();
    if (!p.inside) continue;
    for (uint32_t q = 0; q < cnt; ++q) {
      if (misses_block(s.bound[q], p.bx0, p.by0)) continue;
      const float4 g = s.geo[q], g2 = s.geo2[q], col = s.col[q];
      const float al = fragment(mode, sd2, ell2, g, g2, col, p.fpx, p.fpy, false);
      if (al < 0.0f) continue;
      const float om = 1.0f - al;
      dr = col.x * al + dr * om;
      dg = col.y * al + dg * om;
      db = col.z * al + db * om;
      da = al + da * om;
      if (q8) {
        dr = float(unorm8(dr)) / 255.0f;
        dg = float(unorm8(dg)) / 255.0f;
        db = float(unorm8(db)) / 255.0f;
        da = float(unorm8(da)) / 255.0f;
      }
    }
  }
  if (p.inside) {
    const size_t o = size_t(p.py) * a.grid.width + p.px;
    if (q8)
      reinterpret_cast<uint32_t*>(a.out)[o] = unorm8(dr) | (unorm8(dg) << 8) | (unorm8(db) << 16) | (unorm8(da) << 24);
    else
      reinterpret_cast<float4*>(a.out)[o] = make_float4(dr, dg, db, da);
  }
}

uint32_t composite_max_items(uint32_t num_tiles, uint32_t capacity, uint32_t seg_len) {
  // every tile has one item; a tile of len > seg_len has ceil(len/seg_len)
  // <= len/seg_len + 1 items, and the lens sum to at most capacity
  return num_tiles + capacity / seg_len + 1;
}

uint32_t composite_max_heavy_items(uint32_t capacity, uint32_t seg_len) {
  // multi-segment tiles have len > seg_len, so ceil(len/seg_len) < 2 len/seg_len
  return 2u * (capacity / seg_len) + 2u;
}

void launch_composite(hipStream_t st, const CompositeArgs& a, int num_cus) {
  const uint32_t tiles = a.grid.tiles_x * a.grid.tiles_y;
  const uint32_t threads = a.grid.tile_w * a.grid.tile_h;
  if (tiles == 0) return;
  if (a.blend == kBlendFast) {
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(kPlanThreads), 0, st, a);
    // persistent workgroups take items by ticket; 8 per CU of 256 lanes
    uint32_t grid = uint32_t(num_cus) * 8u * (kCompMaxThreads / threads);
    const uint32_t cap = composite_max_items(tiles, a.capacity, a.seg_len);
    if (grid > cap) grid = cap;
    hipLaunchKernelGGL(k_composite_fast, dim3(grid), dim3(threads), 0, st, a);
    hipLaunchKernelGGL(k_merge, dim3(tiles), dim3(threads), 0, st, a);
  } else {
    hipLaunchKernelGGL(k_composite_exact, dim3(tiles), dim3(threads), 0, st, a);
  }
}

}  // namespace gs3d
