// Tile compositing: replaces the reference's fixed-function rasterizer and
// ALPHA_BLENDING over draw_indirect quads (src/renderer.rs:163-195, 284-308;
// fragment stage render.wesl:143-181).
//
// One workgroup per work item, one lane per pixel of a screen tile (8x8,
// 16x16 or 32x8 pixels); each wave owns an 8x8 block of the tile. The tile's
// splats (its run of the tile-sorted pairs, in draw order) are staged through
// LDS in batches of one splat per lane. While staging, every splat's pixel
// bound is tested against the tile's 8x8 blocks and each wave receives, by
// ballot compaction, the ordered list of the staged splats that can touch its
// block: a wave then evaluates only those (most splats of dense scenes cover
// a few pixels and touch one or two blocks).
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
constexpr int kCompMaxWaves = kCompMaxThreads / kWave;
constexpr int kPlanThreads = 1024;
constexpr float kLog2e = 1.4426950408889634f;

// Lanes stop once what is left behind contributes at most 1e-4 of a colour.
constexpr float kMinTransmittance = 1e-4f;

__device__ __forceinline__ uint32_t unorm8(float x) {
  x = fminf(fmaxf(x, 0.0f), 1.0f);
  return uint32_t(int(x * 255.0f + 0.5f));
}

__device__ __forceinline__ float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

__device__ __forceinline__ float h_lo(uint32_t w) { return __half2float(__ushort_as_half((unsigned short)(w & 0xFFFFu))); }
__device__ __forceinline__ float h_hi(uint32_t w) { return __half2float(__ushort_as_half((unsigned short)(w >> 16))); }

// One batch of staged splats plus, per wave, the ordered slots that touch
// the wave's 8x8 block.
struct Stage {
  float4 geo[kCompMaxThreads];  // splat: cx, cy, ca*log2e, 2*cb*log2e   point: cx, cy, hx, hy
  float4 col[kCompMaxThreads];  // cc*log2e, opacity, r, g
  float blue[kCompMaxThreads];
  uint16_t list[kCompMaxWaves][kCompMaxThreads];
  uint32_t wcount[kCompMaxWaves][kCompMaxWaves];  // [staging wave][target wave]
};

struct Tiling {
  uint32_t px, py;
  bool inside;
  float fpx, fpy;
  uint32_t waves, blocks_x;
  float ox, oy;  // pixel origin of the tile
};

__device__ __forceinline__ Tiling setup_tile(const TileGrid& g, uint32_t tile) {
  Tiling t;
  const uint32_t tx = tile % g.tiles_x, ty = tile / g.tiles_x;
  const uint32_t w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  t.blocks_x = g.tile_w / 8;
  t.waves = blockDim.x / kWave;
  t.ox = float(tx * g.tile_w);
  t.oy = float(ty * g.tile_h);
  t.px = tx * g.tile_w + (w % t.blocks_x) * 8 + (lane & 7u);
  t.py = ty * g.tile_h + (w / t.blocks_x) * 8 + (lane >> 3);
  t.inside = t.px < g.width && t.py < g.height;
  t.fpx = float(t.px) + 0.5f;
  t.fpy = float(t.py) + 0.5f;
  return t;
}

// Stage splat k (or nothing when !valid) into slot threadIdx.x and build the
// per-wave lists. Contains __syncthreads; every lane must call it.
__device__ __forceinline__ void stage_batch(Stage& s, const Tiling& t, const Splat2D* splats, bool valid, uint32_t k,
                                            bool clamp_src, int mode) {
  const uint32_t tid = threadIdx.x, lane = tid % kWave, sw = tid / kWave;
  uint32_t mask = 0u;
  if (valid) {
    const Splat2D r = splats[k];
    float cr = h_lo(r.rg), cg = h_hi(r.rg), cb = h_lo(r.ba);
    const float op = h_hi(r.ba);
    if (clamp_src) {
      cr = clamp01(cr);
      cg = clamp01(cg);
      cb = clamp01(cb);
    }
    if (mode == kModePoint) {
      s.geo[tid] = make_float4(r.cx, r.cy, r.ca, r.cb);
      s.col[tid] = make_float4(0.0f, op, cr, cg);
    } else {
      s.geo[tid] = make_float4(r.cx, r.cy, r.ca * kLog2e, 2.0f * r.cb * kLog2e);
      s.col[tid] = make_float4(r.cc * kLog2e, op, cr, cg);
    }
    s.blue[tid] = cb;
    // which 8x8 blocks of the tile the bound [c - h, c + h] can touch (pixel
    // centres of block (i, j) span [o + 8i + 0.5, o + 8i + 7.5]; half a pixel
    // of margin)
    const float hx = h_lo(r.hxy), hy = h_hi(r.hxy);
    const float x0 = r.cx - hx - t.ox, x1 = r.cx + hx - t.ox;
    const float y0 = r.cy - hy - t.oy, y1 = r.cy + hy - t.oy;
    for (uint32_t w = 0; w < t.waves; ++w) {
      const float bx = float((w % t.blocks_x) * 8), by = float((w / t.blocks_x) * 8);
      const bool hit = !(x1 < bx) && !(x0 > bx + 8.0f) && !(y1 < by) && !(y0 > by + 8.0f);
      mask |= hit ? (1u << w) : 0u;
    }
  }
  uint64_t bal[kCompMaxWaves];
#pragma unroll
  for (int w = 0; w < kCompMaxWaves; ++w) {
    bal[w] = __ballot((mask >> w) & 1u);
    if (lane == 0) s.wcount[sw][w] = uint32_t(__popcll(bal[w]));
  }
  __syncthreads();
#pragma unroll
  for (int w = 0; w < kCompMaxWaves; ++w) {
    if ((mask >> w) & 1u) {
      uint32_t off = 0;
      for (uint32_t q = 0; q < sw; ++q) off += s.wcount[q][w];
      s.list[w][off + mask_rank(bal[w])] = uint16_t(tid);
    }
  }
  __syncthreads();
}

__device__ __forceinline__ uint32_t wave_list_len(const Stage& s, const Tiling& t, uint32_t w) {
  uint32_t n = 0;
  for (uint32_t q = 0; q < t.waves; ++q) n += s.wcount[q][w];
  return n;
}

// Fragment alpha (render.wesl:143-166) of staged slot q at this lane's pixel
// centre; negative when the fragment is discarded. sd2l = sd^2 log2(e),
// ell2l = (sd - 0.1)^2 log2(e).
__device__ __forceinline__ float fragment(const Stage& s, uint32_t q, const Tiling& t, int mode, float sd2l,
                                          float ell2l, float4& col) {
  const float4 g = s.geo[q];
  col = s.col[q];
  const float dx = t.fpx - g.x;
  const float dy = t.fpy - g.y;
  if (mode == kModePoint) return (dx >= -g.z && dx < g.z && dy >= -g.w && dy < g.w) ? 1.0f : -1.0f;
  // |qo|^2 log2(e)
  const float r2 = dx * (g.z * dx + g.w * dy) + col.x * dy * dy;
  if (!(r2 <= sd2l)) return -1.0f;
  if (mode == kModeEllipse) return col.y + (1.0f - col.y) * ((r2 > ell2l) ? 1.0f : 0.0f);
  return col.y * __builtin_amdgcn_exp2f(-r2);
}

// Final pixel: the front-to-back colour plus what is left of the background
// (black with alpha 1 after a clear; the target's contents when loading).
__device__ __forceinline__ void store_pixel(const CompositeArgs& a, const Tiling& t, float r, float g, float b,
                                            float T) {
  const size_t o = size_t(t.py) * a.grid.width + t.px;
  if (a.format == kFormatRgba8Unorm) {
    uint32_t* px = reinterpret_cast<uint32_t*>(a.out) + o;
    float da = 1.0f;
    if (a.load) {
      const uint32_t d = *px;
      r += T * (float(d & 0xFFu) / 255.0f);
      g += T * (float((d >> 8) & 0xFFu) / 255.0f);
      b += T * (float((d >> 16) & 0xFFu) / 255.0f);
      da = (1.0f - T) + T * (float(d >> 24) / 255.0f);
    }
    *px = unorm8(r) | (unorm8(g) << 8) | (unorm8(b) << 16) | (unorm8(da) << 24);
  } else {
    float4* px = reinterpret_cast<float4*>(a.out) + o;
    float da = 1.0f;
    if (a.load) {
      const float4 d = *px;
      r += T * d.x;
      g += T * d.y;
      b += T * d.z;
      da = (1.0f - T) + T * d.w;
    }
    *px = make_float4(r, g, b, da);
  }
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
  const uint32_t max_items = nt + a.capacity / a.seg_len + 1u;  // composite_max_items
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
    for (uint32_t t = t0; t < t1; ++t) {
      const uint32_t ns = segments_of(a.ranges[2 * t + 1] - a.ranges[2 * t], a.seg_len);
      if (pass == 0 && ns > 1) {
        a.seg_base[t] = item;
        for (uint32_t sg = 0; sg < ns; ++sg, ++item) {
          if (item < max_items) {
            a.item_tile[item] = t;
            a.item_seg[item] = sg;
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
        // all items; never more than the list holds (the bound is exact for
        // lists built from at most `capacity` pairs)
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
  __shared__ Stage s;
  __shared__ uint32_t s_item;
  const FrameConsts& fc = *a.fc;
  const int mode = fc.display_mode;
  const float sd2l = fc.max_std_dev * fc.max_std_dev * kLog2e;
  const float e = fc.max_std_dev - 0.1f;
  const float ell2l = e * e * kLog2e;
  const bool q8 = a.format == kFormatRgba8Unorm;
  const uint32_t nt = blockDim.x, tid = threadIdx.x, w = tid / kWave;
  const uint32_t items = a.plan[0], heavy = a.plan[1];

  for (;;) {
    const uint32_t item = take_ticket(a.plan + 2, &s_item);
    if (item >= items) break;
    const uint32_t tile = a.item_tile[item], seg = a.item_seg[item];
    const Tiling t = setup_tile(a.grid, tile);
    const uint32_t tstart = a.ranges[2 * tile], tend = a.ranges[2 * tile + 1];
    // segment counted from the front (nearest end) of the tile's list
    const uint32_t front = tend - seg * a.seg_len;
    const uint32_t back = (item < heavy && front - tstart > a.seg_len) ? front - a.seg_len : tstart;

    float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
    bool done = !t.inside;
    for (uint32_t bend = front; bend > back;) {
      const uint32_t cnt = (bend - back) < nt ? (bend - back) : nt;
      __syncthreads();
      const bool valid = tid < cnt;
      stage_batch(s, t, a.splats, valid, valid ? a.pair_vals[bend - 1 - tid] : 0u, q8, mode);  // nearest first
      const uint32_t len = wave_list_len(s, t, w);
      for (uint32_t j = 0; j < len; ++j) {
        if ((j & 31u) == 0u && __all(done)) break;
        if (done) continue;
        const uint32_t q = s.list[w][j];
        float4 col;
        const float al = fragment(s, q, t, mode, sd2l, ell2l, col);
        if (al < 0.0f) continue;
        const float wgt = T * al;
        cr += wgt * col.z;
        cg += wgt * col.w;
        cb += wgt * s.blue[q];
        T *= 1.0f - al;
        done = T < kMinTransmittance;
      }
      bend -= cnt;
      if (__syncthreads_and(done ? 1 : 0)) break;
    }
    if (t.inside) {
      if (item < heavy)
        a.partial[size_t(item) * nt + tid] = make_float4(cr, cg, cb, T);
      else
        store_pixel(a, t, cr, cg, cb, T);
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
  const Tiling t = setup_tile(a.grid, tile);
  if (!t.inside) return;
  const uint32_t base = a.seg_base[tile];
  float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
  for (uint32_t sg = 0; sg < ns && T >= kMinTransmittance; ++sg) {
    const float4 v = a.partial[size_t(base + sg) * blockDim.x + threadIdx.x];
    cr += T * v.x;
    cg += T * v.y;
    cb += T * v.z;
    T *= v.w;
  }
  store_pixel(a, t, cr, cg, cb, T);
}

__global__ __launch_bounds__(kCompMaxThreads) void k_composite_exact(CompositeArgs a) {
  __shared__ Stage s;
  const uint32_t tile = blockIdx.x;
  const Tiling t = setup_tile(a.grid, tile);
  const uint32_t start = a.ranges[2 * tile], end = a.ranges[2 * tile + 1];
  const FrameConsts& fc = *a.fc;
  const int mode = fc.display_mode;
  const float sd2l = fc.max_std_dev * fc.max_std_dev * kLog2e;
  const float e = fc.max_std_dev - 0.1f;
  const float ell2l = e * e * kLog2e;
  const bool q8 = a.format == kFormatRgba8Unorm;
  const uint32_t nt = blockDim.x, tid = threadIdx.x, w = tid / kWave;

  uint32_t frags = 0;
  float dr = 0.0f, dg = 0.0f, db = 0.0f, da = 1.0f;
  if (a.load && t.inside) {
    const size_t o = size_t(t.py) * a.grid.width + t.px;
    if (q8) {
      const uint32_t d = reinterpret_cast<const uint32_t*>(a.out)[o];
      dr = float(d & 0xFFu) / 255.0f;
      dg = float((d >> 8) & 0xFFu) / 255.0f;
      db = float((d >> 16) & 0xFFu) / 255.0f;
      da = float(d >> 24) / 255.0f;
    } else {
      const float4 d = reinterpret_cast<const float4*>(a.out)[o];
      dr = d.x;
      dg = d.y;
      db = d.z;
      da = d.w;
    }
  }
  for (uint32_t b0 = start; b0 < end; b0 += nt) {
    const uint32_t cnt = (end - b0) < nt ? (end - b0) : nt;
    __syncthreads();
    const bool valid = tid < cnt;
    stage_batch(s, t, a.splats, valid, valid ? a.pair_vals[b0 + tid] : 0u, q8, mode);  // farthest first
    if (!t.inside) continue;
    const uint32_t len = wave_list_len(s, t, w);
    for (uint32_t j = 0; j < len; ++j) {
      const uint32_t q = s.list[w][j];
      float4 col;
      const float al = fragment(s, q, t, mode, sd2l, ell2l, col);
      if (al < 0.0f) continue;
      ++frags;
      const float om = 1.0f - al;
      dr = col.z * al + dr * om;
      dg = col.w * al + dg * om;
      db = s.blue[q] * al + db * om;
      da = al + da * om;
      if (q8) {
        dr = float(unorm8(dr)) / 255.0f;
        dg = float(unorm8(dg)) / 255.0f;
        db = float(unorm8(db)) / 255.0f;
        da = float(unorm8(da)) / 255.0f;
      }
    }
  }
  if (t.inside) {
    const size_t o = size_t(t.py) * a.grid.width + t.px;
    if (q8)
      reinterpret_cast<uint32_t*>(a.out)[o] = unorm8(dr) | (unorm8(dg) << 8) | (unorm8(db) << 16) | (unorm8(da) << 24);
    else
      reinterpret_cast<float4*>(a.out)[o] = make_float4(dr, dg, db, da);
    if (a.frag_count && frags) atomicAdd(a.frag_count, frags);
  }
}

// LoadOp::Clear(BLACK) of a whole target (src/renderer.rs:175), for passes
// that draw nothing.
__global__ __launch_bounds__(256) void k_clear(void* out, uint32_t pixels, int format) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < pixels; i += gridDim.x * blockDim.x) {
    if (format == kFormatRgba8Unorm)
      reinterpret_cast<uint32_t*>(out)[i] = 255u << 24;
    else
      reinterpret_cast<float4*>(out)[i] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
  }
}

void launch_clear(hipStream_t st, void* out, uint32_t pixels, int format) {
  uint32_t blocks = (pixels + 255u) / 256u;
  if (blocks > 4096u) blocks = 4096u;
  if (blocks < 1u) blocks = 1u;
  hipLaunchKernelGGL(k_clear, dim3(blocks), dim3(256), 0, st, out, pixels, format);
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

void launch_composite(hipStream_t st, const CompositeArgs& a, int num_cus, hipEvent_t main_begin,
                      hipEvent_t main_end) {
  const uint32_t tiles = a.grid.tiles_x * a.grid.tiles_y;
  const uint32_t threads = a.grid.tile_w * a.grid.tile_h;
  if (tiles == 0) return;
  if (a.blend == kBlendFast) {
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(kPlanThreads), 0, st, a);
    // persistent workgroups take items by ticket; 8 per CU of 256 lanes
    uint32_t grid = uint32_t(num_cus) * 8u * (kCompMaxThreads / threads);
    const uint32_t cap = composite_max_items(tiles, a.capacity, a.seg_len);
    if (grid > cap) grid = cap;
    if (main_begin) (void)hipEventRecord(main_begin, st);
    hipLaunchKernelGGL(k_composite_fast, dim3(grid), dim3(threads), 0, st, a);
    if (main_end) (void)hipEventRecord(main_end, st);
    hipLaunchKernelGGL(k_merge, dim3(tiles), dim3(threads), 0, st, a);
  } else {
    if (main_begin) (void)hipEventRecord(main_begin, st);
    hipLaunchKernelGGL(k_composite_exact, dim3(tiles), dim3(threads), 0, st, a);
    if (main_end) (void)hipEventRecord(main_end, st);
  }
}

}  // namespace gs3d
