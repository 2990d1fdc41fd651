// Tile compositing: replaces the reference's fixed-function rasterizer and
// ALPHA_BLENDING over draw_indirect quads (src/renderer.rs:163-195, 284-308;
// fragment stage render.wesl:143-181).
//
// One workgroup per work item, one lane per pixel of a screen tile (8x8,
// 16x16 or 32x8 pixels); each wave owns an 8x8 block of the tile. The tile's
// splats (its run of the tile-sorted pairs, in draw order) are staged through
// LDS in batches of one splat per lane. While staging, every splat's pixel
// bound is tested against the tile's 4x4 quadrants, and each quadrant
// receives, by ballot compaction, the ordered list of the staged splats that
// can touch it. Most splats of dense scenes cover a few pixels: each 16-lane
// quadrant blends only its own list.
//
// fast  : front to back in f32 with transmittance. The over operator is
//         associative, so a tile whose list is longer than one segment is
//         split into segments composited in parallel, each to a partial
//         (colour, transmittance) per pixel, and merged in order afterwards:
//         the few very dense tiles of a clustered scene no longer set the
//         kernel's length. Multi-segment tiles are scheduled first. A wave
//         stops when the transmittance of all its lanes is below 1e-4, a
//         workgroup when all its waves have.
// exact : back to front, one workgroup per tile; the destination is rounded
//         to the target format after every splat, as the blender does for an
//         Rgba8Unorm target (rounding is not associative: no segments).
#include <hip/hip_runtime.h>

#include "device_util.hpp"
#include "kernels.hpp"
#include "types.hpp"

namespace gs3d {

constexpr int kCompMaxThreads = 256;
constexpr float kLog2e = 1.4426950408889634f;
constexpr float kSqrtLog2e = 1.2011224087864498f;

// Lanes stop once what is left behind contributes at most 1e-4 of a colour.
constexpr float kMinTransmittance = 1e-4f;

__device__ __forceinline__ uint32_t unorm8(float x) {
  x = fminf(fmaxf(x, 0.0f), 1.0f);
  return uint32_t(int(x * 255.0f + 0.5f));
}

__device__ __forceinline__ float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

__device__ __forceinline__ float h_lo(uint32_t w) { return __half2float(__ushort_as_half((unsigned short)(w & 0xFFFFu))); }
__device__ __forceinline__ float h_hi(uint32_t w) { return __half2float(__ushort_as_half((unsigned short)(w >> 16))); }

// One batch of staged splats plus, per 4x4 quadrant of the tile, the ordered
// slots that touch it.
// Slot NT (one past the staged splats) holds a splat of opacity 0 (blending
// it changes nothing): lanes whose list is shorter than the wave's longest
// blend it. List rows carry kListPad spare entries so that a step may read
// past a list's end.
constexpr uint32_t kListPad = 4;

// LDS of a workgroup of NT lanes (one per pixel of its tile), with one list
// per 4x4 quadrant (NT / 16 lists). Sized per tile shape, so that 8x8 tiles
// (one wave) take a quarter of a 256-lane tile's LDS and as many workgroups
// fit a CU.
template <int NT, bool Z = true>
struct StageT {
  static constexpr int kWaves = NT / kWave;
  static constexpr int kQuads = 4 * kWaves;
  static constexpr uint32_t kDummy = NT;
  // splat / ellipse: |qo|^2 log2(e) = (p.u' - k1)^2 + (p.v' - k2)^2 for the
  // pixel centre p relative to the tile's 32 x 16 cell, u' = u sqrt(log2(e)),
  // v' likewise, k1 = c.u', k2 = c.v' for the splat's centre c (cell-relative)
  float4 geo[NT + 1];  // splat: u'.x, u'.y, v'.x, v'.y   point: c.x, c.y, hx, hy
  float4 col[NT + 1];  // splat: k1, k2, opacity (fast splat: log2 opacity), r   point: 0, 0, opacity, r
  float2 gb[NT + 1];   // g, b
  float z[Z ? NT + 1 : 1];  // fragment depth, for depth-tested draws (Z)
  uint16_t list[kQuads][NT + kListPad];
  uint64_t qbal[kWaves][kQuads];  // [staging wave][quadrant] ballots
};
using Stage = StageT<kCompMaxThreads>;

// The fast path stages log2(opacity) for splat mode (its alpha is then one
// exp2 of log2(opacity) - |qo|^2 log2(e)): the dummy's opacity 0 is -inf.
template <int NT, bool Z>
__device__ __forceinline__ void init_dummy(StageT<NT, Z>& s, int mode) {
  constexpr uint32_t kDummy = StageT<NT, Z>::kDummy;
  if (threadIdx.x == 0) {
    s.geo[kDummy] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    s.col[kDummy] = make_float4(0.0f, 0.0f, mode == kModeSplat ? -__builtin_inff() : 0.0f, 0.0f);
    s.gb[kDummy] = make_float2(0.0f, 0.0f);
    if (Z) s.z[kDummy] = 0.0f;
  }
}

struct Tiling {
  uint32_t px, py;
  bool inside;
  float fpx, fpy;
  uint32_t waves, blocks_x;
  float ox, oy;  // pixel origin of the tile
  // origin of the 32 x 16 cell holding the tile (every tile shape is a
  // sub-tile of one): splat centres and pixel centres are taken relative to
  // it, so a pixel's arithmetic does not depend on the tile shape
  float sx, sy;
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
  const uint32_t cx0 = (tx * g.tile_w) & ~31u, cy0 = (ty * g.tile_h) & ~15u;
  t.sx = float(cx0);
  t.sy = float(cy0);
  t.fpx = float(t.px - cx0) + 0.5f;
  t.fpy = float(t.py - cy0) + 0.5f;
  return t;
}

// Stage one splat record into slot tid. `log2op`: store log2(opacity) (the
// fast path's splat mode) instead of the opacity.
template <int NT, bool Z>
__device__ __forceinline__ void stage_splat(StageT<NT, Z>& s, uint32_t tid, const Splat2D& r, const Tiling& t,
                                            bool clamp_src, int mode, bool log2op) {
  float cr = h_lo(r.rg), cg = h_hi(r.rg), cb = h_lo(r.ba);
  const float op = h_hi(r.ba);
  if (clamp_src) {
    cr = clamp01(cr);
    cg = clamp01(cg);
    cb = clamp01(cb);
  }
  const float cx = r.cx - t.sx, cy = r.cy - t.sy;
  if (mode == kModePoint) {
    s.geo[tid] = make_float4(cx, cy, r.ux, r.uy);
    s.col[tid] = make_float4(0.0f, 0.0f, op, cr);
  } else {
    const float ux = r.ux * kSqrtLog2e, uy = r.uy * kSqrtLog2e;
    const float vx = -r.rho * uy, vy = r.rho * ux;
    s.geo[tid] = make_float4(ux, uy, vx, vy);
    // log2(0) = -inf: exp2 gives alpha 0
    s.col[tid] = make_float4(cx * ux + cy * uy, cx * vx + cy * vy, log2op ? __log2f(op) : op, cr);
  }
  s.gb[tid] = make_float2(cg, cb);
}

// Stage splat k (or nothing) into slot threadIdx.x and build one ordered list
// per 4x4 quadrant of the tile (row-major over the tile's quadrants; each
// wave blends the four of its 8x8 block). (Testing each quadrant against the
// splat's ellipse, as the binning does per tile, shortens the lists but costs
// more in staging than it saves in blending: measured.) Contains
// __syncthreads; every lane must call it.
template <int NT, bool Z>
__device__ __forceinline__ void stage_batch_quads(StageT<NT, Z>& s, const Tiling& t, const Splat2D* splats,
                                                  const float* zsrc, bool valid, uint32_t k, bool clamp_src,
                                                  int mode, bool log2op) {
  const uint32_t tid = threadIdx.x, lane = tid % kWave, sw = tid / kWave;
  uint32_t mask = 0u;
  if (valid) {
    const Splat2D r = splats[k];
    stage_splat(s, tid, r, t, clamp_src, mode, log2op);
    if (Z && zsrc) s.z[tid] = zsrc[k];
    // Quadrant column c of the tile holds pixel centres [4c + 0.5, 4c + 3.5]
    // (tile-relative); the bound [x0, x1] may touch it when x1 >= 4c and
    // x0 <= 4c + 4 (half a pixel of margin). Same for rows. Quadrant (c, r)
    // is list r ncols + c (row-major over the tile's quadrants).
    const float hx = h_lo(r.hxy), hy = h_hi(r.hxy);
    const float x0 = r.cx - hx - t.ox, x1 = r.cx + hx - t.ox;
    const float y0 = r.cy - hy - t.oy, y1 = r.cy + hy - t.oy;
    const int ncols = int(t.blocks_x) * 2, nrows = int(t.waves / t.blocks_x) * 2;
    const int c0 = max(0, int(ceilf((x0 - 4.0f) * 0.25f))), c1 = min(ncols - 1, int(floorf(x1 * 0.25f)));
    const int r0 = max(0, int(ceilf((y0 - 4.0f) * 0.25f))), r1 = min(nrows - 1, int(floorf(y1 * 0.25f)));
    if (c0 <= c1 && r0 <= r1) {
      const uint32_t colbits = (2u << c1) - (1u << c0);  // bits c0 .. c1
      for (int rr = r0; rr <= r1; ++rr) mask |= colbits << (rr * ncols);
    }
  }
  constexpr int kQuads = StageT<NT, Z>::kQuads;
#pragma unroll
  for (int q = 0; q < kQuads; ++q) {
    const uint64_t bal = __ballot((mask >> q) & 1u);
    if (lane == 0) s.qbal[sw][q] = bal;
  }
  __syncthreads();
  // this wave's first entry in each list: the earlier staging waves' counts
  // (lane q < kQuads holds list q's)
  uint32_t first = 0;
  if (lane < uint32_t(kQuads))
    for (uint32_t p = 0; p < sw; ++p) first += uint32_t(__popcll(s.qbal[p][lane]));
  // every lane appends its slot to the lists of its quadrants at first +
  // its rank in this wave's ballot; one uniform step per list (a per-lane
  // walk over its set bits would run as long as the wave's largest splat)
#pragma unroll
  for (int q = 0; q < kQuads; ++q) {
    const uint32_t off = __builtin_amdgcn_readlane(first, q);
    // the ballot again, rather than kept live across the barrier; quadrant
    // lists hold the slot's byte offset in geo / col (16 per slot)
    const uint64_t bal = __ballot((mask >> q) & 1u);
    if ((mask >> q) & 1u) s.list[q][off + mask_rank(bal)] = uint16_t(tid * 16u);
  }
  __syncthreads();
}

// Fragment alpha (render.wesl:143-166) of staged slot q at this lane's pixel
// centre; negative when the fragment is discarded. sd2l = sd^2 log2(e),
// ell2l = (sd - 0.1)^2 log2(e).
__device__ __forceinline__ float fragment(const Stage& s, uint32_t q, const Tiling& t, int mode, float sd2l,
                                          float ell2l, float4& col) {
  const float4 g = s.geo[q];
  col = s.col[q];
  if (mode == kModePoint) {
    const float dx = t.fpx - g.x;
    const float dy = t.fpy - g.y;
    return (dx >= -g.z && dx < g.z && dy >= -g.w && dy < g.w) ? 1.0f : -1.0f;
  }
  // |qo|^2 log2(e)
  const float t1 = (t.fpx * g.x + t.fpy * g.y) - col.x, t2 = (t.fpx * g.z + t.fpy * g.w) - col.y;
  const float r2 = t1 * t1 + t2 * t2;
  if (!(r2 <= sd2l)) return -1.0f;
  if (mode == kModeEllipse) return col.z + (1.0f - col.z) * ((r2 > ell2l) ? 1.0f : 0.0f);
  return col.z * __builtin_amdgcn_exp2f(-r2);
}

// Depth test of a fragment at depth z against the attachment's value d.
__device__ __forceinline__ bool depth_pass(int cmp, float z, float d) {
  switch (cmp) {
    case kCmpNever: return false;
    case kCmpLess: return z < d;
    case kCmpEqual: return z == d;
    case kCmpLessEqual: return z <= d;
    case kCmpGreater: return z > d;
    case kCmpNotEqual: return z != d;
    case kCmpGreaterEqual: return z >= d;
    default: return true;
  }
}

// Fragment alpha for the fast path: 0 for a discarded fragment (blending
// alpha 0 leaves colour and transmittance unchanged), so the loop has no
// branches. `kept` tells whether the fragment survives the discard test.
template <int MODE>
__device__ __forceinline__ float fragment_alpha(float4 g, float4 c, const Tiling& t, float sd2l, float ell2l,
                                                bool& kept) {
#pragma clang fp contract(fast)
  if (MODE == kModePoint) {
    const float dx = t.fpx - g.x;
    const float dy = t.fpy - g.y;
    kept = dx >= -g.z && dx < g.z && dy >= -g.w && dy < g.w;
    return kept ? c.z : 0.0f;
  }
  const float t1 = __builtin_fmaf(t.fpx, g.x, __builtin_fmaf(t.fpy, g.y, -c.x));
  const float t2 = __builtin_fmaf(t.fpx, g.z, __builtin_fmaf(t.fpy, g.w, -c.y));
  const float r2 = __builtin_fmaf(t1, t1, t2 * t2);
  float al;
  if (MODE == kModeEllipse)
    al = c.z + (1.0f - c.z) * ((r2 > ell2l) ? 1.0f : 0.0f);
  else
    al = __builtin_amdgcn_exp2f(c.z - r2);  // c.z = log2(opacity)
  kept = r2 <= sd2l;
  return kept ? al : 0.0f;
}

// Depth state of one lane of a depth-tested draw: the attachment's value
// before the draw and the depth of the nearest kept, passing fragment.
struct DepthLane {
  int cmp;
  float ref;
  float first;
  bool found;
};

// Splats blended per step of the fast loop (1, 2 or 4): more overlaps the
// LDS reads of a step, fewer keeps registers for occupancy.
#ifndef GS3D_COMPOSITE_UNROLL
#define GS3D_COMPOSITE_UNROLL 2
#endif
constexpr int kUnroll = GS3D_COMPOSITE_UNROLL;
static_assert(kUnroll == 1 || kUnroll == 2 || kUnroll == 4, "unroll must be 1, 2 or 4");
static_assert(kListPad % kUnroll == 0, "a step reads at most kListPad entries past a list");

// Front-to-back blend of the lane's quadrant list, kUnroll entries per step;
// a wave steps through its longest quadrant list and the lanes of shorter
// lists blend the dummy slot. Lanes past the transmittance threshold keep
// blending (their colour only gets closer to the exact sum) until the whole
// wave is done.
template <int MODE, bool DEPTH, int NT>
__device__ __forceinline__ void blend_quads(StageT<NT, DEPTH>& s, const Tiling& t, uint32_t w, float sd2l, float ell2l,
                                            float& T, float& cr, float& cg, float& cb, bool& done, DepthLane& dl) {
#pragma clang fp contract(fast)
  const uint32_t lane = threadIdx.x % kWave;
  // this wave's 8x8 block holds quadrant columns 2 bx, 2 bx + 1 and rows
  // 2 by, 2 by + 1; the lane's quadrant is (qx, qy) within it
  const uint32_t ncols = 2u * t.blocks_x, bx = w % t.blocks_x, by = w / t.blocks_x;
  const uint32_t q00 = 2u * by * ncols + 2u * bx;
  const uint32_t myq = q00 + ((lane >> 5) ? ncols : 0u) + ((lane >> 2) & 1u);
  uint32_t lens[4] = {0u, 0u, 0u, 0u};
  for (uint32_t p = 0; p < t.waves; ++p) {
#pragma unroll
    for (uint32_t qd = 0; qd < 4; ++qd) lens[qd] += uint32_t(__popcll(s.qbal[p][q00 + (qd >> 1) * ncols + (qd & 1u)]));
  }
  const uint32_t mylen = lens[((lane >> 2) & 1u) | ((lane >> 5) << 1)];
  // the same for the whole wave: keep the loop bound in a scalar register
  const uint32_t wmax =
      __builtin_amdgcn_readfirstlane(max(max(lens[0], lens[1]), max(lens[2], lens[3])));
  constexpr uint32_t kDummyOff = StageT<NT, DEPTH>::kDummy * 16u;
  // pad this wave's four lists with the dummy up to the longest one (whole
  // steps), so that every step reads list entries without a bounds check;
  // only this wave reads its lists, and its LDS accesses run in order
  {
    const uint32_t wend = (wmax + uint32_t(kUnroll) - 1u) & ~uint32_t(kUnroll - 1);
    const uint32_t qlane = (lane & 3u) | (((lane >> 3) & 3u) << 2);
    for (uint32_t e = mylen + qlane; e < wend; e += 16u) s.list[myq][e] = uint16_t(kDummyOff);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  const char* geo = reinterpret_cast<const char*>(s.geo);
  const char* col = reinterpret_cast<const char*>(s.col);
  const char* gbs = reinterpret_cast<const char*>(s.gb);
  const char* zs = reinterpret_cast<const char*>(s.z);
  for (uint32_t j = 0; j < wmax; j += kUnroll) {
    if ((j & 31u) == 0u && __all(done)) break;
    uint32_t q[kUnroll];  // byte offsets of the slots in geo / col
    if (kUnroll == 1) {
      q[0] = s.list[myq][j];
    } else {
#pragma unroll
      for (int k = 0; k < kUnroll; k += 2) {
        const uint32_t qq = *reinterpret_cast<const uint32_t*>(&s.list[myq][j + k]);
        q[k] = qq & 0xFFFFu;
        q[k + 1 < kUnroll ? k + 1 : 0] = qq >> 16;
      }
    }
    float4 g[kUnroll], c[kUnroll];
    float2 gb[kUnroll];
    float zz[kUnroll];
#pragma unroll
    for (int k = 0; k < kUnroll; ++k) {
      g[k] = *reinterpret_cast<const float4*>(geo + q[k]);
      c[k] = *reinterpret_cast<const float4*>(col + q[k]);
      gb[k] = *reinterpret_cast<const float2*>(gbs + (q[k] >> 1));
      if (DEPTH) zz[k] = *reinterpret_cast<const float*>(zs + (q[k] >> 2));
    }
#pragma unroll
    for (int k = 0; k < kUnroll; ++k) {
      bool kept;
      float al = fragment_alpha<MODE>(g[k], c[k], t, sd2l, ell2l, kept);
      if (DEPTH) {
        // the test is against the attachment as it was before the draw: for
        // Less / LessEqual and back-to-front order, what depth writes by
        // farther splats would change only matters at equal depths
        const bool pass = depth_pass(dl.cmp, zz[k], dl.ref) && q[k] != kDummyOff;
        al = pass ? al : 0.0f;
        if (kept && pass && !dl.found) {
          dl.first = zz[k];
          dl.found = true;
        }
      }
      const float wgt = T * al;
      cr += wgt * c[k].w;
      cg += wgt * gb[k].x;
      cb += wgt * gb[k].y;
      T = __builtin_fmaf(-T, al, T);
    }
    done = T < kMinTransmittance;
  }
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

// Work list: one item per (tile, segment); the segments of multi-segment
// tiles come first (heavy-first), then one item per other tile. Built in two
// launches over blocks of kPlanTiles tiles: per-block counts, then each block
// scans the block counts before it and writes its items in tile order.
constexpr uint32_t kPlanTiles = 256;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t x, uint32_t* s_wave, uint32_t& total) {
  const uint32_t lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  uint32_t inc = x;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t y = __shfl_up(inc, off, kWave);
    if (lane >= uint32_t(off)) inc += y;
  }
  if (lane == kWave - 1) s_wave[w] = inc;
  __syncthreads();
  uint32_t pre = 0;
  total = 0;
  for (uint32_t q = 0; q < kPlanTiles / kWave; ++q) {
    if (q < w) pre += s_wave[q];
    total += s_wave[q];
  }
  __syncthreads();
  return pre + inc - x;
}

__global__ __launch_bounds__(kPlanTiles) void k_plan_count(CompositeArgs a) {
  if (*a.pair_count > a.capacity) return;  // truncated lists: the frame is drawn again
  __shared__ uint32_t s_wave[2][kPlanTiles / kWave];
  const uint32_t nt = a.grid.tiles_x * a.grid.tiles_y;
  const uint32_t t = blockIdx.x * kPlanTiles + threadIdx.x;
  uint32_t heavy = 0, light = 0;
  if (t < nt) {
    const uint32_t ns = segments_of(a.ranges[2 * t + 1] - a.ranges[2 * t], a.seg_len);
    heavy = ns > 1 ? ns : 0u;
    light = ns > 1 ? 0u : 1u;
  }
  uint32_t th, tl;
  block_exclusive_scan(heavy, s_wave[0], th);
  block_exclusive_scan(light, s_wave[1], tl);
  if (threadIdx.x == 0) {
    a.plan_counts[2 * blockIdx.x] = th;
    a.plan_counts[2 * blockIdx.x + 1] = tl;
  }
}

__global__ __launch_bounds__(kPlanTiles) void k_plan(CompositeArgs a) {
  if (*a.pair_count > a.capacity) return;  // truncated lists: the frame is drawn again
  __shared__ uint32_t s_wave[2][kPlanTiles / kWave];
  __shared__ uint32_t s_red[4][kPlanTiles / kWave];
  const uint32_t nt = a.grid.tiles_x * a.grid.tiles_y;
  const uint32_t nb = (nt + kPlanTiles - 1) / kPlanTiles;
  const uint32_t max_items = nt + a.capacity / a.seg_len + 1u;  // composite_max_items
  const uint32_t lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  // heavy / light items of the blocks before this one, and of all blocks
  uint32_t hb = 0, lb = 0, ha = 0, la = 0;
  for (uint32_t b = threadIdx.x; b < nb; b += kPlanTiles) {
    const uint32_t h = a.plan_counts[2 * b], l = a.plan_counts[2 * b + 1];
    ha += h;
    la += l;
    if (b < blockIdx.x) {
      hb += h;
      lb += l;
    }
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    hb += __shfl_xor(hb, off, kWave);
    lb += __shfl_xor(lb, off, kWave);
    ha += __shfl_xor(ha, off, kWave);
    la += __shfl_xor(la, off, kWave);
  }
  if (lane == 0) {
    s_red[0][w] = hb;
    s_red[1][w] = lb;
    s_red[2][w] = ha;
    s_red[3][w] = la;
  }
  __syncthreads();
  hb = lb = ha = la = 0;
  for (uint32_t q = 0; q < kPlanTiles / kWave; ++q) {
    hb += s_red[0][q];
    lb += s_red[1][q];
    ha += s_red[2][q];
    la += s_red[3][q];
  }
  const uint32_t t = blockIdx.x * kPlanTiles + threadIdx.x;
  uint32_t ns = 0, heavy = 0, light = 0;
  if (t < nt) {
    ns = segments_of(a.ranges[2 * t + 1] - a.ranges[2 * t], a.seg_len);
    heavy = ns > 1 ? ns : 0u;
    light = ns > 1 ? 0u : 1u;
  }
  uint32_t th, tl;
  const uint32_t hx = block_exclusive_scan(heavy, s_wave[0], th);
  const uint32_t lx = block_exclusive_scan(light, s_wave[1], tl);
  if (t < nt) {
    if (ns > 1) {
      uint32_t item = hb + hx;
      a.seg_base[t] = item;
      for (uint32_t sg = 0; sg < ns; ++sg, ++item) {
        if (item < max_items) {
          a.item_tile[item] = t;
          a.item_seg[item] = sg;
        }
      }
    } else {
      const uint32_t item = ha + lb + lx;
      if (item < max_items) {
        a.item_tile[item] = t;
        a.item_seg[item] = 0u;
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // all items, never more than the list holds (the bound is exact for lists
    // built from at most `capacity` pairs); multi-segment items
    const uint32_t all = ha + la;
    a.plan[0] = all < max_items ? all : max_items;
    a.plan[1] = ha;
    // the work ticket, in a cache line of its own
    a.plan[kPlanTicketWord] = 0u;
  }
}

template <bool DEPTH, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8))) void k_composite_fast(CompositeArgs a) {
#pragma clang fp contract(fast)
  if (*a.pair_count > a.capacity) return;  // truncated lists: the frame is drawn again
  __shared__ StageT<NT, DEPTH> s;
  __shared__ uint32_t s_item;
  const FrameConsts& fc = *a.fc;
  const int mode = fc.display_mode;
  const float sd2l = fc.max_std_dev * fc.max_std_dev * kLog2e;
  const float e = fc.max_std_dev - 0.1f;
  const float ell2l = e * e * kLog2e;
  const bool q8 = a.format == kFormatRgba8Unorm;
  const uint32_t nt = blockDim.x, tid = threadIdx.x, w = tid / kWave;
  // read once: loads from the ticket's cache line while every workgroup
  // takes tickets from it wait behind the atomics (measured: 3x slower)
  const uint32_t items = a.plan[0], heavy = a.plan[1];
  init_dummy(s, mode);

  for (;;) {
    const uint32_t item = take_ticket(a.plan + kPlanTicketWord, &s_item);
    if (item >= items) break;
    const uint32_t tile = a.item_tile[item], seg = a.item_seg[item];
    const Tiling t = setup_tile(a.grid, tile);
    const uint32_t tstart = a.ranges[2 * tile], tend = a.ranges[2 * tile + 1];
    // segment counted from the front (nearest end) of the tile's list
    const uint32_t front = tend - seg * a.seg_len;
    const uint32_t back = (item < heavy && front - tstart > a.seg_len) ? front - a.seg_len : tstart;
    const size_t o = size_t(t.py) * a.grid.width + t.px;
    DepthLane dl{a.depth_cmp, 0.0f, 0.0f, false};
    if (DEPTH && t.inside) dl.ref = a.depth[o];

    float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
    bool done = !t.inside;
    for (uint32_t bend = front; bend > back;) {
      const uint32_t cnt = (bend - back) < nt ? (bend - back) : nt;
      __syncthreads();
      const bool valid = tid < cnt;
      // nearest first
      stage_batch_quads(s, t, a.splats, DEPTH ? a.splat_depth : nullptr, valid,
                        valid ? a.pair_vals[bend - 1 - tid] : 0u, q8, mode, mode == kModeSplat);
      if (mode == kModeSplat)
        blend_quads<kModeSplat, DEPTH, NT>(s, t, w, sd2l, ell2l, T, cr, cg, cb, done, dl);
      else if (mode == kModeEllipse)
        blend_quads<kModeEllipse, DEPTH, NT>(s, t, w, sd2l, ell2l, T, cr, cg, cb, done, dl);
      else
        blend_quads<kModePoint, DEPTH, NT>(s, t, w, sd2l, ell2l, T, cr, cg, cb, done, dl);
      bend -= cnt;
      if (__syncthreads_and(done ? 1 : 0)) break;
    }
    if (t.inside) {
      if (item < heavy) {
        a.partial[size_t(item) * nt + tid] = make_float4(cr, cg, cb, T);
      } else {
        store_pixel(a, t, cr, cg, cb, T);
        if (DEPTH && a.depth_write && dl.found) a.depth[o] = dl.first;
      }
    }
  }
}

// Front-to-back merge of the segments of multi-segment tiles.
__global__ __launch_bounds__(kCompMaxThreads) void k_merge(CompositeArgs a) {
#pragma clang fp contract(fast)
  if (*a.pair_count > a.capacity) return;  // truncated lists: the frame is drawn again
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
  if (*a.pair_count > a.capacity) return;  // truncated lists: the frame is drawn again
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
  const uint32_t nt = blockDim.x, tid = threadIdx.x, w = tid / kWave, lane = tid % kWave;
  // the lane's 4x4 quadrant list (row-major over the tile's quadrants)
  const uint32_t ncols = 2u * t.blocks_x, bx = w % t.blocks_x, by = w / t.blocks_x;
  const uint32_t myq = (2u * by + (lane >> 5)) * ncols + 2u * bx + ((lane >> 2) & 1u);

  uint32_t frags = 0;
  float dr = 0.0f, dg = 0.0f, db = 0.0f, da = 1.0f;
  const bool dtest = a.depth != nullptr;
  float zcur = (dtest && t.inside) ? a.depth[size_t(t.py) * a.grid.width + t.px] : 0.0f;
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
    // farthest first
    stage_batch_quads(s, t, a.splats, dtest ? a.splat_depth : nullptr, valid, valid ? a.pair_vals[b0 + tid] : 0u,
                      q8, mode, false);
    if (!t.inside) continue;
    uint32_t len = 0;
    for (uint32_t p = 0; p < t.waves; ++p) len += uint32_t(__popcll(s.qbal[p][myq]));
    for (uint32_t j = 0; j < len; ++j) {
      const uint32_t q = uint32_t(s.list[myq][j]) >> 4;  // lists hold byte offsets
      float4 col;
      const float al = fragment(s, q, t, mode, sd2l, ell2l, col);
      if (al < 0.0f) continue;
      if (dtest) {
        // depth test against the attachment as updated by earlier draws
        const float z = s.z[q];
        if (!depth_pass(a.depth_cmp, z, zcur)) continue;
        if (a.depth_write) zcur = z;
      }
      ++frags;
      const float om = 1.0f - al;
      const float2 gb = s.gb[q];
      dr = col.w * al + dr * om;
      dg = gb.x * al + dg * om;
      db = gb.y * al + db * om;
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
    if (dtest && a.depth_write) a.depth[o] = zcur;
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

uint32_t composite_plan_blocks(uint32_t num_tiles) { return (num_tiles + kPlanTiles - 1) / kPlanTiles; }

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
    const uint32_t nb = composite_plan_blocks(tiles);
    hipLaunchKernelGGL(k_plan_count, dim3(nb), dim3(kPlanTiles), 0, st, a);
    hipLaunchKernelGGL(k_plan, dim3(nb), dim3(kPlanTiles), 0, st, a);
    // persistent workgroups take items by ticket; 8 per CU of 256 lanes
    uint32_t grid = uint32_t(num_cus) * 8u * (kCompMaxThreads / threads);
    const uint32_t cap = composite_max_items(tiles, a.capacity, a.seg_len);
    if (grid > cap) grid = cap;
    if (main_begin) (void)hipEventRecord(main_begin, st);
    // 8x8 tiles (one wave per workgroup) use a workgroup of 64 lanes
    if (threads == kWave) {
      if (a.depth)
        hipLaunchKernelGGL((k_composite_fast<true, kWave>), dim3(grid), dim3(threads), 0, st, a);
      else
        hipLaunchKernelGGL((k_composite_fast<false, kWave>), dim3(grid), dim3(threads), 0, st, a);
    } else {
      if (a.depth)
        hipLaunchKernelGGL((k_composite_fast<true, kCompMaxThreads>), dim3(grid), dim3(threads), 0, st, a);
      else
        hipLaunchKernelGGL((k_composite_fast<false, kCompMaxThreads>), dim3(grid), dim3(threads), 0, st, a);
    }
    if (main_end) (void)hipEventRecord(main_end, st);
    hipLaunchKernelGGL(k_merge, dim3(tiles), dim3(threads), 0, st, a);
  } else {
    if (main_begin) (void)hipEventRecord(main_begin, st);
    hipLaunchKernelGGL(k_composite_exact, dim3(tiles), dim3(threads), 0, st, a);
    if (main_end) (void)hipEventRecord(main_end, st);
  }
}

}  // namespace gs3d
