// Tile compositing: replaces the reference's fixed-function rasterizer and
// ALPHA_BLENDING over draw_indirect quads (src/renderer.rs:163-195, 284-308;
// fragment stage render.wesl:143-181).
//
// One workgroup per work item, one lane per pixel of a screen tile (8x8,
// 16x16 or 32x8 pixels); each wave owns an 8x8 block of the tile. The tile's
// splats (its run of the tile-sorted pairs, in draw order) are staged through
// LDS in batches of one splat per lane, and the batch is described per tile
// column and per tile row by a bitset of the slots whose pixel bound reaches
// that column / row (ballots). A pixel's splats are the bits of
// column & row, which its lane walks in slot order: a lane evaluates only
// the splats whose bound holds its pixel.
//
// exact : (the default) back to front; on an Rgba8Unorm target the
//         destination is rounded to 8 bits after every splat, as the blender
//         does (src/renderer.rs:296-300); rounding is not associative, so a
//         tile is one work item.
// fast  : front to back in f32 with transmittance, quantised once. The over
//         operator is associative, so a tile whose list is longer than one
//         segment is split into segments composited in parallel, each to a
//         partial (colour, transmittance) per pixel, and merged in order
//         afterwards. A lane stops once its transmittance is below 1e-4.
// Either way the tiles longer than one segment are scheduled first (a work
// list taken by ticket).
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

__device__ __forceinline__ float h_lo(uint32_t w) { return __half2float(__ushort_as_half((unsigned short)(w & 0xFFFFu))); }
__device__ __forceinline__ float h_hi(uint32_t w) { return __half2float(__ushort_as_half((unsigned short)(w >> 16))); }

// Two f16 values of one word as a vector: converting its elements where they
// feed an FMA lets the compiler use the mixed-precision FMA (f16 sources,
// f32 arithmetic) instead of separate conversions.
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ half2v as_h2(uint32_t w) { return __builtin_bit_cast(half2v, w); }

// f16 bits of clamp01 of an f16 value (exact: [0, 1] of an f16 is an f16).
__device__ __forceinline__ uint32_t clamp01_h2(uint32_t w) {
  const half2v h = as_h2(w);
  const _Float16 lo = h.x > _Float16(1.0f) ? _Float16(1.0f) : (h.x > _Float16(0.0f) ? h.x : _Float16(0.0f));
  const _Float16 hi = h.y > _Float16(1.0f) ? _Float16(1.0f) : (h.y > _Float16(0.0f) ? h.y : _Float16(0.0f));
  half2v r;
  r.x = lo;
  r.y = hi;
  return __builtin_bit_cast(uint32_t, r);
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

// Fragment alpha: 0 for a discarded fragment (blending alpha 0 leaves colour
// and transmittance unchanged), so the loop has no branches. `kept` tells
// whether the fragment survives the discard test. In splat mode the opacity
// is left out and so is the zero for a discarded fragment: the caller applies
// both (its select on `kept` already covers the discard).
template <int MODE>
__device__ __forceinline__ float fragment_alpha(float4 g, float4 c, const Tiling& t, float sd2l, float ell2l,
                                                bool& kept) {
#pragma clang fp contract(fast)
  const half2v bo = __builtin_bit_cast(half2v, c.w);  // b, opacity
  if (MODE == kModePoint) {
    const float dx = t.fpx - g.x;
    const float dy = t.fpy - g.y;
    kept = dx >= -g.z && dx < g.z && dy >= -g.w && dy < g.w;
    return kept ? float(bo.y) : 0.0f;
  }
  const float t1 = __builtin_fmaf(t.fpx, g.x, __builtin_fmaf(t.fpy, g.y, -c.x));
  const float t2 = __builtin_fmaf(t.fpx, g.z, __builtin_fmaf(t.fpy, g.w, -c.y));
  const float r2 = __builtin_fmaf(t1, t1, t2 * t2);
  float al;
  if (MODE == kModeEllipse) {
    const float op = float(bo.y);
    al = op + (1.0f - op) * ((r2 > ell2l) ? 1.0f : 0.0f);
  } else {
    al = __builtin_amdgcn_exp2f(-r2);
  }
  kept = r2 <= sd2l;
  return MODE == kModeSplat ? al : (kept ? al : 0.0f);
}

// Depth state of one lane of a depth-tested draw: the attachment's value
// before the draw and the depth of the nearest kept, passing fragment.
struct DepthLane {
  int cmp;
  float ref;
  float first;
  bool found;
};

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

// Work items of a tile: its segments (fast blend); one for the exact blend,
// which cannot split a tile. Either way a tile longer than one segment is
// "heavy" and scheduled first.
__device__ __forceinline__ void tile_items(const CompositeArgs& a, uint32_t len, uint32_t& heavy, uint32_t& light) {
  const uint32_t ns = segments_of(len, a.seg_len);
  const uint32_t n = a.blend != kBlendFast ? 1u : ns;
  heavy = ns > 1 ? n : 0u;
  light = ns > 1 ? 0u : 1u;
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
  if (t < nt) tile_items(a, a.ranges[2 * t + 1] - a.ranges[2 * t], heavy, light);
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
  uint32_t heavy = 0, light = 0;
  if (t < nt) tile_items(a, a.ranges[2 * t + 1] - a.ranges[2 * t], heavy, light);
  uint32_t th, tl;
  const uint32_t hx = block_exclusive_scan(heavy, s_wave[0], th);
  const uint32_t lx = block_exclusive_scan(light, s_wave[1], tl);
  if (t < nt) {
    if (heavy > 0) {
      uint32_t item = hb + hx;
      a.seg_base[t] = item;
      for (uint32_t sg = 0; sg < heavy; ++sg, ++item) {
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

// ---- per-pixel lists ------------------------------------------------------------
//
// A batch of NT staged splats (one per lane, in list order) is described per
// tile column and per tile row by a bitset of the slots whose pixel bound (the
// pixel centres within the splat's half extents) reaches that column / row:
// NT / 64 ballots per column and row. A pixel's splats are the bits of
// colm[x] & rowm[y], which its lane walks in slot order, 32-slot word after
// word (words with no bit for the pixel skipped), so a lane evaluates only the
// splats whose bound holds its pixel (a discarded fragment inside the bound
// costs a step; a splat beyond it costs nothing) and a wave steps as often as
// its busiest lane has splats. The exact blend keeps the destination as an
// integer 0..255 per channel on an Rgba8Unorm target:
// q' = round(q (1 - a) + 255 c a), the f32 blend followed by the unorm8
// rounding of every draw.

constexpr int kPxMaxCols = 32, kPxMaxRows = 16;

template <int NT, bool Z>
struct PxStage {
  static constexpr int kWords = NT / 32;
  // per slot: geo and col, 16 bytes each, in two arrays (a 16-byte stride
  // spreads the lanes' reads of random slots over more LDS banks than one
  // 32-byte record per slot)
  // Slot j of the batch is at index j + 1; index 0 holds a zero record
  // (alpha 0 in every display mode), written once per workgroup. (col
  // first: its 16 bytes are read as two 8-byte halves, which cannot carry the
  // other array's offset; the 16-byte geo read can.)
  float4 col[NT + 1];
  float4 geo[NT + 1];
  // (the bitsets before z: their rows start 16-byte aligned and are read as
  // 16-byte words)
  uint32_t colm[kPxMaxCols][kWords];
  uint32_t rowm[kPxMaxRows][kWords];
  float z[Z ? NT + 1 : 1];
};

// Transpose of a 32 x 32 bit matrix held one row per lane, in each half of
// the wave: afterwards lane j of a half holds bit j of that half's 32 input
// words (bit i from lane i). Five butterfly steps: at step k, the lanes i and
// i ^ k swap the k-bit blocks that sit off the diagonal of their 2k x 2k
// block. Every lane must be active.
// v_writelane_b32: lane `lane` (wave-uniform) of `old` becomes the
// wave-uniform `val`. Declared through the LLVM intrinsic's name, for which
// this compiler has no builtin; the compiler schedules it like any other
// instruction (no inline assembly).
extern "C" __device__ int gs3d_writelane(int val, int lane, int old) __asm("llvm.amdgcn.writelane.i32");

template <int ST>
__device__ __forceinline__ uint32_t transpose_step(uint32_t x, uint32_t lane) {
  constexpr uint32_t k = 16u >> ST;
  constexpr uint32_t m = ST == 0 ? 0x0000FFFFu : ST == 1 ? 0x00FF00FFu : ST == 2 ? 0x0F0F0F0Fu : ST == 3 ? 0x33333333u
                                                                                                   : 0x55555555u;
  // the word of lane i ^ k (ds_swizzle, bit mode: and 0x1f, xor k)
  const uint32_t p = uint32_t(__builtin_amdgcn_ds_swizzle(int(x), int(0x1Fu | (k << 10))));
  // all ones in the upper lane of the pair (bit k of the lane set)
  const uint32_t up = uint32_t(__builtin_amdgcn_sbfe(int(lane), 4 - ST, 1));
  // the partner's off-diagonal blocks, moved over this lane's: upper lanes
  // take p >> k into their low blocks, lower lanes p << k into their high ones
  const uint32_t sh = ((p >> k) & up) | ((p << k) & ~up);
  const uint32_t keep = m ^ up;  // lower lanes keep the low blocks, upper the high ones
  return (x & keep) | (sh & ~keep);
}

__device__ __forceinline__ uint32_t transpose_half_waves(uint32_t x) {
  uint32_t lane = threadIdx.x % kWave;
  // (the lane id made opaque here: the steps' per-lane masks are then
  // recomputed at each use instead of held in registers across the kernel)
  asm volatile("" : "+v"(lane));
  x = transpose_step<0>(x, lane);
  x = transpose_step<1>(x, lane);
  x = transpose_step<2>(x, lane);
  x = transpose_step<3>(x, lane);
  return transpose_step<4>(x, lane);
}

// Index of the lowest set bit of x, all ones for x = 0: v_ffbl_b32 alone (the
// compiler's count of trailing zeros is undefined at 0, and its defined form
// adds a select after the shift that follows it here). An ALU instruction; no
// memory access.
__device__ __forceinline__ uint32_t lowest_bit_or_ones(uint32_t x) {
  uint32_t b;
  asm("v_ffbl_b32 %0, %1" : "=v"(b) : "v"(x));
  return b;
}

// A lane's splat for a batch, requested before the batch before it is
// blended (the gather's latency then overlaps that blend).
template <bool Z>
struct PxFetch {
  Splat2D r;
  float z;
  bool valid;
};

// k is a valid splat index even for a lane without a splat in the batch (the
// load is then unused): an unconditional load leaves the compiler nothing to
// merge, so it does not wait for the load until the staging uses it. A pair
// value at or above num (a sort that gave up waiting leaves such entries) is
// reported and its splat neither read nor drawn.
template <bool Z>
__device__ __forceinline__ void fetch_px(PxFetch<Z>& f, const Splat2D* splats, const float* zsrc, bool valid,
                                         uint32_t k, uint32_t num, uint32_t* err) {
  const bool ok = k < num;
  if (valid && !ok) atomicOr(err, kErrSplatIndex);
  f.valid = valid && ok;
  const uint32_t kk = ok ? k : 0u;
  f.r = splats[kk];
  if (Z && zsrc) f.z = zsrc[kk];
}

// Stages the fetched splat (or nothing) into slot threadIdx.x and builds the
// column and row bitsets of the batch. Contains __syncthreads; every lane must
// call it.
template <int NT, bool Z>
__device__ __forceinline__ void stage_batch_px(PxStage<NT, Z>& s, const Tiling& t, const TileGrid& g,
                                               const PxFetch<Z>& f, bool has_z, bool clamp_src, int mode) {
  const uint32_t tid = threadIdx.x, lane = tid % kWave, sw = tid / kWave;
  int xa = 1, xb = 0, ya = 1, yb = 0;  // empty
  if (f.valid) {
    const Splat2D& r = f.r;
    const uint32_t rg = clamp_src ? clamp01_h2(r.rg) : r.rg;
    const uint32_t ba = clamp_src ? ((clamp01_h2(r.ba) & 0xFFFFu) | (r.ba & 0xFFFF0000u)) : r.ba;
    const float cx = r.cx - t.sx, cy = r.cy - t.sy;
    if (mode == kModePoint) {
      s.geo[tid + 1] = make_float4(cx, cy, r.ux, r.uy);
      s.col[tid + 1] = make_float4(0.0f, 0.0f, __uint_as_float(rg), __uint_as_float(ba));
    } else {
      const float ux = r.ux * kSqrtLog2e, uy = r.uy * kSqrtLog2e;
      const float vx = -r.rho * uy, vy = r.rho * ux;
      s.geo[tid + 1] = make_float4(ux, uy, vx, vy);
      s.col[tid + 1] = make_float4(cx * ux + cy * uy, cx * vx + cy * vy, __uint_as_float(rg), __uint_as_float(ba));
    }
    if (Z && has_z) s.z[tid + 1] = f.z;
    if (r.hxy != 0u) {
      // pixel x of the tile holds the centre ox + x + 0.5: inside the bound
      // [cx - hx, cx + hx] for x in [ceil(cx - hx - ox - 0.5), floor(cx + hx - ox - 0.5)]
      const float hx = h_lo(r.hxy), hy = h_hi(r.hxy);
      const float fx = r.cx - t.ox - 0.5f, fy = r.cy - t.oy - 0.5f;
      xa = max(0, int(ceilf(fx - hx)));
      xb = min(int(g.tile_w) - 1, int(floorf(fx + hx)));
      ya = max(0, int(ceilf(fy - hy)));
      yb = min(int(g.tile_h) - 1, int(floorf(fy + hy)));
    }
  }
  // this lane's column and row masks (bit c: its pixel bound reaches column
  // c). The bitsets are these masks transposed (bit l of column c's word 2 sw
  // + h is bit c of lane 32 h + l's mask): a butterfly transpose within each
  // half-wave leaves lane 32 h + j holding column (or row) j's word 2 sw + h.
  const bool xe = xa <= xb, ye = ya <= yb;  // then 0 <= a <= b < 32
  const uint32_t cm = xe ? (2u << uint32_t(xb)) - (1u << uint32_t(xa)) : 0u;
  const uint32_t rm = ye ? (2u << uint32_t(yb)) - (1u << uint32_t(ya)) : 0u;
  const uint32_t j = lane & 31u, h = lane >> 5;
  if (g.tile_w <= 16u) {
    // columns in bits 0-15, rows in bits 16-31 of one word
    const uint32_t wd = transpose_half_waves(cm | (rm << 16));
    if (j < 16u) {
      if (j < g.tile_w) s.colm[j][2 * sw + h] = wd;
    } else if (j - 16u < g.tile_h) {
      s.rowm[j - 16u][2 * sw + h] = wd;
    }
  } else {
    // wider tiles: one ballot per column and row, gathered into lanes with
    // v_writelane (column c into lane c, row r into lane 32 + r: at most 32
    // columns and 8 rows here) and stored together
    int w0 = 0, w1 = 0;
    for (int c = 0; c < int(g.tile_w); ++c) {
      const uint64_t bal = __builtin_amdgcn_ballot_w64(((cm >> c) & 1u) != 0u);
      w0 = gs3d_writelane(int(uint32_t(bal)), c, w0);
      w1 = gs3d_writelane(int(uint32_t(bal >> 32)), c, w1);
    }
    int v0 = 0, v1 = 0;
    for (int r = 0; r < int(g.tile_h); ++r) {
      const uint64_t bal = __builtin_amdgcn_ballot_w64(((rm >> r) & 1u) != 0u);
      v0 = gs3d_writelane(int(uint32_t(bal)), r, v0);
      v1 = gs3d_writelane(int(uint32_t(bal >> 32)), r, v1);
    }
    if (lane < g.tile_w) {
      s.colm[lane][2 * sw] = uint32_t(w0);
      s.colm[lane][2 * sw + 1] = uint32_t(w1);
    }
    if (lane < g.tile_h) {
      s.rowm[lane][2 * sw] = uint32_t(v0);
      s.rowm[lane][2 * sw + 1] = uint32_t(v1);
    }
  }
  __syncthreads();
}

// Blend of one batch for this lane's pixel (tile-relative lx, ly), slots in
// increasing order. EXACT: back to front (q8: the 8-bit per-draw rounding);
// otherwise front to back with transmittance T, a lane stopping once below
// the threshold (checked every 16 steps).
// track_a: the destination's alpha may be below 1 (a loaded target); after a
// clear it stays 1 under the over operator. count: count the kept fragments.
// PLAIN: q8 is true and track_a and count are false (the common draw: an
// Rgba8Unorm target after a clear), fixed at compile time so that the
// per-splat step carries no test of them.
template <int MODE, bool EXACT, bool DEPTH, int NT, bool PLAIN>
__device__ __forceinline__ void blend_px(const PxStage<NT, DEPTH>& s, const Tiling& t, uint32_t lx, uint32_t ly,
                                         bool q8_in, bool track_a_in, bool count_in, float sd2l, float ell2l,
                                         float& T, float& cr, float& cg, float& cb, float& ca, bool& done,
                                         DepthLane& dl, uint32_t& frags) {
#pragma clang fp contract(fast)
  constexpr int kWords = PxStage<NT, DEPTH>::kWords;
  const bool q8 = PLAIN || q8_in;
  const bool track_a = !PLAIN && track_a_in;
  const bool count = !PLAIN && count_in;
  // the words (32 slots each) that hold splats for this pixel
  uint32_t nz = 0;
  if (t.inside && (EXACT || !done)) {
#pragma unroll
    for (int wi = 0; wi < kWords; wi += (kWords % 4 == 0 ? 4 : 1)) {
      if constexpr (kWords % 4 == 0) {
        // four words per 16-byte read of each row
        const uint4 c = *reinterpret_cast<const uint4*>(&s.colm[lx][wi]);
        const uint4 r = *reinterpret_cast<const uint4*>(&s.rowm[ly][wi]);
        nz |= (min(c.x & r.x, 1u) | (min(c.y & r.y, 1u) << 1) | (min(c.z & r.z, 1u) << 2) | (min(c.w & r.w, 1u) << 3))
              << wi;
      } else {
        nz |= min(s.colm[lx][wi] & s.rowm[ly][wi], 1u) << wi;
      }
    }
  }
  // each lane walks its own bits in slot order, word after word: the wave
  // steps as often as its busiest lane has splats. The next step's slot is
  // read from LDS before the current one is blended, so that the read's
  // latency overlaps the blend (unrolled by two: no copies between them).
  // base: byte offset of index (first slot of the lane's word) + 1
  uint32_t cur = 0, base = 16u, steps = 0;
  // the lane's next slot: the byte offset of its index, and whether the lane
  // has one. The lowest set bit is -1 for a lane with no bit left: such a
  // lane reads index 32 wi, the slot just below its last word's first
  // (staged: slots fill from 0, so its values are finite and 0 * finite
  // leaves the colour unchanged), or the zero record; its alpha is masked by
  // act. A slot past the batch's end holds stale data (an Inf or NaN there
  // would reach the colour through the zero-alpha blend), and no lane reads
  // one.
  auto advance = [&](uint32_t& q, bool& act) {
    if (cur == 0u && nz != 0u) {
      const uint32_t wi = uint32_t(__builtin_ctz(nz));
      nz &= nz - 1u;
      cur = s.colm[lx][wi] & s.rowm[ly][wi];
      base = (wi * 32u + 1u) * 16u;
    }
    act = cur != 0u;
    const uint32_t b = lowest_bit_or_ones(cur);
    cur &= cur - 1u;
    q = (b << 4) + base;
  };
  auto geo_at = [&](uint32_t q) { return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(s.geo) + q); };
  auto col_at = [&](uint32_t q) { return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(s.col) + q); };
  // one step: the splat in (g, c, zz) for this lane's pixel; true when the
  // wave stops (fast blend: every lane done)
  auto step = [&](const float4& g, const float4& c, bool act, float zz) -> bool {
    bool kept;
    float al = fragment_alpha<MODE>(g, c, t, sd2l, ell2l, kept);
    kept = kept && act;
    const half2v rg = __builtin_bit_cast(half2v, c.z), bo = __builtin_bit_cast(half2v, c.w);
    if (MODE == kModeSplat) al = kept ? __builtin_fmaf(al, float(bo.y), 0.0f) : 0.0f;  // times the opacity
    else al = kept ? al : 0.0f;
    if (DEPTH) {
      if (EXACT) {
        // against the attachment as updated by earlier draws
        const bool pass = kept && depth_pass(dl.cmp, zz, dl.ref);
        al = pass ? al : 0.0f;
        if (pass && dl.found) dl.ref = zz;  // found: depth_write
        kept = pass;
      } else {
        const bool pass = depth_pass(dl.cmp, zz, dl.ref);
        al = pass ? al : 0.0f;
        if (kept && pass && !dl.found) {
          dl.first = zz;
          dl.found = true;
        }
      }
    }
    if (EXACT) {
      if (count) frags += kept ? 1u : 0u;
      const float om = 1.0f - al;
      if (q8) {
        // q' = round(q om + 255 c al), q in 0..255
        const float a255 = 255.0f * al;
        cr = __builtin_rintf(__builtin_fmaf(float(rg.x), a255, cr * om));
        cg = __builtin_rintf(__builtin_fmaf(float(rg.y), a255, cg * om));
        cb = __builtin_rintf(__builtin_fmaf(float(bo.x), a255, cb * om));
        if (track_a) ca = __builtin_rintf(__builtin_fmaf(ca, om, a255));
      } else {
        cr = __builtin_fmaf(float(rg.x), al, cr * om);
        cg = __builtin_fmaf(float(rg.y), al, cg * om);
        cb = __builtin_fmaf(float(bo.x), al, cb * om);
        if (track_a) ca = __builtin_fmaf(ca, om, al);
      }
    } else {
      const float wgt = T * al;
      cr = __builtin_fmaf(wgt, float(rg.x), cr);
      cg = __builtin_fmaf(wgt, float(rg.y), cg);
      cb = __builtin_fmaf(wgt, float(bo.x), cb);
      T = T - wgt;
      if ((++steps & 15u) == 0u) {
        // the wave stops when every lane is done
        if (__all(done || T < kMinTransmittance)) {
          done = true;
          return true;
        }
      }
    }
    return false;
  };
  // two slots in flight (a, b): each half of the loop reads one slot while
  // it blends the other
  uint32_t qa, qb;
  bool acta, actb;
  advance(qa, acta);
  float4 ga = geo_at(qa), ca_ = col_at(qa);
  float zza = DEPTH ? s.z[qa >> 4] : 0.0f;
  for (;;) {
    if (!__any(acta)) break;
    advance(qb, actb);
    const float4 gb = geo_at(qb), cb_ = col_at(qb);
    const float zzb = DEPTH ? s.z[qb >> 4] : 0.0f;
    if (step(ga, ca_, acta, zza)) return;
    if (!__any(actb)) break;
    advance(qa, acta);
    ga = geo_at(qa);
    ca_ = col_at(qa);
    zza = DEPTH ? s.z[qa >> 4] : 0.0f;
    if (step(gb, cb_, actb, zzb)) return;
  }
  if (!EXACT) done = done || T < kMinTransmittance;
}

// 8 waves per SIMD; 7 for depth-tested draws (their depth state and prefetched
// fragment depths need the registers)
template <bool EXACT, bool DEPTH, int NT, bool PLAIN>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(DEPTH ? 7 : 8))) void k_composite_px(
    CompositeArgs a) {
  if (*a.pair_count > a.capacity) return;  // truncated lists: the frame is drawn again
  __shared__ PxStage<NT, DEPTH> s;
  __shared__ uint32_t s_item;
  if (threadIdx.x == 0) {  // the zero record (the first ticket's barrier publishes it)
    s.col[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    s.geo[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (DEPTH) s.z[0] = 0.0f;
  }
  const FrameConsts& fc = *a.fc;
  const int mode = fc.display_mode;
  const float sd2l = fc.max_std_dev * fc.max_std_dev * kLog2e;
  const float e = fc.max_std_dev - 0.1f;
  const float ell2l = e * e * kLog2e;
  const bool q8 = PLAIN || a.format == kFormatRgba8Unorm;
  const uint32_t nt = blockDim.x, tid = threadIdx.x;
  // read once: loads from the ticket's cache line while every workgroup
  // takes tickets from it wait behind the atomics
  const uint32_t items = a.plan[0], heavy = a.plan[1];

  for (;;) {
    const uint32_t item = take_ticket(a.plan + kPlanTicketWord, &s_item);
    if (item >= items) break;
    // in the work list's order, heavy tiles first: taken in raster order, the
    // exact blend's densest tiles (one workgroup each) start late and run on
    // alone after the rest
    const uint32_t tile = a.item_tile[item];
    const uint32_t seg = EXACT ? 0u : a.item_seg[item];
    const Tiling t = setup_tile(a.grid, tile);
    const uint32_t lx = t.px - uint32_t(t.ox), ly = t.py - uint32_t(t.oy);
    const uint32_t tstart = a.ranges[2 * tile], tend = a.ranges[2 * tile + 1];
    if (EXACT) {
      // the longest lists set the kernel's length (one workgroup each): their
      // waves issue first on a SIMD they share
      const uint32_t tl = __builtin_amdgcn_readfirstlane(tend - tstart);
      if (tl > 8192u)
        __builtin_amdgcn_s_setprio(3);
      else if (tl > 4096u)
        __builtin_amdgcn_s_setprio(2);
      else if (tl > 2048u)
        __builtin_amdgcn_s_setprio(1);
      else
        __builtin_amdgcn_s_setprio(0);
    }
    uint32_t front = tend, back = tstart;
    if (!EXACT) {
      // segment counted from the front (nearest end) of the tile's list
      front = tend - seg * a.seg_len;
      back = (item < heavy && front - tstart > a.seg_len) ? front - a.seg_len : tstart;
    }
    const size_t o = size_t(t.py) * a.grid.width + t.px;
    DepthLane dl{a.depth_cmp, 0.0f, 0.0f, false};
    if (DEPTH && t.inside) dl.ref = a.depth[o];
    if (EXACT) dl.found = a.depth_write != 0;  // exact: `found` marks depth writes

    // fast: colour and transmittance; exact: the destination (q8: 0..255)
    float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f, ca = q8 ? 255.0f : 1.0f;
    if (EXACT && !PLAIN && a.load && t.inside) {
      if (q8) {
        const uint32_t d = reinterpret_cast<const uint32_t*>(a.out)[o];
        cr = float(d & 0xFFu);
        cg = float((d >> 8) & 0xFFu);
        cb = float((d >> 16) & 0xFFu);
        ca = float(d >> 24);
      } else {
        const float4 d = reinterpret_cast<const float4*>(a.out)[o];
        cr = d.x;
        cg = d.y;
        cb = d.z;
        ca = d.w;
      }
    }
    bool done = !t.inside;
    uint32_t frags = 0;
    // after a clear the destination's alpha stays 1 (1 a + 1 (1 - a))
    const bool track_a = !PLAIN && a.load != 0;
    const bool count = !PLAIN && a.frag_count != nullptr;
    // batches of nt pairs; exact: farthest first (the list's order), fast:
    // nearest first. Lane tid stages pair `slot(p0)` of the batch at p0; its
    // splat is requested one batch ahead and its pair value two batches ahead.
    const uint32_t len = front - back;
    const float* zsrc = DEPTH ? a.splat_depth : nullptr;
    auto slot = [&](uint32_t p0) { return EXACT ? back + p0 + tid : front - 1u - p0 - tid; };
    auto in_batch = [&](uint32_t p0) { return p0 < len && tid < len - p0; };
    // (a lane past the end of a non-empty list reads pair `back`, written by
    // this frame's binning, so that every load is in bounds and unconditional;
    // an empty list loads nothing)
    PxFetch<DEPTH> f;
    bool v_next = false;
    uint32_t k_next = 0u;
    if (len > 0u) {
      const bool v0 = in_batch(0u);
      fetch_px(f, a.splats, zsrc, v0, a.pair_vals[v0 ? slot(0u) : back], a.num_gaussians, a.err);
      v_next = in_batch(nt);
      k_next = a.pair_vals[v_next ? slot(nt) : back];
    }
    for (uint32_t p0 = 0; p0 < len;) {
      const uint32_t cnt = (len - p0) < nt ? (len - p0) : nt;
      __syncthreads();
      stage_batch_px(s, t, a.grid, f, zsrc != nullptr, q8, mode);
      // the next batch's splat, the one after's pair value
      fetch_px(f, a.splats, zsrc, v_next, k_next, a.num_gaussians, a.err);
      v_next = in_batch(p0 + 2u * nt);
      k_next = a.pair_vals[v_next ? slot(p0 + 2u * nt) : back];
      if (mode == kModeSplat)
        blend_px<kModeSplat, EXACT, DEPTH, NT, PLAIN>(s, t, lx, ly, q8, track_a, count, sd2l, ell2l, T, cr, cg, cb, ca,
                                                       done, dl, frags);
      else if (mode == kModeEllipse)
        blend_px<kModeEllipse, EXACT, DEPTH, NT, PLAIN>(s, t, lx, ly, q8, track_a, count, sd2l, ell2l, T, cr, cg, cb,
                                                         ca, done, dl, frags);
      else
        blend_px<kModePoint, EXACT, DEPTH, NT, PLAIN>(s, t, lx, ly, q8, track_a, count, sd2l, ell2l, T, cr, cg, cb, ca,
                                                       done, dl, frags);
      p0 += cnt;
      if (!EXACT && __syncthreads_and(done ? 1 : 0)) break;
    }
    if (t.inside) {
      if (EXACT) {
        if (q8)
          reinterpret_cast<uint32_t*>(a.out)[o] =
              uint32_t(cr) | (uint32_t(cg) << 8) | (uint32_t(cb) << 16) | (uint32_t(ca) << 24);
        else
          reinterpret_cast<float4*>(a.out)[o] = make_float4(cr, cg, cb, ca);
        if (DEPTH && a.depth_write) a.depth[o] = dl.ref;
        if (a.frag_count && frags) atomicAdd(a.frag_count, frags);
      } else if (item < heavy) {  // a segment of a split tile
        a.partial[size_t(item) * nt + tid] = make_float4(cr, cg, cb, T);
      } else {
        store_pixel(a, t, cr, cg, cb, T);
        if (DEPTH && a.depth_write && dl.found) a.depth[o] = dl.first;
      }
    }
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

template <bool EXACT>
static void launch_px(hipStream_t st, const CompositeArgs& a, uint32_t grid, uint32_t threads) {
  // the exact blend of an Rgba8Unorm target after a clear, not counting
  // fragments, has its own instantiation (no per-splat test of those flags)
  const bool plain = EXACT && !a.depth && a.format == kFormatRgba8Unorm && !a.load && a.frag_count == nullptr;
  if (threads == kWave) {
    if (a.depth)
      hipLaunchKernelGGL((k_composite_px<EXACT, true, kWave, false>), dim3(grid), dim3(threads), 0, st, a);
    else if (plain)
      hipLaunchKernelGGL((k_composite_px<EXACT, false, kWave, EXACT>), dim3(grid), dim3(threads), 0, st, a);
    else
      hipLaunchKernelGGL((k_composite_px<EXACT, false, kWave, false>), dim3(grid), dim3(threads), 0, st, a);
  } else {
    if (a.depth)
      hipLaunchKernelGGL((k_composite_px<EXACT, true, kCompMaxThreads, false>), dim3(grid), dim3(threads), 0, st, a);
    else if (plain)
      hipLaunchKernelGGL((k_composite_px<EXACT, false, kCompMaxThreads, EXACT>), dim3(grid), dim3(threads), 0, st, a);
    else
      hipLaunchKernelGGL((k_composite_px<EXACT, false, kCompMaxThreads, false>), dim3(grid), dim3(threads), 0, st, a);
  }
}

void launch_composite(hipStream_t st, const CompositeArgs& a, int num_cus, hipEvent_t main_begin,
                      hipEvent_t main_end) {
  const uint32_t tiles = a.grid.tiles_x * a.grid.tiles_y;
  const uint32_t threads = a.grid.tile_w * a.grid.tile_h;
  if (tiles == 0) return;
  // work list (heavy tiles first), then persistent workgroups taking items
  // by ticket, 8 per CU of 256 lanes
  const uint32_t nb = composite_plan_blocks(tiles);
  hipLaunchKernelGGL(k_plan_count, dim3(nb), dim3(kPlanTiles), 0, st, a);
  hipLaunchKernelGGL(k_plan, dim3(nb), dim3(kPlanTiles), 0, st, a);
  uint32_t grid = uint32_t(num_cus) * 8u * (kCompMaxThreads / threads);
  const uint32_t cap = composite_max_items(tiles, a.capacity, a.seg_len);
  if (grid > cap) grid = cap;
  if (main_begin) (void)hipEventRecord(main_begin, st);
  if (a.blend == kBlendFast)
    launch_px<false>(st, a, grid, threads);
  else
    launch_px<true>(st, a, grid, threads);
  if (main_end) (void)hipEventRecord(main_end, st);
  if (a.blend == kBlendFast) hipLaunchKernelGGL(k_merge, dim3(tiles), dim3(threads), 0, st, a);
}

}  // namespace gs3d
