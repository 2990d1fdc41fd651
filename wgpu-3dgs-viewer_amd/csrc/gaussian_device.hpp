// Device-side access to the packed Gaussian streams and the per-splat
// vertex-stage computation (render.wesl:46-130) shared by the fused
// projection kernel and the renderer's standalone vertex kernel.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "math.hpp"
#include "types.hpp"

namespace gs3d {

// HBM layout of a GaussiansBuffer (structure of streams, one record per
// Gaussian in each stream):
//   pc   : float4  pos.xyz, color rgba8 bits in .w            16 B
//   kPodSingle: cov  : float2[3]  cov3d xx,xy,xz,yy,yz,zz       24 B
//               sh   : 12 planes of float4: 45 SH floats + 3 pad 192 B
//   kPodHalf:   rs   : uint4  rot xyzw f16, scale xyz f16, pad   16 B
//               sh   : 6 planes of uint4: 45 SH halves + 3 pad   96 B
// The SH stream is plane-major: 16-byte chunk q of Gaussian i is element
// q * n + i, so one load instruction of a wave reads one contiguous 1 KiB
// run of a plane, and lower SH degrees touch only their planes. The culled
// majority only reads `pc` and the covariance stream; the SH stream is read
// for visible splats only.
struct GaussianStreams {
  const float4* pc;
  const float2* cov;  // kPodSingle
  const uint4* rs;    // kPodHalf
  const float4* sh32; // kPodSingle
  const uint4* sh16;  // kPodHalf
  uint32_t n;
};

__device__ __forceinline__ float half_lo(uint32_t w) { return __half2float(__ushort_as_half((unsigned short)(w & 0xFFFFu))); }
__device__ __forceinline__ float half_hi(uint32_t w) { return __half2float(__ushort_as_half((unsigned short)(w >> 16))); }

__device__ __forceinline__ uint32_t pack_half2(float lo, float hi) {
  return uint32_t(__half_as_ushort(__float2half(lo))) | (uint32_t(__half_as_ushort(__float2half(hi))) << 16);
}

// f16 bits of a non-negative extent rounded towards +inf (a conservative
// bound); NaN or values past the f16 range saturate at 65504.
__device__ __forceinline__ uint32_t half_up_bits(float x) {
  if (!(x >= 0.0f)) x = 65504.0f;
  x = fminf(x, 65504.0f);
  uint32_t h = __half_as_ushort(__float2half(x));
  if (__half2float(__ushort_as_half((unsigned short)h)) < x) ++h;
  if (h == 0u) h = 1u;  // smallest subnormal: a drawable splat never encodes as 0
  return h;
}

__device__ __forceinline__ uint32_t pack_half2_up(float lo, float hi) {
  return half_up_bits(lo) | (half_up_bits(hi) << 16);
}

// The covariance stream's raw words for one Gaussian (so that loads can be
// issued before they are needed).
struct CovWords {
  float2 a, b, c;  // kPodSingle
  uint4 rs;        // kPodHalf
};

template <int POD>
__device__ __forceinline__ CovWords load_cov_words(const GaussianStreams& g, uint32_t i) {
  CovWords w{};
  if constexpr (POD == kPodSingle) {
    const float2* p = g.cov + size_t(i) * 3;
    w.a = p[0];
    w.b = p[1];
    w.c = p[2];
  } else {
    w.rs = g.rs[i];
  }
  return w;
}

// cov3d_from_scale_rot contracted (the record's covariance; the bound
// re-test of the cull keeps the uncontracted one).
__device__ __forceinline__ void cov3d_fast(const float s[3], const float q[4], float out[6]) {
#pragma clang fp contract(fast)
  float rm[9], mm[9];
  quat_to_mat3(q, rm);
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) mm[c * 3 + r] = rm[c * 3 + r] * s[c];
  const int ia[6] = {0, 0, 0, 1, 1, 2};
  const int ib[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int e = 0; e < 6; ++e)
    out[e] = mm[0 * 3 + ia[e]] * mm[0 * 3 + ib[e]] + mm[1 * 3 + ia[e]] * mm[1 * 3 + ib[e]] +
             mm[2 * 3 + ia[e]] * mm[2 * 3 + ib[e]];
}

// The cov3d of one Gaussian from its covariance words; kFast contracts the
// half form's product (for the splat record only).
template <int POD, bool kFast = false>
__device__ __forceinline__ void cov3d_of(const CovWords& w, float c[6]) {
  if constexpr (POD == kPodSingle) {
    c[0] = w.a.x;
    c[1] = w.a.y;
    c[2] = w.b.x;
    c[3] = w.b.y;
    c[4] = w.c.x;
    c[5] = w.c.y;
  } else {
    const float q[4] = {half_lo(w.rs.x), half_hi(w.rs.x), half_lo(w.rs.y), half_hi(w.rs.y)};
    const float s[3] = {half_lo(w.rs.z), half_hi(w.rs.z), half_lo(w.rs.w)};
    if constexpr (kFast)
      cov3d_fast(s, q, c);
    else
      cov3d_from_scale_rot(s, q, c);
  }
}

// The ellipse |qo|^2 = (d.u)^2 + (d.v)^2 <= lim of a splat's offsets d from
// its centre, for u = (ux, uy) and v = rho (-uy, ux) (a Splat2D's axes), is
// the quadratic form A dx^2 + 2 B dx dy + C dy^2 <= lim with A = |(ux, vx)|^2,
// B = ux uy + vx vy, C = |(uy, vy)|^2 and determinant D^2, D = rho |u|^2
// (formed from the axes, not as A C - B^2, which cancels for thin splats).
// Its extents are |dx| <= xe = sqrt(lim C) / D and |dy| <= ye = sqrt(lim A) / D;
// the rightmost point lies at dy = yr = -(B / C) xe, the leftmost at -yr.
// Within a horizontal band by0 <= dy <= by1 the ellipse is convex, so it meets
// a tile's box of pixel-centre offsets [bx0, bx1] x [by0, by1] exactly when
// [bx0, bx1] meets the band's x-interval [xl, xr]: xr is the right root
// x+(dy) = (-B dy + sqrt(A lim - D^2 dy^2)) / A, concave in dy, at dy = yr
// clamped to the band, and xl likewise at -yr. One row of tiles thus costs
// two square roots instead of one test per tile. The interval is widened by
// 0.4% of the extents plus 1e-3 pixel, above the rounding of this evaluation
// (the discriminant's cancellation near a tip moves a root by at most ~1e-3
// of sqrt(lim / A) <= xe) and of the blend's, so a fragment the blend keeps is
// never dropped.
struct EllipseRows {
  float A, B, iA, alim, det2;  // det2 = D^2
  float yr, ye, tol;
  bool ok;
};

__device__ __forceinline__ EllipseRows ellipse_rows(float ux, float uy, float rho, float lim) {
  EllipseRows e;
  const float vx = -rho * uy, vy = rho * ux;
  e.A = ux * ux + vx * vx;
  e.B = ux * uy + vx * vy;
  const float C = uy * uy + vy * vy;
  const float D = rho * (ux * ux + uy * uy);
  e.ok = e.A > 0.0f && C > 0.0f && D > 0.0f && e.A < __builtin_inff() && C < __builtin_inff() &&
         D < __builtin_inff();
  const float iD = __builtin_amdgcn_rcpf(D);
  e.iA = __builtin_amdgcn_rcpf(e.A);
  const float xe = __builtin_amdgcn_sqrtf(lim * C) * iD;
  e.ye = __builtin_amdgcn_sqrtf(lim * e.A) * iD;
  e.yr = -(e.B * __builtin_amdgcn_rcpf(C)) * xe;
  e.alim = e.A * lim;
  e.det2 = D * D;
  e.tol = 4e-3f * (xe + e.ye) + 1e-3f;
  return e;
}

// The band's x-interval [xl, xr] (widened by tol); false when the band
// misses the ellipse's rows.
__device__ __forceinline__ bool ellipse_row_span(const EllipseRows& e, float by0, float by1, float& xl, float& xr) {
  if (by0 > e.ye + e.tol || by1 < -e.ye - e.tol) return false;
  const float yR = fminf(fmaxf(e.yr, by0), by1), yL = fminf(fmaxf(-e.yr, by0), by1);
  const float dR = __builtin_amdgcn_sqrtf(fmaxf(e.alim - e.det2 * (yR * yR), 0.0f));
  const float dL = __builtin_amdgcn_sqrtf(fmaxf(e.alim - e.det2 * (yL * yL), 0.0f));
  xr = (dR - e.B * yR) * e.iA + e.tol;
  xl = (-dL - e.B * yL) * e.iA - e.tol;
  return true;
}

// The mask bits of every tile of a w x h rectangle (w, h <= 4).
__device__ __forceinline__ uint32_t rect_bits(uint32_t w, uint32_t h) {
  return ((1u << w) - 1u) * (0x1111u & ((1u << (4u * h)) - 1u));
}

// Inclusive tile rectangle a splat can touch and, for rectangles of at most
// 4 x 4 tiles, which of those tiles its ellipse |qo|^2 <= sd2 reaches:
//   x = tx0 | ty0 << 12 | mask bits 0-7 << 24
//   y = tx1 | ty1 << 12 | mask bits 8-15 << 24
// with bit 4 j + i of the mask for tile (tx0 + i, ty0 + j); point-mode
// splats (rho = 0) set the bits of every tile of the rectangle. Larger
// rectangles keep every tile (mask unused). (0xFFFFFFFF, 0) when it touches
// none. Tile coordinates take 12 bits: textures are at most 16384 pixels wide
// and tiles at least 8. A pixel x is covered only if |x + 0.5 - cx| <= hx, so
// [cx - hx - 1, cx + hx + 1] is a conservative pixel range; hxy holds the f16
// half extents rounded up (0: draws no pixel).

__device__ __forceinline__ uint2 tile_rect(const Splat2D& s, float sd2, const TileGrid& g) {
  const uint2 none = make_uint2(0xFFFFFFFFu, 0u);
  if (s.hxy == 0u) return none;
  const float hx = __half2float(__ushort_as_half((unsigned short)(s.hxy & 0xFFFFu)));
  const float hy = __half2float(__ushort_as_half((unsigned short)(s.hxy >> 16)));
  const float fx0 = floorf(s.cx - hx - 1.0f), fx1 = floorf(s.cx + hx + 1.0f);
  const float fy0 = floorf(s.cy - hy - 1.0f), fy1 = floorf(s.cy + hy + 1.0f);
  const float wmax = float(g.width - 1), hmax = float(g.height - 1);
  if (!(fx1 >= 0.0f) || !(fx0 <= wmax) || !(fy1 >= 0.0f) || !(fy0 <= hmax)) return none;
  const uint32_t x0 = uint32_t(fmaxf(fx0, 0.0f)), x1 = uint32_t(fminf(fx1, wmax));
  const uint32_t y0 = uint32_t(fmaxf(fy0, 0.0f)), y1 = uint32_t(fminf(fy1, hmax));
  // tile sizes are powers of two (8x8, 16x16, 32x8): shifts, not divisions
  const uint32_t sx = uint32_t(__builtin_ctz(g.tile_w)), sy = uint32_t(__builtin_ctz(g.tile_h));
  const uint32_t tx0 = x0 >> sx, tx1 = x1 >> sx, ty0 = y0 >> sy, ty1 = y1 >> sy;
  const bool small = tx1 - tx0 < 4u && ty1 - ty0 < 4u;
  uint32_t mask = small ? rect_bits(tx1 - tx0 + 1u, ty1 - ty0 + 1u) : 0xFFFFu;
  // (a rectangle of one tile keeps it: the bound rarely misses all of it)
  if (s.rho > 0.0f && small && (tx0 != tx1 || ty0 != ty1)) {
    const EllipseRows e = ellipse_rows(s.ux, s.uy, s.rho, sd2 * 1.0001f);
    if (e.ok) {
      const float itw = __uint_as_float((127u - sx) << 23);  // 1 / tile_w, a power of two
      mask = 0u;
      for (uint32_t ty = ty0; ty <= ty1; ++ty) {
        // pixel centres of the tile row: [ty th + 0.5, (ty + 1) th - 0.5]
        const float by0 = float(ty * g.tile_h) + 0.5f - s.cy, by1 = by0 + float(g.tile_h - 1u);
        float xl, xr;
        if (!ellipse_row_span(e, by0, by1, xl, xr)) continue;
        // columns tx whose pixel centres [tx tw + 0.5, tx tw + tw - 0.5] - cx
        // meet [xl, xr], clamped to the rectangle
        const float fhi = fminf(floorf((s.cx + xr - 0.5f) * itw), float(tx1));
        const float flo = fmaxf(ceilf((s.cx + xl + 0.5f) * itw) - 1.0f, float(tx0));
        if (flo <= fhi)
          mask |= ((2u << (uint32_t(fhi) - tx0)) - (1u << (uint32_t(flo) - tx0))) << (4u * (ty - ty0));
      }
    }
  }
  return make_uint2(tx0 | (ty0 << 12) | ((mask & 0xFFu) << 24), tx1 | (ty1 << 12) | ((mask >> 8) << 24));
}

// The tiles of a packed rectangle (tile_rect): count and bounds.
struct TileRect {
  uint32_t tx0, ty0, tx1, ty1, mask;
  bool small;  // at most 4 x 4 tiles: `mask` selects among them
};

__device__ __forceinline__ TileRect unpack_rect(uint2 r) {
  TileRect t;
  t.tx0 = r.x & 0xFFFu;
  t.ty0 = (r.x >> 12) & 0xFFFu;
  t.tx1 = r.y & 0xFFFu;
  t.ty1 = (r.y >> 12) & 0xFFFu;
  t.mask = (r.x >> 24) | ((r.y >> 24) << 8);
  t.small = t.tx1 - t.tx0 < 4u && t.ty1 - t.ty0 < 4u;
  return t;
}

__device__ __forceinline__ uint32_t rect_tiles(const TileRect& t) {
  if (t.tx0 > t.tx1 || t.ty0 > t.ty1) return 0u;
  if (t.small) return uint32_t(__popc(t.mask & rect_bits(t.tx1 - t.tx0 + 1u, t.ty1 - t.ty0 + 1u)));
  return (t.tx1 - t.tx0 + 1u) * (t.ty1 - t.ty0 + 1u);
}

// First R rows of the column-major 4x4 matrix m times v, contracted to fused
// multiply-adds (for the splat record, which is not compared bit for bit).
template <int R, class T>
__device__ __forceinline__ void mv4_fast(const T* m, const float v[4], float* out) {
#pragma clang fp contract(fast)
#pragma unroll
  for (int r = 0; r < R; ++r) out[r] = m[0 * 4 + r] * v[0] + m[1 * 4 + r] * v[1] + m[2 * 4 + r] * v[2] + m[3 * 4 + r] * v[3];
}

template <class T>
__device__ __forceinline__ void mv3_fast(const T* m, const float v[3], float out[3]) {
#pragma clang fp contract(fast)
#pragma unroll
  for (int r = 0; r < 3; ++r) out[r] = m[0 * 3 + r] * v[0] + m[1 * 3 + r] * v[1] + m[2 * 3 + r] * v[2];
}

// The SH stream's words of one Gaussian, requested ahead of their use.
template <int POD>
struct ShWords {
  uint4 v[6];  // kPodHalf: 45 halves + 3 pad
};
template <>
struct ShWords<kPodSingle> {
  float4 v[12];
};

__device__ __forceinline__ int sh_bands(int sh_deg) { return sh_deg >= 3 ? 15 : (sh_deg == 2 ? 8 : (sh_deg == 1 ? 3 : 0)); }

// Requests the 16-byte chunks of coefficients the SH degree uses (all in
// flight at once; lower degrees touch only their planes).
template <int POD>
__device__ __forceinline__ void sh_load(const GaussianStreams& g, uint32_t i, int sh_deg, ShWords<POD>& w) {
  const int kmax = sh_bands(sh_deg);
  if constexpr (POD == kPodSingle) {
    const float4* p = g.sh32 + i;
    const int nq = (3 * kmax + 3) / 4;
#pragma unroll
    for (int q = 0; q < 12; ++q)
      if (q < nq) w.v[q] = p[size_t(q) * g.n];
  } else {
    const uint4* p = g.sh16 + i;
    const int nq = (3 * kmax + 7) / 8;
#pragma unroll
    for (int q = 0; q < 6; ++q)
      if (q < nq) w.v[q] = p[size_t(q) * g.n];
  }
}

// SH colour of one instance (utils.wesl:82-136) for the splat record: the
// 15 basis values of the view direction, then each loaded chunk folded into
// the three channel sums as packed words, without unpacking all 45
// coefficients at once. The sums run in coefficient order, not the shader's
// association (the record is compared within a tolerance, not bit for bit).
template <int POD>
__device__ __forceinline__ void sh_fold(const ShWords<POD>& w, float x, float y, float z, int sh_deg,
                                        float rgb[3]) {
#pragma clang fp contract(fast)
  if (sh_deg < 1) return;
  const float xx = x * x, yy = y * y, zz = z * z;
  const float xy = x * y, yz = y * z, xz = x * z;
  const float basis[15] = {
      -ShC::c1 * y, ShC::c1 * z, -ShC::c1 * x,
      ShC::c2_0 * xy, ShC::c2_1 * yz, ShC::c2_2 * ((2.0f * zz - xx) - yy), ShC::c2_3 * xz, ShC::c2_4 * (xx - yy),
      (ShC::c3_0 * y) * (3.0f * xx - yy), (ShC::c3_1 * xy) * z, (ShC::c3_2 * y) * ((4.0f * zz - xx) - yy),
      (ShC::c3_3 * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy), (ShC::c3_4 * x) * ((4.0f * zz - xx) - yy),
      (ShC::c3_5 * z) * (xx - yy), (ShC::c3_6 * x) * (xx - 3.0f * yy)};
  // coefficient j (0..44) is band-k coefficient k = j / 3 of channel j % 3;
  // bands start at k = 0 (degree 1), 3 (degree 2) and 8 (degree 3)
  const int kmax = sh_bands(sh_deg);
  auto add = [&](int j, float v) {
    if (j < 45 && j / 3 < kmax) rgb[j % 3] += basis[j / 3] * v;
  };
  if constexpr (POD == kPodSingle) {
    const int nq = (3 * kmax + 3) / 4;
#pragma unroll
    for (int q = 0; q < 12; ++q) {
      if (q >= nq) continue;
      add(4 * q + 0, w.v[q].x);
      add(4 * q + 1, w.v[q].y);
      add(4 * q + 2, w.v[q].z);
      add(4 * q + 3, w.v[q].w);
    }
  } else {
    const int nq = (3 * kmax + 7) / 8;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      if (q >= nq) continue;
      add(8 * q + 0, half_lo(w.v[q].x));
      add(8 * q + 1, half_hi(w.v[q].x));
      add(8 * q + 2, half_lo(w.v[q].y));
      add(8 * q + 3, half_hi(w.v[q].y));
      add(8 * q + 4, half_lo(w.v[q].z));
      add(8 * q + 5, half_hi(w.v[q].z));
      add(8 * q + 6, half_lo(w.v[q].w));
      add(8 * q + 7, half_hi(w.v[q].w));
    }
  }
}

// The splat record of one instance (render.wesl:75-130), in pixel space, in
// the functions below. `world` is model_transform_mat * (pos, 1); `c` the
// cov3d.
//
// Unlike the cull and depth-key path (math.hpp, bit-exact with the f32
// oracle), nothing here feeds a bit-exact output: the record only shapes
// pixels, which are compared within a tolerance. It is therefore evaluated
// with contraction, hardware reciprocal / square root (1 ulp), and the
// sparsity of the Jacobian J (utils.wesl:38-42), whose third row and two
// entries are zero: T = J W SR reduces to two rows, cov2d to three dot
// products.

// color(): render.wesl:58-73 — the view direction in model space and the SH
// colour, clamped at 0 (`sw` holds the Gaussian's loaded SH words).
template <int POD, class F>
__device__ __forceinline__ void splat_colour(const F& fc, const ShWords<POD>& sw, const float world[4],
                                             uint32_t color_bits, float rgb[3]) {
#pragma clang fp contract(fast)
  const float wvd[3] = {fc.campos[0] - world[0], fc.campos[1] - world[1], fc.campos[2] - world[2]};
  float mvd[3];
  mv3_fast(fc.inv_sr, wvd, mvd);
  const float il = __builtin_amdgcn_rsqf(mvd[0] * mvd[0] + mvd[1] * mvd[1] + mvd[2] * mvd[2]);
  const float x = -mvd[0] * il, y = -mvd[1] * il, z = -mvd[2] * il;
#pragma unroll
  for (int ch = 0; ch < 3; ++ch)
    rgb[ch] = fc.no_sh0 ? 0.5f : float((color_bits >> (8 * ch)) & 0xFFu) * (1.0f / 255.0f);
  sh_fold<POD>(sw, x, y, z, fc.sh_deg, rgb);
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) rgb[ch] = fmaxf(rgb[ch], 0.0f);
}

// A splat's minor eigenvalue lambda_min = mid - radius (utils.wesl:54-79)
// cancels when the splat is thin: below kThinEigenRatio * lambda_max its
// rounding decides the minor axis (and whether the splat is drawn at all), so
// such splats take their axes from the exact operation order of the oracle
// and the reference's shader (math.hpp's cov2d_axes on the cov3d rebuilt in
// that order) instead of the contracted evaluation.
constexpr float kThinEigenRatio = 1e-3f;
constexpr float kThinRho = 31.6227766f;  // 1 / sqrt(kThinEigenRatio)

// The record's axes (u, rho) and half extents of a thin splat, in that exact
// order (render.wesl:75-130 as restated by the oracle: u = 2 (M.x, -M.y) /
// |M|^2, v = 2 (m.x, -m.y) / |m|^2, of which the record keeps
// rho = |M| / |m|); ok is false when the splat is not drawn.
template <int POD, class F>
__device__ __forceinline__ void thin_splat_axes(const F& fc, const float pos[3], const CovWords& cw,
                                             float& ux, float& uy, float& rho, float& hx, float& hy, bool& ok) {
  float c[6];
  cov3d_of<POD, false>(cw, c);
  const float sd = fc.max_std_dev;
  float ax[4];
  cov2d_axes(fc, pos, c, sd * fc.gsize, ax);
  const float mm = ax[0] * ax[0] + ax[1] * ax[1];
  const float nn = ax[2] * ax[2] + ax[3] * ax[3];
  ok = (mm > 0.0f) && (nn > 0.0f);
  ux = uy = rho = hx = hy = 0.0f;
  if (!ok) return;
  ux = (2.0f * ax[0]) / mm;
  uy = -((2.0f * ax[1]) / mm);
  rho = sqrtf(mm / nn);
  hx = (0.5f * sd) * sqrtf(ax[0] * ax[0] + ax[2] * ax[2]);
  hy = (0.5f * sd) * sqrtf(ax[1] * ax[1] + ax[3] * ax[3]);
}

// The rest of the record: centre, axes, bounds and tile rectangle, with the
// colour from splat_colour. pos and cw (the model-space position and the
// covariance words) serve the exact axes of thin splats: with kInlineThin
// they are computed here; otherwise the record keeps the contracted axes and
// the return value (true) asks the caller to have thin_splat_record redo
// them (the projection kernel queues such splats for a kernel of their own,
// so that the exact evaluation adds nothing to its registers).
template <int POD, bool kInlineThin, class F>
__device__ __forceinline__ bool splat_geometry(const F& fc, const float world[4], const float c[6],
                                               const float pos[3], const CovWords& cw, uint32_t color_bits,
                                               const float rgb[3], Splat2D* out, float* depth_out, uint2* rect_out,
                                               const TileGrid* rect_grid) {
#pragma clang fp contract(fast)
  // the view-space centre (rows 0-2) and the clip-space centre
  // (proj * view) * world
  float view[3], proj[4];
  mv4_fast<3>(fc.view, world, view);
  mv4_fast<4>(fc.pv, world, proj);
  float alpha = float(color_bits >> 24) * (1.0f / 255.0f);

  const float iw = __builtin_amdgcn_rcpf(proj[3]);
  if (depth_out) {
    // fragment depth: every corner of the quad keeps the centre's clip z and
    // w (render.wesl:119-125), so the rasterised depth is z / w, here in the
    // vertex stage's uncontracted association proj * (view * world) and an
    // IEEE divide (depth tests compare it with other draws' depths)
    float v4[4], c4[4];
    m4v(fc.view, world, v4);
    m4v(fc.proj, v4, c4);
    *depth_out = c4[2] / c4[3];
  }
  Splat2D s;
  bool thin = false;
  s.cx = (proj[0] * iw + 1.0f) * 0.5f * fc.size[0];
  s.cy = (1.0f - proj[1] * iw) * 0.5f * fc.size[1];
  s.ux = s.uy = s.rho = 0.0f;
  float hx = 0.0f, hy = 0.0f;
  bool drawable = false;
  // depth clipping of the rasterizer: 0 <= z_clip <= w_clip
  const bool zclip_ok = (proj[2] >= 0.0f) && (proj[2] <= proj[3]);

  if (fc.display_mode == kModePoint) {
    // render.wesl:92-104
    const float aspect = fc.size[1] / fc.size[0];
    const float ivl = __builtin_amdgcn_rsqf(view[0] * view[0] + view[1] * view[1] + view[2] * view[2]);
    const float q = 0.01f * fc.gsize;
    hx = q * aspect * ivl * 0.5f * fc.size[0];
    hy = q * ivl * 0.5f * fc.size[1];
    s.ux = hx;
    s.uy = hy;
    alpha = 1.0f;
    drawable = zclip_ok;
  } else {
    // cov2d (utils.wesl:25-48) with the zero entries of J dropped; its
    // t = (view * model) * pos is the view-space centre computed above as
    // view * (model * pos) (the same point, rounded in another order)
    const float* t = view;
    const float itz = __builtin_amdgcn_rcpf(t[2]);
    const float j00 = fc.focal[0] * itz, j11 = fc.focal[1] * itz;
    const float j20 = -fc.focal[0] * t[0] * itz * itz, j21 = -fc.focal[1] * t[1] * itz * itz;
    // rows of T = J * (W SR): T0 = j00 * R0 + j20 * R2, T1 = j11 * R1 + j21 * R2
    float t0[3], t1[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      t0[k] = j00 * fc.wsr[k * 3 + 0] + j20 * fc.wsr[k * 3 + 2];
      t1[k] = j11 * fc.wsr[k * 3 + 1] + j21 * fc.wsr[k * 3 + 2];
    }
    // Sigma T0^T and Sigma T1^T (Sigma symmetric: xx xy xz yy yz zz)
    const float s0x = c[0] * t0[0] + c[1] * t0[1] + c[2] * t0[2];
    const float s0y = c[1] * t0[0] + c[3] * t0[1] + c[4] * t0[2];
    const float s0z = c[2] * t0[0] + c[4] * t0[1] + c[5] * t0[2];
    const float s1x = c[0] * t1[0] + c[1] * t1[1] + c[2] * t1[2];
    const float s1y = c[1] * t1[0] + c[3] * t1[1] + c[4] * t1[2];
    const float s1z = c[2] * t1[0] + c[4] * t1[1] + c[5] * t1[2];
    const float ca = t0[0] * s0x + t0[1] * s0y + t0[2] * s0z;
    const float cb = t1[0] * s0x + t1[1] * s0y + t1[2] * s0z;
    const float cc = t1[0] * s1x + t1[1] * s1y + t1[2] * s1z;
    // cov2d_axes (utils.wesl:54-79)
    const float sd = fc.max_std_dev;
    const float sds = sd * fc.gsize;
    const float mid = 0.5f * (ca + cc);
    const float hd = 0.5f * (ca - cc);
    const float radius = __builtin_amdgcn_sqrtf(hd * hd + cb * cb);
    const float lmaj = mid + radius, lmin = mid - radius;
    bool ok = zclip_ok && lmin >= 0.0f;
    if (kInlineThin) thin = zclip_ok && !(lmin >= kThinEigenRatio * lmaj);
    if (thin && kInlineThin) {
      // thin (or degenerate): the exact order decides
      thin_splat_axes<POD>(fc, pos, cw, s.ux, s.uy, s.rho, hx, hy, ok);
      thin = false;
    } else if (ok) {
      const float dx0 = cb, dy0 = lmaj - ca;
      float dx = 0.0f, dy = 1.0f;
      if (!(dx0 == 0.0f && dy0 == 0.0f)) {
        const float inv = __builtin_amdgcn_rsqf(dx0 * dx0 + dy0 * dy0);
        dx = dx0 * inv;
        dy = dy0 * inv;
      }
      const float maj = fminf(sds * __builtin_amdgcn_sqrtf(lmaj), 1024.0f);
      const float mnr = fminf(sds * __builtin_amdgcn_sqrtf(lmin), 1024.0f);
      ok = (maj > 0.0f) && (mnr > 0.0f);
      if (ok) {
        // qo = (d.u, d.v) with u = 2 (M.x, -M.y) / |M|^2, v = 2 (m.x, -m.y) / |m|^2,
        // M = maj (dx, dy), m = mnr (dy, -dx): u = (2 / maj) (dx, -dy) and
        // v = (2 / mnr) (dy, dx) = (maj / mnr) (-u.y, u.x)
        const float im = 2.0f * __builtin_amdgcn_rcpf(maj);
        s.ux = dx * im;
        s.uy = -dy * im;
        s.rho = maj * __builtin_amdgcn_rcpf(mnr);
        hx = 0.5f * sd * __builtin_amdgcn_sqrtf(maj * maj * dx * dx + mnr * mnr * dy * dy);
        hy = 0.5f * sd * __builtin_amdgcn_sqrtf(maj * maj * dy * dy + mnr * mnr * dx * dx);
      }
    }
    // queued for the exact axes: lambda_min below kThinEigenRatio lambda_max,
    // i.e. an axis ratio rho = sqrt(lambda_max / lambda_min) above
    // 1 / sqrt(kThinEigenRatio), or no axes at all
    if (!kInlineThin) thin = zclip_ok && !(ok && s.rho <= kThinRho);
    drawable = ok;
  }
  s.rg = pack_half2(rgb[0], rgb[1]);
  s.ba = pack_half2(rgb[2], alpha);
  s.hxy = drawable ? pack_half2_up(hx, hy) : 0u;
  *out = s;
  if (rect_out) *rect_out = tile_rect(s, fc.max_std_dev * fc.max_std_dev, *rect_grid);
  return thin;
}

// Redoes the axes, half extents and tile rectangle of a thin splat's record
// (splat_geometry returned true for it) in the exact order; the centre and
// colour stay.
template <int POD, class F>
__device__ __forceinline__ void thin_splat_record(const F& fc, const float pos[3], const CovWords& cw,
                                                  Splat2D* out, uint2* rect_out, const TileGrid* rect_grid) {
  Splat2D s = *out;
  float hx, hy;
  bool ok;
  thin_splat_axes<POD>(fc, pos, cw, s.ux, s.uy, s.rho, hx, hy, ok);
  s.hxy = ok ? pack_half2_up(hx, hy) : 0u;
  *out = s;
  if (rect_out) *rect_out = tile_rect(s, fc.max_std_dev * fc.max_std_dev, *rect_grid);
}

// render.wesl:75-130 for one instance: the SH words are loaded here.
template <int POD, class F>
__device__ __forceinline__ void compute_splat(const F& fc, const GaussianStreams& g, uint32_t i,
                                              uint32_t color_bits, const float world[4], const float c[6],
                                              const float pos[3], const CovWords& cw, Splat2D* out,
                                              float* depth_out, uint2* rect_out = nullptr,
                                              const TileGrid* rect_grid = nullptr) {
  ShWords<POD> sw;
  sh_load<POD>(g, i, fc.sh_deg, sw);
  float rgb[3];
  splat_colour<POD>(fc, sw, world, color_bits, rgb);
  splat_geometry<POD, true>(fc, world, c, pos, cw, color_bits, rgb, out, depth_out, rect_out, rect_grid);
}

}  // namespace gs3d
