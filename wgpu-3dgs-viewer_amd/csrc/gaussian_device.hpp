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
//               sh   : float4[12] 45 SH floats + 3 pad          192 B
//   kPodHalf:   rs   : uint4  rot xyzw f16, scale xyz f16, pad   16 B
//               sh   : uint4[6] 45 SH halves + 3 pad             96 B
// The culled majority only reads `pc` (and the covariance stream when the
// centre is off screen).
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

template <int POD>
__device__ __forceinline__ void load_cov3d(const GaussianStreams& g, uint32_t i, float c[6]) {
  if constexpr (POD == kPodSingle) {
    const float2* p = g.cov + size_t(i) * 3;
    const float2 a = p[0], b = p[1], d = p[2];
    c[0] = a.x;
    c[1] = a.y;
    c[2] = b.x;
    c[3] = b.y;
    c[4] = d.x;
    c[5] = d.y;
  } else {
    const uint4 w = g.rs[i];
    const float q[4] = {half_lo(w.x), half_hi(w.x), half_lo(w.y), half_hi(w.y)};
    const float s[3] = {half_lo(w.z), half_hi(w.z), half_lo(w.w)};
    cov3d_from_scale_rot(s, q, c);
  }
}

template <int POD>
__device__ __forceinline__ void load_sh(const GaussianStreams& g, uint32_t i, float sh[48]) {
  if constexpr (POD == kPodSingle) {
    const float4* p = g.sh32 + size_t(i) * 12;
#pragma unroll
    for (int q = 0; q < 12; ++q) {
      const float4 v = p[q];
      sh[4 * q + 0] = v.x;
      sh[4 * q + 1] = v.y;
      sh[4 * q + 2] = v.z;
      sh[4 * q + 3] = v.w;
    }
  } else {
    const uint4* p = g.sh16 + size_t(i) * 6;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const uint4 v = p[q];
      sh[8 * q + 0] = half_lo(v.x);
      sh[8 * q + 1] = half_hi(v.x);
      sh[8 * q + 2] = half_lo(v.y);
      sh[8 * q + 3] = half_hi(v.y);
      sh[8 * q + 4] = half_lo(v.z);
      sh[8 * q + 5] = half_hi(v.z);
      sh[8 * q + 6] = half_lo(v.w);
      sh[8 * q + 7] = half_hi(v.w);
    }
  }
}

// render.wesl:75-130 for one instance, in pixel space. `world` is
// model_transform_mat * (pos, 1); `c` the cov3d.
template <int POD>
__device__ __forceinline__ void compute_splat(const FrameConsts& fc, const GaussianStreams& g, uint32_t i,
                                              const float pos[3], uint32_t color_bits, const float world[4],
                                              const float c[6], Splat2D* out) {
  float view[4], proj[4];
  m4v(fc.view, world, view);
  m4v(fc.proj, view, proj);

  // color(): render.wesl:58-73
  const float wvd[3] = {fc.campos[0] - world[0], fc.campos[1] - world[1], fc.campos[2] - world[2]};
  float mvd[3];
  m3v(fc.inv_sr, wvd, mvd);
  const float len = sqrtf((mvd[0] * mvd[0] + mvd[1] * mvd[1]) + mvd[2] * mvd[2]);
  const float x = -(mvd[0] / len), y = -(mvd[1] / len), z = -(mvd[2] / len);
  float sh[48];
  if (fc.sh_deg >= 1) load_sh<POD>(g, i, sh);
  float rgb[3];
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    const float base = fc.no_sh0 ? 0.5f : float((color_bits >> (8 * ch)) & 0xFFu) / 255.0f;
    rgb[ch] = view_color_channel(base, x, y, z, fc.sh_deg, [&](int k) { return sh[k * 3 + ch]; });
  }
  float alpha = float(color_bits >> 24) / 255.0f;

  const float ndx = proj[0] / proj[3];
  const float ndy = proj[1] / proj[3];
  Splat2D s;
  s.cx = ((ndx + 1.0f) * 0.5f) * fc.size[0];
  s.cy = ((1.0f - ndy) * 0.5f) * fc.size[1];
  s.ux = s.uy = s.vx = s.vy = 0.0f;
  s.hx = s.hy = 0.0f;
  s.pad[0] = s.pad[1] = s.pad[2] = 0u;
  // depth clipping of the rasterizer: 0 <= z_clip <= w_clip
  const bool zclip_ok = (proj[2] >= 0.0f) && (proj[2] <= proj[3]);

  if (fc.display_mode == kModePoint) {
    // render.wesl:92-104
    const float aspect = fc.size[1] / fc.size[0];
    const float vlen = sqrtf((view[0] * view[0] + view[1] * view[1]) + view[2] * view[2]);
    const float q = 0.01f * fc.gsize;
    s.hx = (((q * aspect) / vlen) * 0.5f) * fc.size[0];
    s.hy = ((q / vlen) * 0.5f) * fc.size[1];
    alpha = 1.0f;
    s.drawable = zclip_ok ? 1u : 0u;
  } else {
    const float sd = fc.max_std_dev;
    float ax[4];
    cov2d_axes(fc, pos, c, sd * fc.gsize, ax);
    const float mm = ax[0] * ax[0] + ax[1] * ax[1];
    const float nn = ax[2] * ax[2] + ax[3] * ax[3];
    const bool ok = zclip_ok && (mm > 0.0f) && (nn > 0.0f);
    s.drawable = ok ? 1u : 0u;
    if (ok) {
      s.ux = (2.0f * ax[0]) / mm;
      s.uy = -((2.0f * ax[1]) / mm);
      s.vx = (2.0f * ax[2]) / nn;
      s.vy = -((2.0f * ax[3]) / nn);
      s.hx = (0.5f * sd) * sqrtf(ax[0] * ax[0] + ax[2] * ax[2]);
      s.hy = (0.5f * sd) * sqrtf(ax[1] * ax[1] + ax[3] * ax[3]);
    }
  }
  s.r = rgb[0];
  s.g = rgb[1];
  s.b = rgb[2];
  s.a = alpha;
  *out = s;
}

}  // namespace gs3d
