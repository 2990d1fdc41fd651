// f32 arithmetic of the splat path in a fixed, documented operation order.
//
// The library is compiled with -ffp-contract=off so that every expression
// below is evaluated exactly as written (no fused multiply-add, no
// reassociation). Matrices are column-major (m[c*4+r]) as in WGSL.
//   dot3(a,b)  = (a0*b0 + a1*b1) + a2*b2
//   mat * vec  = row r: ((m[0][r]*v0 + m[1][r]*v1) + m[2][r]*v2) + m[3][r]*v3
//   mat * mat  = (A*B)[c][r] = sum over k of A[k][r]*B[c][k], left to right
//   normalize  = v / sqrt(dot(v,v)) per component; sqrt and / correctly rounded
//   min / max  = IEEE minNum / maxNum (a NaN operand yields the other operand)
#pragma once

#include <cmath>
#include <cstdint>

#include "types.hpp"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GS3D_HD __host__ __device__ __forceinline__
#else
#define GS3D_HD inline
#endif

namespace gs3d {

// (T: float in any address space)
template <class T>
GS3D_HD void m4v(const T* m, const float v[4], float out[4]) {
#pragma unroll
  for (int r = 0; r < 4; ++r)
    out[r] = ((m[0 * 4 + r] * v[0] + m[1 * 4 + r] * v[1]) + m[2 * 4 + r] * v[2]) + m[3 * 4 + r] * v[3];
}

GS3D_HD void m4m4(const float* a, const float* b, float* out) {
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r)
      out[c * 4 + r] =
          ((a[0 * 4 + r] * b[c * 4 + 0] + a[1 * 4 + r] * b[c * 4 + 1]) + a[2 * 4 + r] * b[c * 4 + 2]) +
          a[3 * 4 + r] * b[c * 4 + 3];
}

template <class A, class B>
GS3D_HD void m3m3(const A* a, const B* b, float* out) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r)
      out[c * 3 + r] = (a[0 * 3 + r] * b[c * 3 + 0] + a[1 * 3 + r] * b[c * 3 + 1]) + a[2 * 3 + r] * b[c * 3 + 2];
}

GS3D_HD void m3v(const float* m, const float v[3], float out[3]) {
#pragma unroll
  for (int r = 0; r < 3; ++r) out[r] = (m[0 * 3 + r] * v[0] + m[1 * 3 + r] * v[1]) + m[2 * 3 + r] * v[2];
}

// Rotation of a quaternion (x, y, z, w), glam's Mat3::from_quat expansion.
GS3D_HD void quat_to_mat3(const float q[4], float m[9]) {
  const float x2 = q[0] + q[0], y2 = q[1] + q[1], z2 = q[2] + q[2];
  const float xx = q[0] * x2, xy = q[0] * y2, xz = q[0] * z2;
  const float yy = q[1] * y2, yz = q[1] * z2, zz = q[2] * z2;
  const float wx = q[3] * x2, wy = q[3] * y2, wz = q[3] * z2;
  m[0] = 1.0f - (yy + zz);
  m[1] = xy + wz;
  m[2] = xz - wy;
  m[3] = xy - wz;
  m[4] = 1.0f - (xx + zz);
  m[5] = yz + wx;
  m[6] = xz + wy;
  m[7] = yz - wx;
  m[8] = 1.0f - (xx + yy);
}

// cov3d = (R S)(R S)^T as (xx, xy, xz, yy, yz, zz), the order
// gaussian_unpack_cov3d feeds utils.wesl:29-33.
GS3D_HD void cov3d_from_scale_rot(const float s[3], const float q[4], float out[6]) {
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
    out[e] = (mm[0 * 3 + ia[e]] * mm[0 * 3 + ib[e]] + mm[1 * 3 + ia[e]] * mm[1 * 3 + ib[e]]) +
             mm[2 * 3 + ia[e]] * mm[2 * 3 + ib[e]];
}

// model_transform_mat, model_scale_rot_mat, model_transform_inv_sr_mat
// (wgpu-3dgs-core 0.6.0): M = T R S, SR = R S, INV_SR = S^-1 R^T.
GS3D_HD void model_mats(const ModelTransformPod& t, float model[16], float sr[9], float inv_sr[9]) {
  float rm[9];
  quat_to_mat3(t.rot, rm);
  for (int c = 0; c < 3; ++c) {
    for (int r = 0; r < 3; ++r) {
      sr[c * 3 + r] = rm[c * 3 + r] * t.scale[c];
      model[c * 4 + r] = sr[c * 3 + r];
      inv_sr[c * 3 + r] = rm[r * 3 + c] / t.scale[r];
    }
    model[c * 4 + 3] = 0.0f;
  }
  model[12] = t.pos[0];
  model[13] = t.pos[1];
  model[14] = t.pos[2];
  model[15] = 1.0f;
}

// Everything a frame needs that does not depend on the Gaussian.
GS3D_HD void derive_frame_consts(const CameraPod& cam, const ModelTransformPod& mt, const GaussianTransformPod& gt,
                                 FrameConsts& fc) {
  for (int i = 0; i < 16; ++i) {
    fc.view[i] = cam.view[i];
    fc.proj[i] = cam.proj[i];
  }
  model_mats(mt, fc.model, fc.sr, fc.inv_sr);
  m4m4(fc.proj, fc.view, fc.pv);
  m4m4(fc.view, fc.model, fc.vm);
  const float w3[9] = {cam.view[0], cam.view[1], cam.view[2], cam.view[4], cam.view[5],
                       cam.view[6], cam.view[8], cam.view[9], cam.view[10]};
  m3m3(w3, fc.sr, fc.wsr);
  for (int i = 0; i < 3; ++i)
    fc.campos[i] = -((cam.view[i * 4 + 0] * cam.view[12] + cam.view[i * 4 + 1] * cam.view[13]) +
                     cam.view[i * 4 + 2] * cam.view[14]);
  fc.size[0] = cam.size[0];
  fc.size[1] = cam.size[1];
  fc.focal[0] = (cam.proj[0] * cam.size[0]) * 0.5f;
  fc.focal[1] = (cam.proj[5] * cam.size[1]) * 0.5f;
  fc.gsize = gt.size;
  fc.max_std_dev = gt.max_std_dev;
  fc.display_mode = int32_t(gt.flags & 3u);
  fc.sh_deg = int32_t((gt.flags >> 2) & 3u);
  fc.no_sh0 = int32_t((gt.flags >> 4) & 1u);
}

// utils.wesl:18-20.
GS3D_HD bool cull(float x, float y, float z) {
  return !((x >= -1.0f && y >= -1.0f && z >= 0.0f) && (x <= 1.0f && y <= 1.0f && z <= 1.0f));
}

// utils.wesl:25-48: (cov.xx, cov.xy, cov.yy) in pixels^2.
// (F: FrameConsts in any address space)
template <class F>
GS3D_HD void cov2d(const F& fc, const float pos[3], const float c[6], float out[3]) {
  const float vrk[9] = {c[0], c[1], c[2], c[1], c[3], c[4], c[2], c[4], c[5]};
  const float p4[4] = {pos[0], pos[1], pos[2], 1.0f};
  float t[4];
  m4v(fc.vm, p4, t);
  const float tz2 = t[2] * t[2];
  const float j[9] = {fc.focal[0] / t[2], 0.0f, 0.0f, 0.0f, fc.focal[1] / t[2], 0.0f,
                      (-(fc.focal[0] * t[0])) / tz2, (-(fc.focal[1] * t[1])) / tz2, 0.0f};
  const float wm[9] = {fc.view[0], fc.view[1], fc.view[2], fc.view[4], fc.view[5],
                       fc.view[6], fc.view[8], fc.view[9], fc.view[10]};
  float jw[9], tm[9], tv[9], tt[9], cv[9];
  m3m3(j, wm, jw);
  m3m3(jw, fc.sr, tm);
  m3m3(tm, vrk, tv);
#pragma unroll
  for (int cc = 0; cc < 3; ++cc)
#pragma unroll
    for (int r = 0; r < 3; ++r) tt[cc * 3 + r] = tm[r * 3 + cc];
  m3m3(tv, tt, cv);
  out[0] = cv[0];
  out[1] = cv[1];
  out[2] = cv[4];
}

// utils.wesl:54-79: (major.xy, minor.xy); all zero when lambda_min < 0.
template <class F>
GS3D_HD void cov2d_axes(const F& fc, const float pos[3], const float c[6], float sd, float out[4]) {
  float cv[3];
  cov2d(fc, pos, c, cv);
  const float mid = 0.5f * (cv[0] + cv[2]);
  const float hx = 0.5f * (cv[0] - cv[2]);
  const float radius = sqrtf(hx * hx + cv[1] * cv[1]);
  const float lmaj = mid + radius;
  const float lmin = mid - radius;
  if (lmin < 0.0f) {
    out[0] = out[1] = out[2] = out[3] = 0.0f;
    return;
  }
  const float dx0 = cv[1];
  const float dy0 = lmaj - cv[0];
  float dx = 0.0f, dy = 1.0f;
  if (!(dx0 == 0.0f && dy0 == 0.0f)) {
    const float len = sqrtf(dx0 * dx0 + dy0 * dy0);
    dx = dx0 / len;
    dy = dy0 / len;
  }
  const float maj = fminf(sd * sqrtf(lmaj), 1024.0f);
  const float mnr = fminf(sd * sqrtf(lmin), 1024.0f);
  out[0] = maj * dx;
  out[1] = maj * dy;
  out[2] = mnr * dy;
  out[3] = mnr * (-dx);
}

// SH band constants (utils.wesl:83-87), as f32.
struct ShC {
  static constexpr float c1 = 0.4886025f;
  static constexpr float c2_0 = 1.0925484f, c2_1 = -1.0925484f, c2_2 = 0.3153916f, c2_3 = -1.0925484f,
                         c2_4 = 0.5462742f;
  static constexpr float c3_0 = -0.5900436f, c3_1 = 2.8906114f, c3_2 = -0.4570458f, c3_3 = 0.3731763f,
                         c3_4 = -0.4570458f, c3_5 = 1.4453057f, c3_6 = -0.5900436f;
};

// utils.wesl:82-136 for one colour channel. sh(k) returns coefficient k of
// this channel.
template <typename ShFn>
GS3D_HD float view_color_channel(float base, float x, float y, float z, int sh_deg, ShFn sh) {
  float res = base;
  if (sh_deg >= 1) {
    res = res + ShC::c1 * (((-sh(0)) * y + sh(1) * z) - sh(2) * x);
    if (sh_deg >= 2) {
      const float xx = x * x, yy = y * y, zz = z * z;
      const float xy = x * y, yz = y * z, xz = x * z;
      res = res + (((((ShC::c2_0 * xy) * sh(3) + (ShC::c2_1 * yz) * sh(4)) +
                     (ShC::c2_2 * ((2.0f * zz - xx) - yy)) * sh(5)) +
                    (ShC::c2_3 * xz) * sh(6)) +
                   (ShC::c2_4 * (xx - yy)) * sh(7));
      if (sh_deg >= 3) {
        res = res + ((((((((ShC::c3_0 * y) * (3.0f * xx - yy)) * sh(8) + ((ShC::c3_1 * xy) * z) * sh(9)) +
                         ((ShC::c3_2 * y) * ((4.0f * zz - xx) - yy)) * sh(10)) +
                        ((ShC::c3_3 * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy)) * sh(11)) +
                       ((ShC::c3_4 * x) * ((4.0f * zz - xx) - yy)) * sh(12)) +
                      ((ShC::c3_5 * z) * (xx - yy)) * sh(13)) +
                     ((ShC::c3_6 * x) * (xx - 3.0f * yy)) * sh(14));
      }
    }
  }
  return fmaxf(res, 0.0f);
}

}  // namespace gs3d
