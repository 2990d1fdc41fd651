/*
 * CPU oracle for the per-frame splat path of wgpu-3dgs-viewer 0.6.1.
 *
 * TEST INFRASTRUCTURE ONLY. Tests compare the product against it: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product (libgs3d.so) never links, loads or calls it.
 *
 * It restates, in plain C, the arithmetic of the reference's WGSL shaders and
 * of the fixed-function stages the reference relies on:
 *   src/shader/preprocess.wesl  (pre / main / post)
 *   src/shader/utils.wesl       (cull, cov2d, cov2d_axes, view_color)
 *   src/shader/camera.wesl      (world_to_camera)
 *   src/shader/render.wesl      (vert_main, color, splat, ellipse, point)
 *   src/shader/radix_sort.wgsl  (semantics only: stable ascending u32 sort)
 *   src/renderer.rs:169-181, 296-303 (clear to black, ALPHA_BLENDING)
 *   src/camera.rs:85-92, src/buffer/camera.rs:61-79 (camera matrices, pod)
 *
 * Floating-point convention (the "documented order" both this oracle and the
 * HIP kernels follow, compiled without FP contraction):
 *   dot2(a,b)   = a0*b0 + a1*b1
 *   dot3(a,b)   = (a0*b0 + a1*b1) + a2*b2
 *   mat*vec     = row r: ((m[0][r]*v0 + m[1][r]*v1) + m[2][r]*v2) + m[3][r]*v3
 *   mat*mat     = (A*B)[c][r] = sum over k = 0.. of A[k][r]*B[c][k], left to right
 *   length(v)   = sqrt(dot(v,v)); normalize(v) = v / length(v) (per component)
 *   min         = IEEE minNum (fminf): a NaN operand yields the other operand
 * Matrices are column-major (m[c*4+r]) as in WGSL and glam.
 *
 * Every function is available in float ("_f32", bit-exact target of the GPU
 * path) and double ("_f64", the FP64 host oracle used to state the error of
 * the FP16 storage form).
 *
 * Parity status per function (how each is pinned is in tests/test_oracle.py):
 *   orc_camera_*            pinned by closed forms (orthonormal view, eye maps
 *                           to origin, known projection entries)
 *   orc_cov3d_*             pinned against numpy R S S^T R^T in float64
 *   orc_preprocess_*        pinned by closed-form depth keys, frustum special
 *                           cases, the reference's own predicates
 *   orc_cov2d_axes_*        pinned against numpy float64 eigendecomposition
 *   orc_view_color_*        pinned against real spherical harmonics built
 *                           from scipy.special.sph_harm
 *   orc_stable_sort         pinned against numpy.argsort(kind="stable")
 *   orc_render_*            pinned by the closed-form single-splat image and
 *                           the two-splat over-operator; the f64 twin by the
 *                           same closed form
 *   orc_model_mats          pinned by model (p,1) = R(q) diag(s) p + t with the
 *                           textbook rotation matrix in float64, a known
 *                           90-degree rotation, and inv_sr sr = I
 *   orc_splats_* (colour)   the view direction fed to view_color pinned by
 *                           axis-aligned cameras (one degree-1 term each) and
 *                           by numpy's inv(R S) (campos - world) under random
 *                           model transforms
 *   orc_splats_* (geometry) the pixel-space footprint pinned against an EWA
 *                           footprint built from the projection's numerical
 *                           Jacobian (pixel y down) for tilted 10:1 needles
 *   display modes           ellipse ring and interior and the point quad's
 *                           half extent 0.01 size H / (2 |view|) pinned by
 *                           closed forms
 *   orc_frame_f32           equal to the composition of the pinned stages
 * Each of these checks was shown to fail on planted slips (a swapped or
 * transposed factor, a dropped sign or y flip, a wrong constant).
 * The wgpu-3dgs-core 0.6.0 helpers (Gaussian packing, flag decoding) are
 * absent from the reference tree: their restatement here is "parity unpinned"
 * against the reference itself and is pinned only by the closed forms above.
 */
#ifndef GS3D_ORACLE_H
#define GS3D_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Display modes, as core's GaussianDisplayMode (render.wesl:12-14). */
enum { ORC_MODE_SPLAT = 0, ORC_MODE_ELLIPSE = 1, ORC_MODE_POINT = 2 };

/* Blend targets (src/renderer.rs:296-300 with the caller's texture format). */
enum {
    ORC_TARGET_RGBA8_STEP = 0,  /* Rgba8Unorm: clamp src, blend, store unorm8 per step */
    ORC_TARGET_RGBA8_F32 = 1,   /* clamp src, blend in f32, quantise once at the end */
    ORC_TARGET_RGBA32F = 2      /* Rgba32Float: no clamp, blend in f32 */
};

/* Per-frame uniforms: camera pod, model transform matrices and gaussian
 * transform, as bound at preprocess.wesl:11-18 and render.wesl:31-38. */
typedef struct orc_frame {
    float view[16];
    float proj[16];
    float size[2];
    float model[16];   /* model_transform_mat */
    float sr[9];       /* model_scale_rot_mat */
    float inv_sr[9];   /* model_transform_inv_sr_mat */
    float gsize;       /* gaussian_transform.size */
    int32_t display_mode;
    int32_t sh_deg;
    int32_t no_sh0;
    float max_std_dev;
} orc_frame;

/* Per visible instance, what vert_main (render.wesl:75-130) hands the
 * rasterizer, restated in pixel space: centre c, inverse axis rows u and v
 * (qo = (d.u, d.v) for d = pixel - c), colour, and a drawable flag. */
typedef struct orc_splat_f32 {
    float cx, cy;
    float ux, uy, vx, vy;
    float r, g, b, a;
    float hx, hy;      /* point mode: half extents in pixels */
    int32_t drawable;
    int32_t pad;
    float z;           /* fragment depth: clip z / w of the centre */
} orc_splat_f32;

typedef struct orc_splat_f64 {
    double cx, cy;
    double ux, uy, vx, vy;
    double r, g, b, a;
    double hx, hy;
    int32_t drawable;
    int32_t pad;
    double z;
} orc_splat_f64;

/* ---- camera (src/camera.rs:85-92; glam look_to_rh / perspective_rh) ---- */
void orc_camera_forward(float pitch, float yaw, float out[3]);
void orc_look_to_rh(const float eye[3], const float dir[3], const float up[3], float out[16]);
void orc_perspective_rh(float fov_y, float aspect, float z_near, float z_far, float out[16]);

/* ---- core restatements (wgpu-3dgs-core 0.6.0, absent: parity unpinned) ---- */
void orc_cov3d_f32(const float scale[3], const float rot_xyzw[4], float out[6]);
void orc_cov3d_f64(const double scale[3], const double rot_xyzw[4], double out[6]);
/* the same over n records (scale n x 3, rot n x 4, out n x 6) */
void orc_cov3d_f32_n(uint64_t n, const float* scale, const float* rot_xyzw, float* out);
void orc_cov3d_f64_n(uint64_t n, const double* scale, const double* rot_xyzw, double* out);
void orc_model_mats(const float pos[3], const float rot_xyzw[4], const float scale[3],
                    float model[16], float sr[9], float inv_sr[9]);

/* ---- preprocess.wesl:60-106 for every Gaussian; output in index order ----
 * Writes the visible indices and their depth keys (1 - ndc.z) in ascending
 * index order and returns the visible count. */
uint32_t orc_preprocess_f32(const orc_frame* f, uint32_t n, const float* pos,
                            const float* cov3d, uint32_t* out_idx, float* out_key);

/* utils.wesl:54-79 on one Gaussian: (major.xy, minor.xy). */
void orc_cov2d_axes_f32(const orc_frame* f, const float pos[3], const float cov3d[6],
                        float std_dev, float out[4]);
void orc_cov2d_axes_f64(const orc_frame* f, const double pos[3], const double cov3d[6],
                        double std_dev, double out[4]);
void orc_cov2d_f32(const orc_frame* f, const float pos[3], const float cov3d[6], float out[3]);

/* utils.wesl:82-136 on one Gaussian. */
void orc_view_color_f32(const uint8_t color[4], const float sh[45], const float dir[3],
                        int32_t sh_deg, int32_t no_sh0, float out[4]);
void orc_view_color_f64(const uint8_t color[4], const double sh[45], const double dir[3],
                        int32_t sh_deg, int32_t no_sh0, double out[4]);

/* render.wesl:75-130 for instances idx[0..v). */
void orc_splats_f32(const orc_frame* f, uint32_t v, const uint32_t* idx, const float* pos,
                    const uint8_t* color, const float* sh, const float* cov3d,
                    orc_splat_f32* out);
void orc_splats_f64(const orc_frame* f, uint32_t v, const uint32_t* idx, const double* pos,
                    const uint8_t* color, const double* sh, const double* cov3d,
                    orc_splat_f64* out);

/* Rasterise + blend instances in draw order 0..v-1 (back to front) over the
 * clear colour (0,0,0,1). target selects the blend model. out is w*h*4 u8
 * (RGBA8 targets) or w*h*4 f32 (RGBA32F). threads <= 0 means 1. */
void orc_render_f32(const orc_frame* f, uint32_t w, uint32_t h, uint32_t v,
                    const orc_splat_f32* s, int32_t target, void* out, int32_t threads);
/* orc_render_f32 with a depth attachment (a pipeline given a depth_stencil,
 * src/renderer.rs:300-305): `depth` (w*h f32, row-major) is tested with the
 * wgpu::CompareFunction `cmp` (1 never .. 8 always) for every fragment that
 * survives the discard, in draw order, and written when `write`. */
void orc_render_depth_f32(const orc_frame* f, uint32_t w, uint32_t h, uint32_t v, const orc_splat_f32* s,
                          int32_t target, void* out, int32_t threads, float* depth, int32_t cmp, int32_t write);
void orc_render_f64(const orc_frame* f, uint32_t w, uint32_t h, uint32_t v,
                    const orc_splat_f64* s, double* out, int32_t threads);

/* radix_sort.wgsl semantics: stable ascending sort of (u32 key, u32 value). */
void orc_stable_sort(uint32_t n, uint32_t* keys, uint32_t* vals);

/* Whole frame (Viewer::render, src/lib.rs:266-275), f32, for timing and
 * end-to-end checks. Returns V; out_idx/out_key hold the sorted instances. */
uint32_t orc_frame_f32(const orc_frame* f, uint32_t n, const float* pos, const uint8_t* color,
                       const float* sh, const float* cov3d, uint32_t w, uint32_t h,
                       int32_t target, uint32_t* out_idx, float* out_key, void* out_image,
                       int32_t threads);

#ifdef __cplusplus
}
#endif
#endif
