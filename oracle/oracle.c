/*
 * CPU oracle for the per-frame splat path (TEST INFRASTRUCTURE ONLY; see
 * oracle.h for scope, citations and the floating-point convention).
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define R float
#define SUF f32
#define SQRT sqrtf
#define EXP expf
#define FMIN fminf
#define FMAX fmaxf
#include "oracle_impl.inc"
#undef R
#undef SUF
#undef SQRT
#undef EXP
#undef FMIN
#undef FMAX

#define R double
#define SUF f64
#define SQRT sqrt
#define EXP exp
#define FMIN fmin
#define FMAX fmax
#include "oracle_impl.inc"
#undef R
#undef SUF
#undef SQRT
#undef EXP
#undef FMIN
#undef FMAX

/* ------------------------------------------------------------------------ */
/* camera: src/camera.rs:71-92 (glam 0.30 look_to_rh, perspective_rh)        */
/* ------------------------------------------------------------------------ */

void orc_camera_forward(float pitch, float yaw, float out[3]) {
    /* src/camera.rs:71-77 */
    out[0] = cosf(pitch) * sinf(yaw);
    out[1] = sinf(pitch);
    out[2] = cosf(pitch) * cosf(yaw);
}

static void cross3(const float a[3], const float b[3], float o[3]) {
    o[0] = a[1] * b[2] - b[1] * a[2];
    o[1] = a[2] * b[0] - b[2] * a[0];
    o[2] = a[0] * b[1] - b[0] * a[1];
}

static void normalize3(const float a[3], float o[3]) {
    const float l = sqrtf((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]);
    const float inv = 1.0f / l;
    o[0] = a[0] * inv;
    o[1] = a[1] * inv;
    o[2] = a[2] * inv;
}

static float dot3f(const float a[3], const float b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

void orc_look_to_rh(const float eye[3], const float dir[3], const float up[3], float out[16]) {
    /* glam 0.30 takes dir as already normalised (it only asserts it) */
    const float* f = dir;
    float s0[3], s[3], u[3];
    cross3(f, up, s0);
    normalize3(s0, s);
    cross3(s, f, u);
    const float cols[16] = {s[0], u[0], -f[0], 0.0f, s[1], u[1], -f[1], 0.0f,
                            s[2], u[2], -f[2], 0.0f, -dot3f(eye, s), -dot3f(eye, u), dot3f(eye, f), 1.0f};
    memcpy(out, cols, sizeof cols);
}

void orc_perspective_rh(float fov_y, float aspect, float z_near, float z_far, float out[16]) {
    const float h = 0.5f * fov_y;
    const float c = cosf(h), sn = sinf(h);
    const float fh = c / sn;
    const float fw = fh / aspect;
    const float r = z_far / (z_near - z_far);
    const float cols[16] = {fw, 0, 0, 0, 0, fh, 0, 0, 0, 0, r, -1.0f, 0, 0, r * z_near, 0};
    memcpy(out, cols, sizeof cols);
}

/* ------------------------------------------------------------------------ */
/* wgpu-3dgs-core restatements (absent from the reference tree)             */
/* ------------------------------------------------------------------------ */

/* Rotation matrix of a quaternion (x,y,z,w), column-major, glam's
 * Mat3::from_quat expansion. */
#define QUAT_TO_MAT3(T, q, m)                                                              \
    do {                                                                                   \
        const T x2 = q[0] + q[0], y2 = q[1] + q[1], z2 = q[2] + q[2];                       \
        const T xx = q[0] * x2, xy = q[0] * y2, xz = q[0] * z2;                             \
        const T yy = q[1] * y2, yz = q[1] * z2, zz = q[2] * z2;                             \
        const T wx = q[3] * x2, wy = q[3] * y2, wz = q[3] * z2;                             \
        m[0] = (T)1 - (yy + zz); m[1] = xy + wz; m[2] = xz - wy;                            \
        m[3] = xy - wz; m[4] = (T)1 - (xx + zz); m[5] = yz + wx;                            \
        m[6] = xz + wy; m[7] = yz - wx; m[8] = (T)1 - (xx + yy);                            \
    } while (0)

/* cov3d = (R S)(R S)^T, stored xx, xy, xz, yy, yz, zz (the order utils.wesl:29-33
 * unpacks). */
#define COV3D(T, scale, rot, out)                                                          \
    do {                                                                                   \
        T rm[9], mm[9];                                                                    \
        QUAT_TO_MAT3(T, rot, rm);                                                          \
        for (int c_ = 0; c_ < 3; ++c_)                                                     \
            for (int r_ = 0; r_ < 3; ++r_) mm[c_ * 3 + r_] = rm[c_ * 3 + r_] * scale[c_];   \
        /* sigma[a][b] = sum_k M[k][a] * M[k][b] (rows a, b of M) */                       \
        const int ia[6] = {0, 0, 0, 1, 1, 2}, ib[6] = {0, 1, 2, 1, 2, 2};                   \
        for (int e_ = 0; e_ < 6; ++e_)                                                     \
            out[e_] = (mm[0 * 3 + ia[e_]] * mm[0 * 3 + ib[e_]] +                            \
                       mm[1 * 3 + ia[e_]] * mm[1 * 3 + ib[e_]]) +                           \
                      mm[2 * 3 + ia[e_]] * mm[2 * 3 + ib[e_]];                              \
    } while (0)

void orc_cov3d_f32(const float scale[3], const float rot[4], float out[6]) { COV3D(float, scale, rot, out); }
void orc_cov3d_f64(const double scale[3], const double rot[4], double out[6]) { COV3D(double, scale, rot, out); }
void orc_cov3d_f32_n(uint64_t n, const float* scale, const float* rot, float* out) {
    for (uint64_t i = 0; i < n; ++i) orc_cov3d_f32(scale + 3 * i, rot + 4 * i, out + 6 * i);
}
void orc_cov3d_f64_n(uint64_t n, const double* scale, const double* rot, double* out) {
    for (uint64_t i = 0; i < n; ++i) orc_cov3d_f64(scale + 3 * i, rot + 4 * i, out + 6 * i);
}

void orc_model_mats(const float pos[3], const float rot[4], const float scale[3], float model[16],
                    float sr[9], float inv_sr[9]) {
    float rm[9];
    QUAT_TO_MAT3(float, rot, rm);
    for (int c = 0; c < 3; ++c) {
        for (int r = 0; r < 3; ++r) {
            sr[c * 3 + r] = rm[c * 3 + r] * scale[c];
            model[c * 4 + r] = sr[c * 3 + r];
            /* inverse(R S) = S^-1 R^T */
            inv_sr[c * 3 + r] = rm[r * 3 + c] / scale[r];
        }
        model[c * 4 + 3] = 0.0f;
    }
    model[12] = pos[0];
    model[13] = pos[1];
    model[14] = pos[2];
    model[15] = 1.0f;
}

/* ------------------------------------------------------------------------ */
/* preprocess.wesl:60-106                                                   */
/* ------------------------------------------------------------------------ */

static int visible_one_f32(const orc_frame* f, const mats_f32* m, const float* pos, const float* cov3d,
                           float* key) {
    const float p4[4] = {pos[0], pos[1], pos[2], 1.0f};
    float world[4], clip[4];
    m4v_f32(m->model, p4, world);
    m4v_f32(m->pv, world, clip);
    const float nx = clip[0] / clip[3];
    const float ny = clip[1] / clip[3];
    const float nz = clip[2] / clip[3];
    if (cull_f32(nx, ny, nz)) {
        const float sd = f->max_std_dev;
        float ax[4];
        axes_m_f32(m, pos, cov3d, sd * f->gsize, ax);
        const float mx = (ax[0] * sd) / m->w;
        const float my = (ax[1] * sd) / m->h;
        const float mlen = sqrtf(mx * mx + my * my);
        const float nlen = sqrtf((-nx) * (-nx) + (-ny) * (-ny));
        const float dx = (-nx) / nlen;
        const float dy = (-ny) / nlen;
        const float t = fminf(mlen, nlen);
        const float bx = nx + t * dx;
        const float by = ny + t * dy;
        if (cull_f32(bx, by, nz)) return 0;
    }
    *key = 1.0f - nz;
    return 1;
}

uint32_t orc_preprocess_f32(const orc_frame* f, uint32_t n, const float* pos, const float* cov3d,
                            uint32_t* out_idx, float* out_key) {
    mats_f32 m;
    load_mats_f32(f, &m);
    uint32_t v = 0;
    for (uint32_t i = 0; i < n; ++i) {
        float key;
        if (visible_one_f32(f, &m, pos + 3 * (size_t)i, cov3d + 6 * (size_t)i, &key)) {
            out_idx[v] = i;
            out_key[v] = key;
            ++v;
        }
    }
    return v;
}

void orc_cov2d_f32(const orc_frame* f, const float pos[3], const float cov3d[6], float out[3]) {
    mats_f32 m;
    load_mats_f32(f, &m);
    cov2d_m_f32(&m, pos, cov3d, out);
}

/* ------------------------------------------------------------------------ */
/* radix_sort.wgsl semantics: stable ascending (key, value) sort             */
/* ------------------------------------------------------------------------ */

void orc_stable_sort(uint32_t n, uint32_t* keys, uint32_t* vals) {
    if (n < 2) return;
    uint32_t* kb = (uint32_t*)malloc((size_t)n * 4);
    uint32_t* vb = (uint32_t*)malloc((size_t)n * 4);
    uint32_t *ks = keys, *vs = vals, *kd = kb, *vd = vb;
    for (int pass = 0; pass < 4; ++pass) {
        const int shift = 8 * pass;
        size_t cnt[257];
        memset(cnt, 0, sizeof cnt);
        for (uint32_t i = 0; i < n; ++i) cnt[((ks[i] >> shift) & 0xFF) + 1]++;
        for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
        for (uint32_t i = 0; i < n; ++i) {
            const size_t p = cnt[(ks[i] >> shift) & 0xFF]++;
            kd[p] = ks[i];
            vd[p] = vs[i];
        }
        uint32_t* t = ks; ks = kd; kd = t;
        t = vs; vs = vd; vd = t;
    }
    /* four passes: result is back in keys/vals */
    free(kb);
    free(vb);
}

/* ------------------------------------------------------------------------ */
/* rasterise + blend (render.wesl:143-181, src/renderer.rs:169-181,296-300)  */
/* ------------------------------------------------------------------------ */

static uint8_t unorm8(float x) {
    x = fminf(fmaxf(x, 0.0f), 1.0f);
    return (uint8_t)(int)(x * 255.0f + 0.5f);
}

static float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

/* Pixel rectangle a splat may touch: its half extents plus a 2 px margin,
 * clipped to [y0,y1) rows of the band being rendered. */
static int splat_rect(const orc_splat_f32* s, uint32_t w, int32_t y0, int32_t y1, int32_t* rx0,
                      int32_t* rx1, int32_t* ry0, int32_t* ry1) {
    if (!s->drawable) return 0;
    const double ex = (double)s->hx + 2.0, ey = (double)s->hy + 2.0;
    double a = floor((double)s->cx - ex), b = ceil((double)s->cx + ex);
    double c = floor((double)s->cy - ey), d = ceil((double)s->cy + ey);
    if (!(a == a) || !(c == c) || !(b == b) || !(d == d)) return 0;
    if (a < 0) a = 0;
    if (b > (double)w) b = (double)w;
    if (c < y0) c = y0;
    if (d > y1) d = y1;
    if (a >= b || c >= d) return 0;
    *rx0 = (int32_t)a; *rx1 = (int32_t)b; *ry0 = (int32_t)c; *ry1 = (int32_t)d;
    return 1;
}

static int depth_pass(int32_t cmp, float z, float d) {
    switch (cmp) {
        case 1: return 0;
        case 2: return z < d;
        case 3: return z == d;
        case 4: return z <= d;
        case 5: return z > d;
        case 6: return z != d;
        case 7: return z >= d;
        default: return 1;
    }
}

static void render_f32(const orc_frame* f, uint32_t w, uint32_t h, uint32_t v, const orc_splat_f32* s,
                       int32_t target, void* out, int32_t threads, float* depth, int32_t cmp, int32_t write) {
    const int32_t band = 16;
    const int32_t nb = (int32_t)((h + band - 1) / band);
    if (threads <= 0) threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int32_t bi = 0; bi < nb; ++bi) {
        const int32_t y0 = bi * band;
        const int32_t y1 = (y0 + band < (int32_t)h) ? y0 + band : (int32_t)h;
        const size_t npx = (size_t)(y1 - y0) * w;
        float* acc = (float*)malloc(npx * 4 * sizeof(float));
        uint8_t* acc8 = (uint8_t*)malloc(npx * 4);
        for (size_t p = 0; p < npx; ++p) {
            acc[4 * p + 0] = acc[4 * p + 1] = acc[4 * p + 2] = 0.0f;
            acc[4 * p + 3] = 1.0f;
            acc8[4 * p + 0] = acc8[4 * p + 1] = acc8[4 * p + 2] = 0;
            acc8[4 * p + 3] = 255;
        }
        for (uint32_t k = 0; k < v; ++k) {
            const orc_splat_f32* sp = s + k;
            int32_t x0, x1, ry0, ry1;
            if (!splat_rect(sp, w, y0, y1, &x0, &x1, &ry0, &ry1)) continue;
            float src[3] = {sp->r, sp->g, sp->b};
            if (target != ORC_TARGET_RGBA32F) {
                src[0] = clamp01(src[0]);
                src[1] = clamp01(src[1]);
                src[2] = clamp01(src[2]);
            }
            for (int32_t y = ry0; y < ry1; ++y) {
                for (int32_t x = x0; x < x1; ++x) {
                    float al = frag_alpha_f32(f, sp, (float)x + 0.5f, (float)y + 0.5f);
                    if (al < 0.0f) continue;
                    if (depth) {
                        float* dz = depth + (size_t)y * w + (size_t)x;
                        if (!depth_pass(cmp, sp->z, *dz)) continue;
                        if (write) *dz = sp->z;
                    }
                    if (target != ORC_TARGET_RGBA32F) al = clamp01(al);
                    const float om = 1.0f - al;
                    const size_t p = (size_t)(y - y0) * w + (size_t)x;
                    if (target == ORC_TARGET_RGBA8_STEP) {
                        uint8_t* d = acc8 + 4 * p;
                        for (int ch = 0; ch < 3; ++ch)
                            d[ch] = unorm8(src[ch] * al + ((float)d[ch] / 255.0f) * om);
                        d[3] = unorm8(al + ((float)d[3] / 255.0f) * om);
                    } else {
                        float* d = acc + 4 * p;
                        for (int ch = 0; ch < 3; ++ch) d[ch] = src[ch] * al + d[ch] * om;
                        d[3] = al + d[3] * om;
                    }
                }
            }
        }
        const size_t base = (size_t)y0 * w;
        if (target == ORC_TARGET_RGBA32F) {
            memcpy((float*)out + 4 * base, acc, npx * 4 * sizeof(float));
        } else if (target == ORC_TARGET_RGBA8_F32) {
            uint8_t* o = (uint8_t*)out + 4 * base;
            for (size_t p = 0; p < 4 * npx; ++p) o[p] = unorm8(acc[p]);
        } else {
            memcpy((uint8_t*)out + 4 * base, acc8, npx * 4);
        }
        free(acc);
        free(acc8);
    }
}

void orc_render_f32(const orc_frame* f, uint32_t w, uint32_t h, uint32_t v, const orc_splat_f32* s,
                    int32_t target, void* out, int32_t threads) {
    render_f32(f, w, h, v, s, target, out, threads, NULL, 8, 0);
}

void orc_render_depth_f32(const orc_frame* f, uint32_t w, uint32_t h, uint32_t v, const orc_splat_f32* s,
                          int32_t target, void* out, int32_t threads, float* depth, int32_t cmp, int32_t write) {
    render_f32(f, w, h, v, s, target, out, threads, depth, cmp, write);
}

void orc_render_f64(const orc_frame* f, uint32_t w, uint32_t h, uint32_t v, const orc_splat_f64* s,
                    double* out, int32_t threads) {
    const int32_t band = 16;
    const int32_t nb = (int32_t)((h + band - 1) / band);
    if (threads <= 0) threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int32_t bi = 0; bi < nb; ++bi) {
        const int32_t y0 = bi * band;
        const int32_t y1 = (y0 + band < (int32_t)h) ? y0 + band : (int32_t)h;
        for (int32_t y = y0; y < y1; ++y)
            for (uint32_t x = 0; x < w; ++x) {
                double* d = out + 4 * ((size_t)y * w + x);
                d[0] = d[1] = d[2] = 0.0;
                d[3] = 1.0;
            }
        for (uint32_t k = 0; k < v; ++k) {
            const orc_splat_f64* sp = s + k;
            if (!sp->drawable) continue;
            const double ex = sp->hx + 2.0, ey = sp->hy + 2.0;
            double a = floor(sp->cx - ex), b = ceil(sp->cx + ex);
            double c = floor(sp->cy - ey), dd = ceil(sp->cy + ey);
            if (!(a == a) || !(b == b) || !(c == c) || !(dd == dd)) continue;
            if (a < 0) a = 0;
            if (b > (double)w) b = (double)w;
            if (c < y0) c = y0;
            if (dd > y1) dd = y1;
            for (int32_t y = (int32_t)c; y < (int32_t)dd; ++y)
                for (int32_t x = (int32_t)a; x < (int32_t)b; ++x) {
                    const double al = frag_alpha_f64(f, sp, (double)x + 0.5, (double)y + 0.5);
                    if (al < 0.0) continue;
                    const double om = 1.0 - al;
                    double* d = out + 4 * ((size_t)y * w + (size_t)x);
                    d[0] = sp->r * al + d[0] * om;
                    d[1] = sp->g * al + d[1] * om;
                    d[2] = sp->b * al + d[2] * om;
                    d[3] = al + d[3] * om;
                }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* whole frame: Viewer::render (src/lib.rs:266-275)                          */
/* ------------------------------------------------------------------------ */

uint32_t orc_frame_f32(const orc_frame* f, uint32_t n, const float* pos, const uint8_t* color,
                       const float* sh, const float* cov3d, uint32_t w, uint32_t h, int32_t target,
                       uint32_t* out_idx, float* out_key, void* out_image, int32_t threads) {
    mats_f32 m;
    load_mats_f32(f, &m);
    if (threads <= 0) threads = 1;
    /* preprocess: visibility in parallel, compaction in index order */
    uint8_t* vis = (uint8_t*)malloc(n ? n : 1);
    float* keys = (float*)malloc((size_t)(n ? n : 1) * sizeof(float));
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int64_t i = 0; i < (int64_t)n; ++i)
        vis[i] = (uint8_t)visible_one_f32(f, &m, pos + 3 * i, cov3d + 6 * i, &keys[i]);
    uint32_t v = 0;
    for (uint32_t i = 0; i < n; ++i)
        if (vis[i]) {
            out_idx[v] = i;
            out_key[v] = keys[i];
            ++v;
        }
    free(vis);
    free(keys);
    /* sort */
    orc_stable_sort(v, (uint32_t*)out_key, out_idx);
    /* vertex stage */
    orc_splat_f32* sp = (orc_splat_f32*)malloc((size_t)(v ? v : 1) * sizeof(orc_splat_f32));
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int64_t s = 0; s < (int64_t)v; ++s) {
        const size_t i = out_idx[s];
        splat_one_f32(f, &m, pos + 3 * i, color + 4 * i, sh + 45 * i, cov3d + 6 * i, sp + s);
    }
    orc_render_f32(f, w, h, v, sp, target, out_image, threads);
    free(sp);
    return v;
}
