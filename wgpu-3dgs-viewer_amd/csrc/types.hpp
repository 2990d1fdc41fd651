// Device-visible record layouts shared by the HIP kernels and the host library.
//
// Everything here is plain data: the C ABI in include/gs3d.h exposes only the
// host-side records (gs3d_gaussian, gs3d_camera_pod, ...); these are the
// HBM layouts behind it.
#pragma once

#include <cstdint>

namespace gs3d {

// wgpu_sort::HISTO_BLOCK_KVS (src/radix_sorter.rs:175-178): the reference's
// sort dispatch granularity. Kept only to reproduce the indirect-dispatch
// arguments and the depth-buffer padding the reference writes
// (preprocess.wesl:113-125, radix_sort.wgsl:92-110).
constexpr uint32_t kHistoBlockKvs = 3840;

// f32 bits of the padding key the reference writes after the visible keys
// (preprocess.wesl:124, `2.0`).
constexpr uint32_t kPadKeyBits = 0x40000000u;

enum PodKind : int32_t {
  kPodSingle = 0,  // f32 SH + f32 cov3d (reference default pod, src/lib.rs:44)
  kPodHalf = 1,    // f16 SH + f16 rotation/scale, cov3d built on the GPU
};

enum DisplayMode : int32_t { kModeSplat = 0, kModeEllipse = 1, kModePoint = 2 };

enum TextureFormat : int32_t {
  kFormatRgba8Unorm = 0,
  kFormatRgba32Float = 1,
  kFormatDepth32Float = 2,  // depth attachment: one f32 per pixel
};

// wgpu::CompareFunction of a depth test (same numbering).
enum CompareFn : int32_t {
  kCmpNever = 1,
  kCmpLess = 2,
  kCmpEqual = 3,
  kCmpLessEqual = 4,
  kCmpGreater = 5,
  kCmpNotEqual = 6,
  kCmpGreaterEqual = 7,
  kCmpAlways = 8,
};

enum BlendMode : int32_t {
  // front-to-back in f32 with early termination, one quantisation at the end
  kBlendFast = 0,
  // (the default) back-to-front, src clamped, destination stored as unorm8
  // after every splat: the fixed-function ALPHA_BLENDING into an Rgba8Unorm
  // target
  kBlendExactUnorm8 = 1,
};

// CameraPod (src/buffer/camera.rs:61-79): 144 bytes.
struct CameraPod {
  float view[16];
  float proj[16];
  float size[2];
  uint32_t padding[2];
};
static_assert(sizeof(CameraPod) == 144, "CameraPod is 144 bytes");

// Model transform uniform: position, rotation quaternion (x,y,z,w), scale.
struct ModelTransformPod {
  float pos[3];
  float pad0;
  float rot[4];
  float scale[3];
  float pad1;
};
static_assert(sizeof(ModelTransformPod) == 48, "ModelTransformPod is 48 bytes");

// Gaussian transform uniform: size, display mode / SH degree / no_sh0 flags,
// max standard deviation.
struct GaussianTransformPod {
  float size;
  uint32_t flags;  // bits 0-1 display mode, 2-3 sh degree, 4 no_sh0
  float max_std_dev;
  uint32_t pad;
};
static_assert(sizeof(GaussianTransformPod) == 16, "GaussianTransformPod is 16 bytes");

inline uint32_t pack_gaussian_flags(int32_t mode, uint32_t sh_deg, bool no_sh0) {
  return (uint32_t(mode) & 3u) | ((sh_deg & 3u) << 2) | (no_sh0 ? 16u : 0u);
}

// Per-frame constants derived on the GPU from the three uniform buffers at the
// start of every frame (the reference recomputes them in every invocation).
struct FrameConsts {
  float view[16];
  float proj[16];
  float pv[16];     // proj * view           (camera.wesl:14)
  float vm[16];     // view * model          (utils.wesl:37)
  float model[16];  // model_transform_mat
  float sr[9];      // model_scale_rot_mat
  float inv_sr[9];  // model_transform_inv_sr_mat
  float wsr[9];     // mat3(view) * model_scale_rot_mat
  float campos[3];  // render.wesl:59-63
  float size[2];
  float focal[2];   // utils.wesl:35
  float gsize;
  float max_std_dev;
  int32_t display_mode;
  int32_t sh_deg;
  int32_t no_sh0;
};

#if defined(__HIPCC__)
// The frame constants as kernels read them: in the constant address space.
// An earlier launch writes them and this one only reads them, so the
// compiler may use scalar loads (their own counter, no vector registers)
// anywhere in the kernel, also after its stores and atomics; read through a
// generic pointer, they would be vector loads that count on the same in-order
// counter as the Gaussian loads requested ahead, so every wait for a matrix
// would also wait for those.
using ConstFrame = __attribute__((address_space(4))) const FrameConsts;
__host__ __device__ inline const ConstFrame& const_frame(const FrameConsts* p) { return *(ConstFrame*)p; }
#endif

// Per visible Gaussian, the 2D splat the compositor draws (the pixel-space
// restatement of what vert_main hands the rasterizer, render.wesl:75-130).
// Indexed by Gaussian index; 32 bytes (one aligned 32-byte HBM sector).
//   splat / ellipse: |qo|^2 = (d.u)^2 + (d.v)^2 for d = pixel centre -
//                    (cx, cy), where qo is the interpolated quad offset of
//                    render.wesl:116-121: u = (ux, uy) = 2 (M.x, -M.y) /
//                    |M|^2 for the major axis M, and v = rho (-uy, ux) with
//                    rho = |M| / |m| (v is the minor axis' term, u rotated
//                    by 90 degrees and scaled). A sum of squares, as the
//                    shader's, so it never goes negative for thin splats.
//   point          : ux, uy hold the half extents in pixels (f32)
// rg, ba: colour and opacity as f16 pairs; hxy: f16 pair of the half extents
// of the pixel bound, rounded up, 0 when the splat draws no pixel.
struct alignas(16) Splat2D {
  float cx, cy;
  float ux, uy, rho;
  uint32_t rg;
  uint32_t ba;
  uint32_t hxy;
};
static_assert(sizeof(Splat2D) == 32, "Splat2D is 32 bytes");

// Look-back status words: 2-bit flag + 30-bit value.
constexpr uint32_t kFlagAggregate = 1u << 30;
constexpr uint32_t kFlagInclusive = 2u << 30;
constexpr uint32_t kFlagMask = 3u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1u;

// Error bits the kernels raise in the per-viewer status word.
constexpr uint32_t kErrLookbackSpin = 1u;
constexpr uint32_t kErrTileOverflow = 2u;
constexpr uint32_t kErrTileKey = 4u;  // a tile pair with a tile id outside the grid
constexpr uint32_t kErrSplatIndex = 8u;  // a sorted list held a Gaussian index outside the model

// Screen tiles of a render target.
struct TileGrid {
  uint32_t width, height;   // target pixels
  uint32_t tile_w, tile_h;  // tile pixels (powers of two)
  uint32_t tiles_x, tiles_y;
};

inline bool same_grid(const TileGrid& a, const TileGrid& b) {
  return a.width == b.width && a.height == b.height && a.tile_w == b.tile_w && a.tile_h == b.tile_h;
}

}  // namespace gs3d
