/*
 * gs3d.h — C ABI of the MI355X-native per-frame splat pipeline
 * (libgs3d.so, built from wgpu-3dgs-viewer_amd/csrc).
 *
 * Drop-in boundary for wgpu-3dgs-viewer 0.6.1's Viewer / Preprocessor /
 * RadixSorter / Renderer (Rust, src/lib.rs, src/preprocessor.rs,
 * src/radix_sorter.rs, src/renderer.rs). Each entry point names the
 * reference interface it replaces. Plain pointers and sizes only; every call
 * returns a gs3d_status and gs3d_last_error() describes the last failure on
 * the calling thread.
 *
 * Execution model: a gs3d_device owns one HIP stream (the reference's
 * wgpu::Queue). Recording calls (render, preprocess, sort, draw) only enqueue
 * work, as the reference's calls only record into a CommandEncoder;
 * gs3d_device_poll() waits for it (queue.submit + device.poll(wait)). A
 * device, and everything created from it, must be used by one thread at a
 * time.
 */
#ifndef GS3D_H
#define GS3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status (src/error.rs:7-50) --------------------------------------- */
typedef enum gs3d_status {
    GS3D_OK = 0,
    GS3D_ERROR_INVALID_ARGUMENT = 1,
    /* RendererCreateError / PreprocessorCreateError::ModelSizeExceedsDeviceLimit
     * (src/error.rs:7-33; checks at src/preprocessor.rs:239-246,
     * src/renderer.rs:130-135) */
    GS3D_ERROR_MODEL_SIZE_EXCEEDS_DEVICE_LIMIT = 2,
    GS3D_ERROR_OUT_OF_MEMORY = 3,
    GS3D_ERROR_DEVICE = 4,
    /* MultiModelViewerAccessError::ModelNotFound (src/error.rs:47-50) */
    GS3D_ERROR_MODEL_NOT_FOUND = 5,
    /* a kernel reported an internal error (look-back wait limit exceeded, or
     * a tile pair with a tile id outside the grid); no counterpart in the
     * reference, whose recording is infallible */
    GS3D_ERROR_KERNEL = 6
} gs3d_status;

/* ---- records ------------------------------------------------------------ */

/* wgpu_3dgs_core::Gaussian (used at tests/e2e/viewer.rs:42-48): rotation
 * quaternion (x, y, z, w), position, RGBA8 colour (rgb = 0.5 + SH_C0 * dc,
 * a = opacity), 15 SH coefficients per channel (degree 1..3), scale. */
typedef struct gs3d_gaussian {
    float rot[4];
    float pos[3];
    uint8_t color[4];
    float sh[15][3];
    float scale[3];
} gs3d_gaussian;

/* CameraPod (src/buffer/camera.rs:61-79): column-major view and projection,
 * target size in pixels; 144 bytes. */
typedef struct gs3d_camera_pod {
    float view[16];
    float proj[16];
    float size[2];
    uint32_t padding[2];
} gs3d_camera_pod;

/* Camera (src/camera.rs:18-29). */
typedef struct gs3d_camera {
    float pos[3];
    float z_start;
    float z_end;
    float vertical_fov;
    float pitch;
    float yaw;
} gs3d_camera;

/* GaussianPod choice (src/lib.rs:44). */
typedef enum gs3d_gaussian_pod {
    /* SH and cov3d in f32 (the reference's DefaultGaussianPod) */
    GS3D_POD_SINGLE = 0,
    /* SH, rotation and scale in f16, opacity u8; cov3d built on the GPU */
    GS3D_POD_HALF = 1
} gs3d_gaussian_pod;

/* wgpu::TextureFormat of the render target (src/renderer.rs:120-125). */
typedef enum gs3d_texture_format {
    GS3D_FORMAT_RGBA8_UNORM = 0,
    GS3D_FORMAT_RGBA32_FLOAT = 1,
    /* a depth attachment (wgpu::TextureFormat::Depth32Float), one f32 per
     * pixel; usable only as the depth texture of a draw */
    GS3D_FORMAT_DEPTH32_FLOAT = 2
} gs3d_texture_format;

/* wgpu::CompareFunction (same numbering) of a depth test. */
typedef enum gs3d_compare {
    GS3D_COMPARE_NEVER = 1,
    GS3D_COMPARE_LESS = 2,
    GS3D_COMPARE_EQUAL = 3,
    GS3D_COMPARE_LESS_EQUAL = 4,
    GS3D_COMPARE_GREATER = 5,
    GS3D_COMPARE_NOT_EQUAL = 6,
    GS3D_COMPARE_GREATER_EQUAL = 7,
    GS3D_COMPARE_ALWAYS = 8
} gs3d_compare;

/* GaussianDisplayMode (render.wesl:12-14). */
typedef enum gs3d_display_mode {
    GS3D_DISPLAY_SPLAT = 0,
    GS3D_DISPLAY_ELLIPSE = 1,
    GS3D_DISPLAY_POINT = 2
} gs3d_display_mode;

/* How the compositor blends (the reference uses fixed-function
 * ALPHA_BLENDING, src/renderer.rs:296-300). EXACT (the default): back to
 * front, destination rounded to the target format after every splat, as the
 * blender does. FAST: front to back in f32 with early termination, rounded
 * once (on an Rgba8Unorm target it differs from the reference's image where
 * many faint splats overlap: up to a few /255). */
typedef enum gs3d_blend_mode {
    GS3D_BLEND_FAST = 0,
    GS3D_BLEND_EXACT = 1
} gs3d_blend_mode;

/* Per-stage GPU times of the last rendered frame (gs3d_viewer_stage_times). */
typedef enum gs3d_stage {
    GS3D_STAGE_PREPROCESS = 0,
    GS3D_STAGE_SORT = 1,
    GS3D_STAGE_BIN = 2,
    GS3D_STAGE_TILE_SORT = 3,
    GS3D_STAGE_COMPOSITE = 4,
    /* the main compositing kernel alone, an interval inside COMPOSITE */
    GS3D_STAGE_COMPOSITE_MAIN = 5,
    /* the projection kernel alone, an interval inside PREPROCESS */
    GS3D_STAGE_PREPROCESS_MAIN = 6,
    GS3D_STAGE_COUNT = 7
} gs3d_stage;

typedef struct gs3d_device gs3d_device;
typedef struct gs3d_texture gs3d_texture;
typedef struct gs3d_viewer gs3d_viewer;
typedef struct gs3d_radix_sorter gs3d_radix_sorter;

/* ---- errors ------------------------------------------------------------- */

/* Message of the last failed call on this thread ("" if none). */
const char* gs3d_last_error(void);

/* ---- device and queue (wgpu::Device + wgpu::Queue) ---------------------- */

gs3d_status gs3d_device_create(int hip_device, gs3d_device** out);
void gs3d_device_destroy(gs3d_device* dev);
/* queue.submit + device.poll(wait) (tests/e2e/viewer.rs:31-34): waits for all
 * enqueued work. A frame whose tile lists overflowed is re-rendered with
 * larger lists before this returns. */
gs3d_status gs3d_device_poll(gs3d_device* dev);
/* The largest single buffer a model may need, the counterpart of
 * max_storage_buffer_binding_size (src/preprocessor.rs:239). Defaults to the
 * device's HBM size. */
gs3d_status gs3d_device_set_buffer_limit(gs3d_device* dev, uint64_t bytes);
gs3d_status gs3d_device_compute_units(gs3d_device* dev, int* out);

/* ---- render targets (wgpu::Texture / TextureView) ----------------------- */

gs3d_status gs3d_texture_create(gs3d_device* dev, uint32_t width, uint32_t height, gs3d_texture_format format,
                                gs3d_texture** out);
void gs3d_texture_destroy(gs3d_texture* tex);
/* Copies the texture to host memory (row-major, top row first, 4 channels);
 * waits for the queue first. */
gs3d_status gs3d_texture_read(gs3d_texture* tex, void* dst, size_t bytes);
/* queue.write_texture: replaces a texture's contents (bytes = its size, f32
 * per pixel for a depth texture); waits for the device first. */
gs3d_status gs3d_texture_write(gs3d_texture* tex, const void* src, size_t bytes);

/* ---- Viewer (src/lib.rs:65-276) ------------------------------------------ */

/* Viewer::new (src/lib.rs:86-97): uploads the Gaussians in the chosen pod
 * and creates every buffer, the preprocessor, the sorter and the renderer. */
gs3d_status gs3d_viewer_new(gs3d_device* dev, gs3d_texture_format format, const gs3d_gaussian* gaussians,
                            uint32_t count, gs3d_gaussian_pod pod, gs3d_viewer** out);
void gs3d_viewer_destroy(gs3d_viewer* v);
/* Viewer::update_camera (src/lib.rs:202-209). */
gs3d_status gs3d_viewer_update_camera(gs3d_viewer* v, const gs3d_camera* camera, uint32_t width, uint32_t height);
/* Viewer::update_camera_with_pod (src/lib.rs:212-214). */
gs3d_status gs3d_viewer_update_camera_with_pod(gs3d_viewer* v, const gs3d_camera_pod* pod);
/* Viewer::update_model_transform (src/lib.rs:217-225). */
gs3d_status gs3d_viewer_update_model_transform(gs3d_viewer* v, const float pos[3], const float rot[4],
                                               const float scale[3]);
/* Viewer::update_gaussian_transform (src/lib.rs:237-254). */
gs3d_status gs3d_viewer_update_gaussian_transform(gs3d_viewer* v, float size, gs3d_display_mode display_mode,
                                                  uint32_t sh_deg, int no_sh0, float max_std_dev);
/* Selection early-out of the preprocessor (src/shader/preprocess.wesl:69-78,
 * feature "viewer-selection"; buffers src/lib.rs:74-77): the SelectionBuffer
 * contents, bit i of word i / 32 for Gaussian i, exactly ceil(N / 32) words.
 * Gaussian i is skipped when its bit equals the invert flag. Until the first
 * call the bitset is all zero. */
gs3d_status gs3d_viewer_update_selection(gs3d_viewer* v, const uint32_t* words, uint32_t word_count);
/* PreprocessorInvertSelectionBuffer::update (src/selection/buffer.rs:167-171);
 * the initial value is true (selected Gaussians are hidden). */
gs3d_status gs3d_viewer_update_invert_selection(gs3d_viewer* v, int invert);
/* Viewer::render (src/lib.rs:266-275): preprocess -> sort -> render. */
gs3d_status gs3d_viewer_render(gs3d_viewer* v, gs3d_texture* target);
/* ViewerCreateOptions::depth_stencil (src/lib.rs:279-290, passed to
 * Renderer::new, src/renderer.rs:120-137 and 300-305): enabled != 0 gives the
 * renderer a depth test with `compare` and `depth_write`; its frames are then
 * drawn with gs3d_viewer_render_with_depth only. A fragment's depth is the
 * clip z / w of its splat's centre. */
gs3d_status gs3d_viewer_set_depth_stencil(gs3d_viewer* v, int enabled, gs3d_compare compare, int depth_write);
/* Viewer::render drawn through Renderer::render_with_pass into a caller's
 * pass that has a depth attachment (src/renderer.rs:186-195): `depth` is a
 * GS3D_FORMAT_DEPTH32_FLOAT texture of the target's size; load != 0 keeps the
 * target's colour (LoadOp::Load), otherwise it is cleared to black. */
gs3d_status gs3d_viewer_render_with_depth(gs3d_viewer* v, gs3d_texture* target, gs3d_texture* depth, int load);
/* Preprocessor::preprocess (src/preprocessor.rs:283-290). */
gs3d_status gs3d_viewer_preprocess(gs3d_viewer* v);
/* RadixSorter::sort (src/radix_sorter.rs:56-66). */
gs3d_status gs3d_viewer_sort(gs3d_viewer* v);
/* Renderer::render (src/renderer.rs:163-184), using the splats the fused
 * preprocessor computed. */
gs3d_status gs3d_viewer_draw(gs3d_viewer* v, gs3d_texture* target);
/* Renderer::render driven as a standalone operator: recomputes every drawn
 * instance's splat from the Gaussians and the sorted indices (the vertex
 * stage, render.wesl:75-130) before binning and compositing. */
gs3d_status gs3d_viewer_draw_standalone(gs3d_viewer* v, gs3d_texture* target);

gs3d_status gs3d_viewer_set_blend_mode(gs3d_viewer* v, gs3d_blend_mode mode);
/* No counterpart in the reference (its rasterizer needs no lists): the
 * renderer's (tile, splat) pair lists start at max(2^20, 2N) pairs and grow
 * when a frame needs more; a draw that overflows them draws nothing, and
 * gs3d_device_poll grows the lists and records every draw since the last
 * poll again (an error when one loaded its target's contents or wrote its
 * depth attachment). This replaces the lists with ones of `pairs` pairs
 * (1 .. 2^30 - 1); it polls the device first. */
gs3d_status gs3d_viewer_set_pair_capacity(gs3d_viewer* v, uint32_t pairs);
/* Compositing tile size: 8x8, 16x16 (default) or 32x8 pixels. */
gs3d_status gs3d_viewer_set_tile_size(gs3d_viewer* v, uint32_t tile_w, uint32_t tile_h);
/* No counterpart in the reference (a tuning choice of this sort's kernels):
 * keys per thread of the depth and tile sorts' tiles, 8, 12, 16, 20 (default)
 * or 24.
 * Every choice gives the same sorted order. */
gs3d_status gs3d_viewer_set_sort_items(gs3d_viewer* v, uint32_t items);
/* Record HIP events around each stage of every render (off by default). */
gs3d_status gs3d_viewer_set_profiling(gs3d_viewer* v, int enabled);
/* Milliseconds of each gs3d_stage in the last profiled frame; waits. */
gs3d_status gs3d_viewer_stage_times(gs3d_viewer* v, float ms[GS3D_STAGE_COUNT]);
/* Renders one frame (preprocess, sort, exact-blend compositing) counting the
 * fragments that pass the discard test of render.wesl:143-166 (the
 * reference's per-fragment work, for roofline accounting); waits. */
gs3d_status gs3d_viewer_count_fragments(gs3d_viewer* v, gs3d_texture* target, uint64_t* out);

/* Buffer downloads (the Viewer's pub buffers, src/lib.rs:66-73); each waits
 * for the queue first. */
/* (tile, splat) pairs in the tile lists: the largest count of the frames the
 * viewer drew before this call, which waits for them (no reference
 * counterpart: the reference rasterises instead of binning). */
gs3d_status gs3d_viewer_pair_count(gs3d_viewer* v, uint32_t* out);
gs3d_status gs3d_viewer_gaussian_count(gs3d_viewer* v, uint32_t* out);
/* IndirectArgsBuffer {vertex_count 6, instance_count, 0, 0}
 * (src/buffer/indirect_args.rs:12-26). */
gs3d_status gs3d_viewer_read_indirect_args(gs3d_viewer* v, uint32_t out[4]);
/* RadixSortIndirectArgsBuffer {x, y, z} (src/buffer/indirect_args.rs:63-71). */
gs3d_status gs3d_viewer_read_radix_sort_indirect_args(gs3d_viewer* v, uint32_t out[3]);
/* IndirectIndicesBuffer, first `count` entries. */
gs3d_status gs3d_viewer_read_indirect_indices(gs3d_viewer* v, uint32_t* out, uint32_t count);
/* GaussiansDepthBuffer as f32 keys, first `count` entries (the buffer holds
 * gs3d_depth_buffer_len(N) keys). */
gs3d_status gs3d_viewer_read_gaussians_depth(gs3d_viewer* v, float* out, uint32_t count);
/* Number of keys in a GaussiansDepthBuffer for N Gaussians: N rounded up to
 * the reference's sort block of 3840 (src/radix_sorter.rs:922-939). */
uint32_t gs3d_depth_buffer_len(uint32_t gaussian_count);

/* ---- RadixSorter as a standalone operator (src/radix_sorter.rs:9-96) ---- */

/* RadixSorter::new with its own key/payload buffers of `capacity` entries. */
gs3d_status gs3d_radix_sorter_new(gs3d_device* dev, uint32_t capacity, gs3d_radix_sorter** out);
void gs3d_radix_sorter_destroy(gs3d_radix_sorter* s);
/* Uploads n (key, value) pairs. */
gs3d_status gs3d_radix_sorter_write(gs3d_radix_sorter* s, const uint32_t* keys, const uint32_t* vals, uint32_t n);
/* RadixSorter::sort: stable ascending sort of the n pairs by key. */
gs3d_status gs3d_radix_sorter_sort(gs3d_radix_sorter* s);
/* No counterpart in the reference: keys per thread of the sort's tiles, 8, 12,
 * 16, 20 (default) or 24; every choice gives the same order. */
gs3d_status gs3d_radix_sorter_set_items(gs3d_radix_sorter* s, uint32_t items);
gs3d_status gs3d_radix_sorter_read(gs3d_radix_sorter* s, uint32_t* keys, uint32_t* vals, uint32_t n);

/* Most pairs gs3d_radix_sort_noalloc sorts in one call. */
#define GS3D_RADIX_SORT_NOALLOC_MAX_KEYS 4194304u
/* No counterpart in the reference. An allocation-free variant of
 * gs3d_radix_sorter_sort for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): sorts the n (key, value) pairs at device addresses keys_in /
 * vals_in into keys_out / vals_out, stable and ascending by key (the same
 * kernels and result as gs3d_radix_sorter_sort), on the device's queue 0;
 * asynchronous, like every launch (gs3d_device_poll waits and reports kernel
 * errors). The inputs are left unchanged; keys_in may equal keys_out and
 * vals_in vals_out (an in-place sort). It allocates nothing: its ping-pong
 * pairs, histograms and look-back words live in one fixed device array per GPU
 * (41.9 MB, held until the process ends), so n above
 * GS3D_RADIX_SORT_NOALLOC_MAX_KEYS is GS3D_ERROR_INVALID_ARGUMENT. Every call
 * on one GPU shares that array: calls that overlap on one GPU (from several
 * threads, or on streams the caller orders differently) race. */
gs3d_status gs3d_radix_sort_noalloc(gs3d_device* dev, const uint32_t* keys_in, const uint32_t* vals_in,
                                    uint32_t* keys_out, uint32_t* vals_out, uint32_t n);

/* ---- resident models ------------------------------------------------------ */

/* A GaussiansBuffer (core's GaussiansBuffer<G>) kept in HBM and shared by
 * batched views and served requests. */
typedef struct gs3d_model gs3d_model;
gs3d_status gs3d_model_new(gs3d_device* dev, const gs3d_gaussian* gaussians, uint32_t count, gs3d_gaussian_pod pod,
                           gs3d_model** out);
void gs3d_model_destroy(gs3d_model* m);

/* ---- MultiModelViewer (src/multi_model.rs:289-531) ------------------------ */

typedef struct gs3d_multi_model_viewer gs3d_multi_model_viewer;
/* MultiModelViewer::new (src/multi_model.rs:307-317). */
gs3d_status gs3d_multi_model_viewer_new(gs3d_device* dev, gs3d_texture_format format, gs3d_multi_model_viewer** out);
void gs3d_multi_model_viewer_destroy(gs3d_multi_model_viewer* v);
/* MultiModelViewer::insert_model (src/multi_model.rs:353-391). */
gs3d_status gs3d_multi_model_viewer_insert_model(gs3d_multi_model_viewer* v, const char* key,
                                                 const gs3d_gaussian* gaussians, uint32_t count,
                                                 gs3d_gaussian_pod pod);
/* MultiModelViewer::remove_model (src/multi_model.rs:393-396);
 * GS3D_ERROR_MODEL_NOT_FOUND for an unknown key. */
gs3d_status gs3d_multi_model_viewer_remove_model(gs3d_multi_model_viewer* v, const char* key);
/* MultiModelViewer::update_camera / _with_pod (src/multi_model.rs:398-412). */
gs3d_status gs3d_multi_model_viewer_update_camera(gs3d_multi_model_viewer* v, const gs3d_camera* camera,
                                                  uint32_t width, uint32_t height);
gs3d_status gs3d_multi_model_viewer_update_camera_with_pod(gs3d_multi_model_viewer* v, const gs3d_camera_pod* pod);
/* MultiModelViewer::update_model_transform (src/multi_model.rs:414-429);
 * GS3D_ERROR_MODEL_NOT_FOUND for an unknown key. */
gs3d_status gs3d_multi_model_viewer_update_model_transform(gs3d_multi_model_viewer* v, const char* key,
                                                           const float pos[3], const float rot[4],
                                                           const float scale[3]);
/* The selection buffers of one model (MultiModelViewerGaussianBuffers,
 * src/multi_model.rs:81-84), as gs3d_viewer_update_selection and
 * gs3d_viewer_update_invert_selection; ModelNotFound for an unknown key. */
gs3d_status gs3d_multi_model_viewer_update_selection(gs3d_multi_model_viewer* v, const char* key,
                                                     const uint32_t* words, uint32_t word_count);
gs3d_status gs3d_multi_model_viewer_update_invert_selection(gs3d_multi_model_viewer* v, const char* key,
                                                            int invert);
/* MultiModelViewer::update_gaussian_transform (src/multi_model.rs:446-464). */
gs3d_status gs3d_multi_model_viewer_update_gaussian_transform(gs3d_multi_model_viewer* v, float size,
                                                              gs3d_display_mode display_mode, uint32_t sh_deg,
                                                              int no_sh0, float max_std_dev);
gs3d_status gs3d_multi_model_viewer_set_blend_mode(gs3d_multi_model_viewer* v, gs3d_blend_mode mode);
/* MultiModelViewer::render (src/multi_model.rs:476-530): the models named by
 * keys[0..count), drawn in that order into one pass cleared to black. */
gs3d_status gs3d_multi_model_viewer_render(gs3d_multi_model_viewer* v, gs3d_texture* target,
                                           const char* const* keys, uint32_t count);

/* ---- batched views (no counterpart in the reference: the batched form) --- */

typedef struct gs3d_view_spec {
    uint32_t model;  /* index into the batch's model list */
    uint32_t width;
    uint32_t height;
} gs3d_view_spec;

typedef struct gs3d_batch gs3d_batch;
/* A batch renderer over resident models, spreading its views over `queues`
 * device queues. */
gs3d_status gs3d_batch_new(gs3d_device* dev, gs3d_model* const* models, uint32_t model_count, int queues,
                           gs3d_batch** out);
void gs3d_batch_destroy(gs3d_batch* b);
/* One view slot (viewer + RGBA8 target) per view. */
gs3d_status gs3d_batch_set_layout(gs3d_batch* b, const gs3d_view_spec* views, uint32_t count);
/* Picks, for every view, the fastest compositing tile (8x8, 16x16, 32x8),
 * timing `reps` draws of each with the given camera pods, then the fastest
 * sort tile (8, 12, 16, 20 or 24 keys per thread, the same for every view),
 * timing `reps` whole batches of each (the views run concurrently). */
gs3d_status gs3d_batch_autotune(gs3d_batch* b, const gs3d_camera_pod* pods, int reps);
/* Renders every view with its camera pod. use_graph: the first call records
 * the whole batch as one HIP graph, later calls replay it. Enqueue only. */
gs3d_status gs3d_batch_render(gs3d_batch* b, const gs3d_camera_pod* pods, int use_graph);
/* Waits for the batch (re-rendering views whose tile lists overflowed). */
gs3d_status gs3d_batch_finish(gs3d_batch* b);
gs3d_status gs3d_batch_read_target(gs3d_batch* b, uint32_t view, void* dst, size_t bytes);
gs3d_status gs3d_batch_tile_size(gs3d_batch* b, uint32_t view, uint32_t* tile_w, uint32_t* tile_h);
/* Keys per thread of the view's sort tiles (what gs3d_batch_autotune picked). */
gs3d_status gs3d_batch_sort_items(gs3d_batch* b, uint32_t view, uint32_t* items);

/* ---- serving (no counterpart in the reference: the serving form) --------- */

typedef struct gs3d_request {
    uint32_t model;
    uint32_t width;
    uint32_t height;
    uint32_t padding;
    double arrival_s;  /* seconds from the start of the run */
    gs3d_camera_pod pod;
} gs3d_request;

typedef struct gs3d_serve_stats {
    uint64_t requests;
    uint64_t completed;
    double seconds;       /* first arrival to last completion */
    double frames_per_s;
    double p50_ms;        /* latency: arrival to completion */
    double p99_ms;
    double mean_ms;
    double max_ms;
    double mean_in_flight;
    uint64_t peak_hbm_bytes;
    uint64_t deferred_for_memory;
} gs3d_serve_stats;

typedef struct gs3d_server gs3d_server;
/* A continuous-batching server over resident models: up to 2 requests per
 * queue in flight, per-request viewers and targets drawn from pools bounded
 * by hbm_budget bytes. */
gs3d_status gs3d_server_new(gs3d_device* dev, gs3d_model* const* models, uint32_t model_count, int queues,
                            uint64_t hbm_budget, gs3d_server** out);
void gs3d_server_destroy(gs3d_server* s);
/* Serves requests[0..count) (sorted by arrival) in real time. images may be
 * null; otherwise images[i], when non-null, receives request i's RGBA8 frame. */
gs3d_status gs3d_server_run(gs3d_server* s, const gs3d_request* requests, uint32_t count, uint8_t* const* images,
                            gs3d_serve_stats* out);
/* One request with nothing else in flight; its RGBA8 frame into out. */
gs3d_status gs3d_server_render_alone(gs3d_server* s, const gs3d_request* request, uint8_t* out);

/* ---- host helpers -------------------------------------------------------- */

/* CameraPod::new (src/buffer/camera.rs:72-79) from a Camera. */
gs3d_status gs3d_camera_pod_new(const gs3d_camera* camera, uint32_t width, uint32_t height, gs3d_camera_pod* out);
/* A camera pod looking from `eye` at `target` (+Y up). */
gs3d_status gs3d_camera_pod_look_at(const float eye[3], const float target[3], float vertical_fov, float z_near,
                                    float z_far, uint32_t width, uint32_t height, gs3d_camera_pod* out);
/* cov3d of a Gaussian as the f32 pod stores it: (R S)(R S)^T as
 * xx, xy, xz, yy, yz, zz. */
gs3d_status gs3d_gaussian_cov3d(const gs3d_gaussian* g, float out[6]);
/* The f16 pod's rotation/scale (8 halves per Gaussian: rot xyzw, scale xyz,
 * 0) and SH (48 halves: 45 coefficients, 3 zeros) as uploaded. */
gs3d_status gs3d_pack_half(const gs3d_gaussian* g, uint32_t count, uint16_t* rot_scale, uint16_t* sh);

/* PLY ingestion, replacing wgpu_3dgs_core::Gaussians::read_from_file with
 * GaussiansSource::Ply (examples/simple.rs:157-160, examples/multi_model.rs:
 * 131-134): the standard 3DGS vertex layout (x y z, f_dc_0..2, f_rest_*,
 * opacity, scale_0..2, rot_0..3; other properties ignored), binary
 * little-endian or ASCII. With out == NULL, *count receives the number of
 * Gaussians; otherwise capacity must be at least that and *count receives the
 * number written. rot_0 is w; scale and opacity are stored as log and logit;
 * rgb = 0.5 + SH_C0 * f_dc; f_rest is channel-major. */
gs3d_status gs3d_ply_read_file(const char* path, gs3d_gaussian* out, uint64_t capacity, uint64_t* count);
gs3d_status gs3d_ply_read_memory(const void* data, uint64_t bytes, gs3d_gaussian* out, uint64_t capacity,
                                 uint64_t* count);

/* SPZ ingestion, replacing Gaussians::read_from_file with GaussiansSource::Spz
 * (examples/simple.rs:157-160): Niantic's gzip-compressed layout, versions 2
 * and 3, same contract as gs3d_ply_read_file. No coordinate conversion. */
gs3d_status gs3d_spz_read_file(const char* path, gs3d_gaussian* out, uint64_t capacity, uint64_t* count);
gs3d_status gs3d_spz_read_memory(const void* data, uint64_t bytes, gs3d_gaussian* out, uint64_t capacity,
                                 uint64_t* count);

/* Synthetic scenes. config 0: means U[-1,1]^3; 1: means U[-4,4]^3 (both
 * log-scale N(-4,0.5)); 2: 64 anisotropic clusters + 20% on the plane
 * y = -2, log-scale N(-5,0.7). Counter-based RNG: record i depends only on
 * (seed, i). */
gs3d_status gs3d_synth_gaussians(int config, uint64_t seed, uint32_t first, uint32_t count, gs3d_gaussian* out);

#ifdef __cplusplus
}
#endif
#endif /* GS3D_H */
