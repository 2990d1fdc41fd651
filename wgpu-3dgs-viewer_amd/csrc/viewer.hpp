// C++ host mirror of the reference's Viewer / Preprocessor / RadixSorter /
// Renderer and their buffers (src/lib.rs, src/preprocessor.rs,
// src/radix_sorter.rs, src/renderer.rs, src/buffer/*). Type names and call
// shapes follow the reference; wgpu handles become a Device (HIP device +
// stream) and device buffers. The C ABI in include/gs3d.h wraps these.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gs3d.h"
#include "kernels.hpp"
#include "types.hpp"

// splats per segment of a long tile list in the fast blend (split-K)
#ifndef GS3D_COMPOSITE_SEG_LEN
#define GS3D_COMPOSITE_SEG_LEN 2048
#endif

namespace gs3d {

// Radix passes of the (tile, splat) pair sort the renderer's scratch holds:
// tile ids of the largest grid (16384 x 16384 pixels in 8x8 tiles, 2^22
// tiles) take 22 bits.
constexpr int kTileSortPasses = 3;

// Error carrying a gs3d_status (the reference's error enums, src/error.rs).
struct Error : std::runtime_error {
  gs3d_status code;
  Error(gs3d_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what);

class Renderer;
class Texture;

// A unit of recorded work that Device::poll records again when one of its
// draws overflowed its tile lists (such a draw draws nothing): a whole
// Viewer::render with the uniforms and settings it was recorded with, one
// MultiModelViewer pass, one stage-level draw. It is dropped when a texture
// or renderer it uses is destroyed.
struct RecordedUnit {
  std::function<void()> replay;
  std::vector<Renderer*> renderers;
  std::vector<const Texture*> textures;
  // why the unit cannot be recorded again as it was (null: it can), asked
  // when a replay is needed
  std::function<const char*()> blocker;
};

// wgpu::Device + wgpu::Queue: one HIP device and one stream.
// Raises the error a kernel reported through the device's error word (the
// look-back wait limit, a tile id outside the grid); overflow is not one.
void throw_kernel_errors(uint32_t e);

class Device {
 public:
  explicit Device(int ordinal);
  ~Device();
  Device(const Device&) = delete;
  Device& operator=(const Device&) = delete;

  // Queue q's stream; queue 0 always exists (all at the default priority).
  hipStream_t stream(int q = 0) const { return streams_.at(size_t(q)); }
  // Adds a queue (a new stream) and returns its index.
  int add_queue();
  int num_queues() const { return int(streams_.size()); }
  int ordinal() const { return ordinal_; }
  int num_cus() const { return num_cus_; }
  uint64_t buffer_limit() const { return buffer_limit_; }
  void set_buffer_limit(uint64_t b) { buffer_limit_ = b; }
  // Device-side error word every kernel may OR bits into.
  uint32_t* err() const { return err_; }
  void make_current() const;

  // queue.submit + device.poll(wait). When a draw since the last poll
  // overflowed its tile lists, grows them and records every unit since the
  // last poll again, in order; throws on kernel-reported errors, and when a
  // unit that needs recording again cannot be.
  void poll();
  void wait();
  // Reads and clears the kernels' error bits; the caller has waited.
  uint32_t take_errors();
  // Keeps a unit of work until the next poll.
  void record_unit(RecordedUnit u);
  // Drops the recorded units that use a renderer / texture being destroyed.
  void forget_renderer(const Renderer* r);
  void forget_texture(const Texture* t);
  // Bytes held by DeviceBuffers on this device.
  uint64_t allocated() const { return allocated_; }
  void track(int64_t delta) { allocated_ = uint64_t(int64_t(allocated_) + delta); }

 private:
  int ordinal_ = 0;
  int num_cus_ = 0;
  uint64_t buffer_limit_ = 0;
  std::vector<hipStream_t> streams_;
  uint32_t* err_ = nullptr;
  uint64_t allocated_ = 0;
  std::vector<RecordedUnit> units_;
};

// Owning device allocation.
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  DeviceBuffer(Device& dev, size_t bytes);
  ~DeviceBuffer();
  DeviceBuffer(DeviceBuffer&& o) noexcept;
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;

  void* get() const { return ptr_; }
  template <typename T>
  T* as() const {
    return static_cast<T*>(ptr_);
  }
  size_t size() const { return bytes_; }
  void upload(const void* src, size_t bytes, size_t offset = 0);
  void download(void* dst, size_t bytes, size_t offset = 0) const;

 private:
  Device* dev_ = nullptr;
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
};

// Small uniform buffer written from the host in stream order (the
// reference's queue.write_buffer): the bytes are copied into pinned staging
// memory and from there to the device on the stream, so the caller's memory
// may change as soon as update() returns.
class UniformBuffer {
 public:
  UniformBuffer(Device& dev, size_t bytes, int queue = 0);
  ~UniformBuffer();
  UniformBuffer(const UniformBuffer&) = delete;
  UniformBuffer& operator=(const UniformBuffer&) = delete;
  void update(const void* src);
  void* dev() const { return buf_.get(); }

 private:
  Device& device_;
  hipStream_t st_;
  DeviceBuffer buf_;
  size_t bytes_;
  void* staging_ = nullptr;
  hipEvent_t done_ = nullptr;
};

// ---- host camera (src/camera.rs) -------------------------------------------

struct Camera {
  float pos[3] = {0, 0, 0};
  float z_start = 0.1f, z_end = 1e4f;
  float vertical_fov = 1.0471976f;
  float pitch = 0, yaw = 0;
  void forward(float out[3]) const;
  void view(float out[16]) const;
  void projection(float aspect, float out[16]) const;
};

void look_to_rh(const float eye[3], const float dir[3], const float up[3], float out[16]);
void perspective_rh(float fov_y, float aspect, float z_near, float z_far, float out[16]);
CameraPod camera_pod_new(const Camera& c, uint32_t w, uint32_t h);  // CameraPod::new

// ---- buffers (src/buffer/*) --------------------------------------------------

class CameraBuffer {
 public:
  explicit CameraBuffer(Device& dev, int queue = 0);
  void update(const Camera& c, uint32_t w, uint32_t h);
  void update_with_pod(const CameraPod& pod);
  const CameraPod* dev() const { return static_cast<const CameraPod*>(buf_.dev()); }
  const CameraPod& host() const { return pod_; }

 private:
  UniformBuffer buf_;
  CameraPod pod_{};
};

class ModelTransformBuffer {
 public:
  explicit ModelTransformBuffer(Device& dev, int queue = 0);
  void update(const float pos[3], const float rot[4], const float scale[3]);
  void update_with_pod(const ModelTransformPod& pod);
  const ModelTransformPod* dev() const { return static_cast<const ModelTransformPod*>(buf_.dev()); }
  const ModelTransformPod& host() const { return pod_; }

 private:
  UniformBuffer buf_;
  ModelTransformPod pod_{};
};

class GaussianTransformBuffer {
 public:
  explicit GaussianTransformBuffer(Device& dev, int queue = 0);
  void update(float size, int32_t mode, uint32_t sh_deg, bool no_sh0, float max_std_dev);
  void update_with_pod(const GaussianTransformPod& pod);
  const GaussianTransformPod* dev() const { return static_cast<const GaussianTransformPod*>(buf_.dev()); }
  const GaussianTransformPod& host() const { return pod_; }

 private:
  UniformBuffer buf_;
  GaussianTransformPod pod_{};
};

// GaussiansBuffer<G>: the packed model in HBM (structure of streams).
class GaussiansBuffer {
 public:
  GaussiansBuffer(Device& dev, const gs3d_gaussian* g, uint32_t n, PodKind pod);
  uint32_t len() const { return n_; }
  PodKind pod() const { return pod_; }
  uint64_t size_bytes() const;
  GaussianStreams streams() const;
  static uint64_t bytes_per_gaussian(PodKind pod);

 private:
  PodKind pod_;
  uint32_t n_;
  DeviceBuffer pc_, cov_, rs_, sh_;
};

class IndirectArgsBuffer {
 public:
  explicit IndirectArgsBuffer(Device& dev);
  uint32_t* dev() const { return buf_.as<uint32_t>(); }
  void read(uint32_t out[4]) const { buf_.download(out, 16); }

 private:
  DeviceBuffer buf_;
};

class RadixSortIndirectArgsBuffer {
 public:
  explicit RadixSortIndirectArgsBuffer(Device& dev);
  uint32_t* dev() const { return buf_.as<uint32_t>(); }
  void read(uint32_t out[3]) const { buf_.download(out, 12); }

 private:
  DeviceBuffer buf_;
};

class IndirectIndicesBuffer {
 public:
  IndirectIndicesBuffer(Device& dev, uint32_t n);
  uint32_t* dev() const { return buf_.as<uint32_t>(); }
  uint32_t len() const { return n_; }
  void read(uint32_t* out, uint32_t count) const { buf_.download(out, size_t(count) * 4); }

 private:
  uint32_t n_;
  DeviceBuffer buf_;
};

uint32_t depth_buffer_len(uint32_t n);

class GaussiansDepthBuffer {
 public:
  GaussiansDepthBuffer(Device& dev, uint32_t n);
  float* dev() const { return buf_.as<float>(); }
  uint32_t len() const { return len_; }
  void read(float* out, uint32_t count) const { buf_.download(out, size_t(count) * 4); }

 private:
  uint32_t len_;
  DeviceBuffer buf_;
};

// Per-Gaussian 2D splat records (no counterpart in the reference: its vertex
// shader recomputes them for each of the 6 vertices of every instance).
class SplatsBuffer {
 public:
  SplatsBuffer(Device& dev, uint32_t n);
  Splat2D* dev() const { return buf_.as<Splat2D>(); }
  uint32_t len() const { return n_; }  // Gaussians (records, depths, rectangles)
  // Per-Gaussian fragment depth, kept only for depth-tested draws (null
  // until enable_depth).
  float* depth() const { return depth_.as<float>(); }
  void enable_depth();
  // Per-Gaussian tile rectangles on one grid, written with the splats so
  // that binning reads 8 bytes per instance instead of the whole splat.
  // Set by whoever knows the next frame's target and tile size.
  void set_rect_grid(const TileGrid& g);
  bool rects_on() const { return rects_.size() != 0; }
  const TileGrid& rect_grid() const { return rect_grid_; }
  uint2* rects() const { return rects_.as<uint2>(); }

 private:
  Device& dev_;
  uint32_t n_;
  DeviceBuffer buf_, depth_, rects_;
  TileGrid rect_grid_{};
};

// Render target (wgpu::Texture).
class Texture {
 public:
  Texture(Device& dev, uint32_t w, uint32_t h, TextureFormat fmt);
  ~Texture();
  Texture(const Texture&) = delete;
  Texture& operator=(const Texture&) = delete;
  uint32_t width() const { return w_; }
  uint32_t height() const { return h_; }
  TextureFormat format() const { return fmt_; }
  void* dev() const { return buf_.get(); }
  size_t bytes() const { return buf_.size(); }
  void read(void* dst, size_t bytes) const;
  // queue.write_texture: replaces the contents (waits for the device first).
  void write(const void* src, size_t bytes);

 private:
  Device* device_;
  uint32_t w_, h_;
  TextureFormat fmt_;
  DeviceBuffer buf_;
};

// ---- operators -------------------------------------------------------------

// Stage timing with HIP events on the device stream.
class StageTimer {
 public:
  explicit StageTimer(Device& dev, int queue = 0);
  ~StageTimer();
  void set_enabled(bool on) { enabled_ = on; }
  bool enabled() const { return enabled_; }
  void mark(int stage, bool end);
  // The event pair of a stage for a launcher to record itself (null when
  // timing is off); marks the stage as recorded.
  hipEvent_t event(int stage, bool end);
  void read(float ms[GS3D_STAGE_COUNT]);

 private:
  Device* dev_;
  hipStream_t st_;
  bool enabled_ = false;
  hipEvent_t ev_[GS3D_STAGE_COUNT][2];
  bool used_[GS3D_STAGE_COUNT];
};

// Preprocessor (src/preprocessor.rs): per-frame constants ("pre"), the fused
// projection ("main" + "post" + the vertex stage's per-instance work).
class Preprocessor {
 public:
  Preprocessor(Device& dev, const CameraBuffer& camera, const ModelTransformBuffer& model,
               const GaussianTransformBuffer& gt, const GaussiansBuffer& gaussians, IndirectArgsBuffer& indirect_args,
               RadixSortIndirectArgsBuffer& rs_args, IndirectIndicesBuffer& indices, GaussiansDepthBuffer& depth,
               SplatsBuffer& splats, int queue = 0);
  void preprocess(uint32_t gaussian_count);
  // preprocess() in two steps: the per-frame constants and resets, then the
  // projection kernel (timed alone by the viewer's profiling)
  void frame_pre(uint32_t gaussian_count);
  void project(uint32_t gaussian_count);
  const FrameConsts* frame_consts() const { return consts_.as<FrameConsts>(); }
  // Selection early-out (preprocess.wesl:69-78): `words` is a device bitset
  // of the Gaussians (null: none); a Gaussian is skipped when its bit equals
  // `invert`.
  void set_selection(const uint32_t* words) { selection_ = words; }
  // Where the fused vertex work writes each splat's fragment depth (null:
  // nowhere), for depth-tested draws.
  void set_splat_depth(float* p) { splat_depth_ = p; }
  // Histogram of the depth sort's first pass, written with the keys
  // (`sort_max`: the sorter's capacity; null: not written).
  void set_sort_hist(uint32_t* hist, uint32_t sort_max) {
    sort_hist_ = hist;
    sort_max_ = sort_max;
  }
  void set_invert_selection(bool invert) { invert_selection_ = invert; }
  // Counts the frames this preprocessor has recorded (a draw recorded after
  // frame k draws frame k's splats).
  uint64_t generation() const { return generation_; }

 private:
  Device& dev_;
  hipStream_t st_;
  const CameraBuffer& camera_;
  const ModelTransformBuffer& model_;
  const GaussianTransformBuffer& gt_;
  const GaussiansBuffer& gaussians_;
  IndirectArgsBuffer& indirect_args_;
  RadixSortIndirectArgsBuffer& rs_args_;
  IndirectIndicesBuffer& indices_;
  GaussiansDepthBuffer& depth_;
  SplatsBuffer& splats_;
  DeviceBuffer consts_, status_, ticket_, thin_idx_;
  uint32_t max_tiles_;
  const uint32_t* selection_ = nullptr;
  float* splat_depth_ = nullptr;
  uint32_t* sort_hist_ = nullptr;
  uint32_t sort_max_ = 0;
  bool invert_selection_ = true;  // PreprocessorInvertSelectionBuffer::new: initial value true
  uint64_t generation_ = 0;
};

// Reusable sort scratch for up to `capacity` pairs and `passes` passes.
class SortScratchBuffers {
 public:
  SortScratchBuffers(Device& dev, uint32_t capacity, int passes);
  SortScratch get() const;
  uint32_t capacity() const { return capacity_; }
  // Keys per thread of the onesweep tiles the sorts with this scratch use:
  // 8, 12, 16, 20 (the default) or 24; every choice gives the same order.
  void set_items(uint32_t items);
  uint32_t items() const { return items_; }
  // Replaces the lists with ones of exactly `pairs` pairs (they still grow
  // when a frame needs more); waits for the device.
  void set_capacity(uint32_t pairs);

 private:
  Device& dev_;
  uint32_t capacity_;
  int passes_;
  uint32_t items_ = 20;
  DeviceBuffer hist_, lookback_, tickets_;
};

// RadixSorter (src/radix_sorter.rs:9-96): stable ascending (depth key,
// index) sort in place, count from the RadixSortIndirectArgsBuffer.
class RadixSorter {
 public:
  RadixSorter(Device& dev, GaussiansDepthBuffer& depth, IndirectIndicesBuffer& indices, int queue = 0);
  // hist_ready: the preprocessor that wrote the keys also wrote the first
  // pass's histogram into hist() (Preprocessor::set_sort_hist).
  void sort(const RadixSortIndirectArgsBuffer& rs_args, bool hist_ready = false);
  // Sort an explicit device count (count = *count_dev, at most len).
  void sort_count(const uint32_t* count_dev, uint32_t count_mul, bool hist_ready = false);
  uint32_t* hist() const { return scratch_.get().hist; }
  uint32_t len() const { return indices_.len(); }
  // Keys per thread of the sort's onesweep tiles (8, 12, 16, 20 or 24).
  void set_items(uint32_t items) { scratch_.set_items(items); }
  uint32_t items() const { return scratch_.items(); }

 private:
  Device& dev_;
  hipStream_t st_;
  GaussiansDepthBuffer& depth_;
  IndirectIndicesBuffer& indices_;
  DeviceBuffer keys_b_, vals_b_;
  SortScratchBuffers scratch_;
};

// Renderer (src/renderer.rs): tile binning + compositing into a Texture.
class Renderer {
 public:
  Renderer(Device& dev, TextureFormat format, const CameraBuffer& camera, const ModelTransformBuffer& model,
           const GaussianTransformBuffer& gt, const GaussiansBuffer& gaussians, const IndirectIndicesBuffer& indices,
           SplatsBuffer& splats, const FrameConsts* shared_consts, int queue = 0);
  ~Renderer();
  // Renderer::render; `standalone` recomputes the splats (vertex stage).
  // clear: LoadOp::Clear(BLACK) (src/renderer.rs:175); otherwise the splats
  // are blended over what the target holds (LoadOp::Load, as the draws of
  // several models share one pass, src/multi_model.rs:505-527).
  // depth: the depth attachment (Depth32Float) of a draw; required exactly
  // when the renderer was given a depth-stencil state.
  void render(Texture& target, const IndirectArgsBuffer& indirect_args, bool standalone, bool clear = true,
              Texture* depth = nullptr);
  // Renderer::new's depth_stencil (src/renderer.rs:120-137, 300-305): the
  // depth compare function and depth write of the pipeline, or none.
  void set_depth_stencil(bool enabled, CompareFn cmp, bool write);
  bool depth_enabled() const { return depth_enabled_; }
  bool depth_write() const { return depth_enabled_ && depth_write_; }
  // Most (tile, splat) pairs a frame's lists may hold (30-bit status values).
  static constexpr uint32_t kMaxPairs = 0x3FFFFFFFu;
  // The tile grid of a frame drawn into `target`.
  TileGrid grid_for(const Texture& target) const;
  void set_seg_len(uint32_t s) { seg_len_ = s; }
  // Renders once in exact mode counting the fragments that are not
  // discarded (the reference's per-fragment work); waits.
  uint64_t count_fragments(Texture& target, const IndirectArgsBuffer& indirect_args);
  uint32_t tile_w() const { return tile_w_; }
  uint32_t tile_h() const { return tile_h_; }
  void set_blend(BlendMode b) { blend_ = b; }
  BlendMode blend() const { return blend_; }
  void set_tile_size(uint32_t w, uint32_t h);
  // Keys per thread of the tile sort's onesweep tiles (8, 12, 16, 20 or 24).
  void set_sort_items(uint32_t items);
  uint32_t sort_items() const { return sort_items_; }
  StageTimer& timer() { return timer_; }
  void set_timer(StageTimer* t) { ext_timer_ = t; }
  // True if a draw since the last call overflowed the tile lists: they are
  // grown to hold it, and the draw must be recorded again. Throws when no
  // capacity can hold it. The draws have completed.
  bool grow_if_overflowed();
  // Records the last draw again (for owners whose every renderer draws once
  // between waits: the batched and served views).
  void rerender();
  // rerender() until the last draw fits its lists (at most 3 times more),
  // waiting for each; the draws so far have completed.
  void settle();
  // (tile, splat) pairs of the last completed frame, and the list capacity
  uint32_t last_pair_total() const { return *static_cast<volatile uint32_t*>(pair_total_host_); }
  // the largest pair total of the frames drawn before the last poll
  uint32_t checked_pair_total() const { return checked_pair_total_; }
  uint32_t capacity() const { return capacity_; }
  // Replaces the lists with ones of exactly `pairs` pairs (they still grow
  // when a frame needs more); waits for the device.
  void set_capacity(uint32_t pairs);

 private:
  void ensure_capacity(uint32_t pairs);
  StageTimer& active_timer() { return ext_timer_ ? *ext_timer_ : timer_; }

  Device& dev_;
  hipStream_t st_;
  TextureFormat format_;
  const CameraBuffer& camera_;
  const ModelTransformBuffer& model_;
  const GaussianTransformBuffer& gt_;
  const GaussiansBuffer& gaussians_;
  const IndirectIndicesBuffer& indices_;
  SplatsBuffer& splats_;
  const FrameConsts* shared_consts_;
  DeviceBuffer own_consts_;
  BlendMode blend_ = kBlendExactUnorm8;  // the reference's blend
  uint32_t tile_w_ = 16, tile_h_ = 16;
  uint32_t sort_items_ = 20;
  uint32_t capacity_ = 0;
  DeviceBuffer pk_[2], pv_[2];
  std::unique_ptr<SortScratchBuffers> scratch_;
  DeviceBuffer ranges_, bin_status_, bin_ticket_, pair_count_;
  uint32_t ranges_words_ = 0;
  // compositing work list (fast blend)
  uint32_t seg_len_ = GS3D_COMPOSITE_SEG_LEN;
  DeviceBuffer item_tile_, item_seg_, seg_base_, plan_, plan_counts_, partial_;
  uint32_t* pair_total_host_ = nullptr;  // pinned, mapped
  uint32_t checked_pair_total_ = 0;
  uint32_t* pair_total_dev_ = nullptr;   // its device alias
  DeviceBuffer frag_count_;
  bool counting_ = false;
  StageTimer timer_;
  StageTimer* ext_timer_ = nullptr;
  // last frame, for re-rendering after an overflow
  Texture* last_target_ = nullptr;
  const IndirectArgsBuffer* last_args_ = nullptr;
  bool last_standalone_ = false;
  bool last_clear_ = true;
  Texture* last_depth_ = nullptr;
  bool depth_enabled_ = false;
  CompareFn depth_cmp_ = kCmpLess;
  bool depth_write_ = false;
};

// Viewer (src/lib.rs:65-276).
class Viewer {
 public:
  // Everything a recorded frame depends on besides the model: the uniforms
  // and the renderer's settings.
  struct State {
    CameraPod camera;
    ModelTransformPod model;
    GaussianTransformPod gaussian;
    BlendMode blend;
    uint32_t tile_w, tile_h;
  };

  Viewer(Device& dev, TextureFormat format, const gs3d_gaussian* g, uint32_t n, PodKind pod);
  // A viewer over a resident model shared with other viewers, recording on
  // the device's queue `queue`.
  Viewer(Device& dev, TextureFormat format, std::shared_ptr<GaussiansBuffer> model, int queue = 0);
  void update_camera(const Camera& c, uint32_t w, uint32_t h) { camera_buffer.update(c, w, h); }
  void update_camera_with_pod(const CameraPod& pod) { camera_buffer.update_with_pod(pod); }
  void update_model_transform(const float pos[3], const float rot[4], const float scale[3]) {
    model_transform_buffer.update(pos, rot, scale);
  }
  void update_gaussian_transform(float size, int32_t mode, uint32_t sh_deg, bool no_sh0, float max_std_dev) {
    gaussian_transform_buffer.update(size, mode, sh_deg, no_sh0, max_std_dev);
  }
  // record: keep the frame as a unit Device::poll may record again (owners
  // that handle overflow themselves pass false).
  void render(Texture& target, bool record = true);
  State state() const;
  // Writes the uniforms (in stream order) and settings of a recorded state.
  void restore(const State& s);
  // SelectionBuffer contents (one bit per Gaussian, ceil(N / 32) words) and
  // PreprocessorInvertSelectionBuffer::update (src/selection/buffer.rs:
  // 151-172). The bitset starts all zero and invert true: nothing is hidden.
  void update_selection(const uint32_t* words, uint32_t count);
  void update_invert_selection(bool invert) { preprocessor.set_invert_selection(invert); }
  // Keys per thread of the depth and tile sorts' onesweep tiles: 8, 12, 16,
  // 20 (the default) or 24. A tuning choice only: every choice sorts the same way.
  void set_sort_items(uint32_t items) {
    radix_sorter.set_items(items);
    renderer.set_sort_items(items);
  }
  uint32_t sort_items() const { return radix_sorter.items(); }
  // Viewer::new_with_options(depth_stencil) (src/lib.rs:100-199, 279-284):
  // the renderer's depth test, and a frame drawn with a depth attachment
  // (Renderer::render_with_pass in a pass that has one, src/renderer.rs:
  // 186-195); load keeps the target's colour (LoadOp::Load).
  void set_depth_stencil(bool enabled, CompareFn cmp, bool write);
  void render_with_depth(Texture& target, Texture& depth, bool load);
  // The stage-level draw (Renderer::render after separate preprocess and
  // sort calls), kept as a unit like render().
  void draw(Texture& target, bool standalone);

  Device& device;
  int queue;
  std::shared_ptr<GaussiansBuffer> model;
  CameraBuffer camera_buffer;
  ModelTransformBuffer model_transform_buffer;
  GaussianTransformBuffer gaussian_transform_buffer;
  GaussiansBuffer& gaussians_buffer;
  IndirectArgsBuffer indirect_args_buffer;
  RadixSortIndirectArgsBuffer radix_sort_indirect_args_buffer;
  IndirectIndicesBuffer indirect_indices_buffer;
  GaussiansDepthBuffer gaussians_depth_buffer;
  SplatsBuffer splats_buffer;
  Preprocessor preprocessor;
  RadixSorter radix_sorter;
  Renderer renderer;
  StageTimer timer;
  DeviceBuffer selection_buffer;  // allocated by the first update_selection
};

// ---- host-side packing helpers ---------------------------------------------

uint16_t float_to_half(float f);  // round to nearest even
void gaussian_cov3d(const gs3d_gaussian& g, float out[6]);
uint32_t pack_color(const uint8_t c[4]);

// PLY ingestion (ply.cpp). With out == nullptr only the vertex count is
// returned; otherwise capacity must hold every vertex.
uint64_t ply_read(const uint8_t* data, size_t size, gs3d_gaussian* out, uint64_t capacity);
uint64_t ply_read_file(const char* path, gs3d_gaussian* out, uint64_t capacity);
// SPZ ingestion (spz.cpp), same contract.
uint64_t spz_read(const uint8_t* data, size_t size, gs3d_gaussian* out, uint64_t capacity);
uint64_t spz_read_file(const char* path, gs3d_gaussian* out, uint64_t capacity);

// Synthetic scenes (synth.cpp).
void synth_gaussians(int config, uint64_t seed, uint32_t first, uint32_t count, gs3d_gaussian* out);

}  // namespace gs3d
