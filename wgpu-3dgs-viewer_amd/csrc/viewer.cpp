// Host side of the C++ mirror (see viewer.hpp).
#include "viewer.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "math.hpp"

namespace gs3d {

void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  const gs3d_status code = (e == hipErrorOutOfMemory) ? GS3D_ERROR_OUT_OF_MEMORY : GS3D_ERROR_DEVICE;
  throw Error(code, std::string(what) + ": " + hipGetErrorString(e));
}

// ---- Device -----------------------------------------------------------------

Device::Device(int ordinal) : ordinal_(ordinal) {
  int count = 0;
  hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
  if (ordinal < 0 || ordinal >= count) throw Error(GS3D_ERROR_INVALID_ARGUMENT, "no such HIP device");
  hip_check(hipSetDevice(ordinal), "hipSetDevice");
  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, ordinal), "hipGetDeviceProperties");
  num_cus_ = prop.multiProcessorCount;
  buffer_limit_ = prop.totalGlobalMem;
  add_queue();
  hip_check(hipMalloc(&err_, 256), "hipMalloc(err)");
  hip_check(hipMemset(err_, 0, 256), "hipMemset(err)");
  // the allocation-free sort's fixed array: its address on this device
  hip_check(noalloc_sort_init(ordinal), "hipGetSymbolAddress(scratch of the allocation-free sort)");
}

Device::~Device() {
  for (hipStream_t s : streams_) {
    hipStreamSynchronize(s);
    hipStreamDestroy(s);
  }
  if (err_) hipFree(err_);
}

int Device::add_queue() {
  make_current();
  hipStream_t s = nullptr;
  hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  streams_.push_back(s);
  return int(streams_.size()) - 1;
}

void Device::make_current() const { hip_check(hipSetDevice(ordinal_), "hipSetDevice"); }

void Device::wait() {
  for (hipStream_t s : streams_) hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

uint32_t Device::take_errors() {
  uint32_t e = 0;
  hip_check(hipMemcpy(&e, err_, 4, hipMemcpyDeviceToHost), "read error word");
  if (e) hip_check(hipMemset(err_, 0, 4), "clear error word");
  return e;
}

void Device::record_unit(RecordedUnit u) { units_.push_back(std::move(u)); }

void Device::forget_renderer(const Renderer* r) {
  units_.erase(std::remove_if(units_.begin(), units_.end(),
                              [r](const RecordedUnit& u) {
                                return std::find(u.renderers.begin(), u.renderers.end(), r) != u.renderers.end();
                              }),
               units_.end());
}

void Device::forget_texture(const Texture* t) {
  units_.erase(std::remove_if(units_.begin(), units_.end(),
                              [t](const RecordedUnit& u) {
                                return std::find(u.textures.begin(), u.textures.end(), t) != u.textures.end();
                              }),
               units_.end());
}

void throw_kernel_errors(uint32_t e) {
  if (e & kErrLookbackSpin) throw Error(GS3D_ERROR_KERNEL, "a kernel exceeded its look-back wait limit");
  if (e & kErrTileKey) throw Error(GS3D_ERROR_KERNEL, "a tile pair held a tile id outside the grid");
  if (e & kErrSplatIndex) throw Error(GS3D_ERROR_KERNEL, "a sorted list held a Gaussian index outside the model");
}

void Device::poll() {
  make_current();
  for (int attempt = 0;; ++attempt) {
    wait();
    // the overflow bit is not needed here: every renderer keeps its own
    // largest pair total since it was last asked
    const uint32_t e = take_errors();
    if (e & (kErrLookbackSpin | kErrTileKey | kErrSplatIndex)) {
      units_.clear();
      throw_kernel_errors(e);
    }
    std::vector<Renderer*> asked;
    bool grown = false;
    try {
      for (const RecordedUnit& u : units_)
        for (Renderer* r : u.renderers)
          if (std::find(asked.begin(), asked.end(), r) == asked.end()) {
            asked.push_back(r);
            grown = r->grow_if_overflowed() || grown;
          }
    } catch (...) {
      units_.clear();
      throw;
    }
    if (!grown) break;
    // A draw that overflowed drew nothing, and later draws into its target
    // blended over what was there: every unit since the last poll is
    // recorded again, in its order, with the state it was recorded with.
    if (attempt >= 3) {
      units_.clear();
      throw Error(GS3D_ERROR_KERNEL, "tile lists still overflowed after growing them three times");
    }
    for (const RecordedUnit& u : units_) {
      const char* why = u.blocker ? u.blocker() : nullptr;
      if (why) {
        units_.clear();
        throw Error(GS3D_ERROR_KERNEL, std::string("a draw overflowed its tile lists and cannot be recorded again: ") +
                                           why + " (render it again, or draw it alone)");
      }
    }
    std::vector<RecordedUnit> again;
    again.swap(units_);
    for (RecordedUnit& u : again) u.replay();  // records each unit again
  }
  units_.clear();
}

// ---- DeviceBuffer -------------------------------------------------------------

DeviceBuffer::DeviceBuffer(Device& dev, size_t bytes) : dev_(&dev), bytes_(bytes) {
  dev.make_current();
  const size_t alloc = bytes ? bytes : 4;
  hip_check(hipMalloc(&ptr_, alloc), "hipMalloc");
  dev.track(int64_t(alloc));
  hip_check(hipMemsetAsync(ptr_, 0, alloc, dev.stream()), "hipMemsetAsync");
  hip_check(hipStreamSynchronize(dev.stream()), "hipStreamSynchronize");
}

DeviceBuffer::~DeviceBuffer() {
  if (ptr_) {
    hipFree(ptr_);
    dev_->track(-int64_t(bytes_ ? bytes_ : 4));
  }
}

DeviceBuffer::DeviceBuffer(DeviceBuffer&& o) noexcept : dev_(o.dev_), ptr_(o.ptr_), bytes_(o.bytes_) {
  o.ptr_ = nullptr;
  o.bytes_ = 0;
}

DeviceBuffer& DeviceBuffer::operator=(DeviceBuffer&& o) noexcept {
  if (this != &o) {
    if (ptr_) {
      hipFree(ptr_);
      dev_->track(-int64_t(bytes_ ? bytes_ : 4));
    }
    dev_ = o.dev_;
    ptr_ = o.ptr_;
    bytes_ = o.bytes_;
    o.ptr_ = nullptr;
    o.bytes_ = 0;
  }
  return *this;
}

void DeviceBuffer::upload(const void* src, size_t bytes, size_t offset) {
  if (offset + bytes > bytes_) throw Error(GS3D_ERROR_INVALID_ARGUMENT, "upload past the end of a buffer");
  if (bytes == 0) return;
  // stream-ordered like queue.write_buffer; pageable sources are staged
  // before the call returns
  hip_check(hipMemcpyAsync(static_cast<char*>(ptr_) + offset, src, bytes, hipMemcpyHostToDevice, dev_->stream()),
            "hipMemcpyAsync(H2D)");
}

void DeviceBuffer::download(void* dst, size_t bytes, size_t offset) const {
  if (offset + bytes > bytes_) throw Error(GS3D_ERROR_INVALID_ARGUMENT, "download past the end of a buffer");
  if (bytes == 0) return;
  hip_check(hipMemcpyAsync(dst, static_cast<const char*>(ptr_) + offset, bytes, hipMemcpyDeviceToHost,
                           dev_->stream()),
            "hipMemcpyAsync(D2H)");
  hip_check(hipStreamSynchronize(dev_->stream()), "hipStreamSynchronize");
}

// ---- UniformBuffer ---------------------------------------------------------------

UniformBuffer::UniformBuffer(Device& dev, size_t bytes, int queue)
    : device_(dev), st_(dev.stream(queue)), buf_(dev, bytes), bytes_(bytes) {
  hip_check(hipHostMalloc(&staging_, bytes, hipHostMallocDefault), "hipHostMalloc");
  hip_check(hipEventCreateWithFlags(&done_, hipEventDisableTiming), "hipEventCreate");
  hip_check(hipEventRecord(done_, st_), "hipEventRecord");
}

UniformBuffer::~UniformBuffer() {
  if (done_) {
    hipEventSynchronize(done_);
    hipEventDestroy(done_);
  }
  if (staging_) hipHostFree(staging_);
}

void UniformBuffer::update(const void* src) {
  // the previous copy out of the staging memory must have finished
  hip_check(hipEventSynchronize(done_), "hipEventSynchronize");
  std::memcpy(staging_, src, bytes_);
  hip_check(hipMemcpyAsync(buf_.get(), staging_, bytes_, hipMemcpyHostToDevice, st_), "hipMemcpyAsync(uniform)");
  hip_check(hipEventRecord(done_, st_), "hipEventRecord");
}

// ---- camera (src/camera.rs:71-92, glam 0.30) --------------------------------------

static void cross3(const float a[3], const float b[3], float o[3]) {
  o[0] = a[1] * b[2] - b[1] * a[2];
  o[1] = a[2] * b[0] - b[2] * a[0];
  o[2] = a[0] * b[1] - b[0] * a[1];
}

static float dot3(const float a[3], const float b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

void look_to_rh(const float eye[3], const float dir[3], const float up[3], float out[16]) {
  float s0[3], s[3], u[3];
  cross3(dir, up, s0);
  const float inv = 1.0f / std::sqrt(dot3(s0, s0));
  for (int i = 0; i < 3; ++i) s[i] = s0[i] * inv;
  cross3(s, dir, u);
  const float m[16] = {s[0], u[0], -dir[0], 0.0f, s[1], u[1], -dir[1], 0.0f,
                       s[2], u[2], -dir[2], 0.0f, -dot3(eye, s), -dot3(eye, u), dot3(eye, dir), 1.0f};
  std::memcpy(out, m, sizeof m);
}

void perspective_rh(float fov_y, float aspect, float z_near, float z_far, float out[16]) {
  const float half = 0.5f * fov_y;
  const float h = std::cos(half) / std::sin(half);
  const float w = h / aspect;
  const float r = z_far / (z_near - z_far);
  const float m[16] = {w, 0, 0, 0, 0, h, 0, 0, 0, 0, r, -1.0f, 0, 0, r * z_near, 0};
  std::memcpy(out, m, sizeof m);
}

void Camera::forward(float out[3]) const {
  out[0] = std::cos(pitch) * std::sin(yaw);
  out[1] = std::sin(pitch);
  out[2] = std::cos(pitch) * std::cos(yaw);
}

void Camera::view(float out[16]) const {
  float f[3];
  forward(f);
  const float up[3] = {0.0f, 1.0f, 0.0f};
  look_to_rh(pos, f, up, out);
}

void Camera::projection(float aspect, float out[16]) const { perspective_rh(vertical_fov, aspect, z_start, z_end, out); }

CameraPod camera_pod_new(const Camera& c, uint32_t w, uint32_t h) {
  CameraPod p{};
  c.view(p.view);
  c.projection(float(w) / float(h), p.proj);
  p.size[0] = float(w);
  p.size[1] = float(h);
  return p;
}

// ---- buffers -------------------------------------------------------------------

CameraBuffer::CameraBuffer(Device& dev, int queue) : buf_(dev, sizeof(CameraPod), queue) {}

void CameraBuffer::update(const Camera& c, uint32_t w, uint32_t h) { update_with_pod(camera_pod_new(c, w, h)); }

void CameraBuffer::update_with_pod(const CameraPod& pod) {
  pod_ = pod;
  buf_.update(&pod_);
}

ModelTransformBuffer::ModelTransformBuffer(Device& dev, int queue) : buf_(dev, sizeof(ModelTransformPod), queue) {
  const float p[3] = {0, 0, 0}, r[4] = {0, 0, 0, 1}, s[3] = {1, 1, 1};
  update(p, r, s);
}

void ModelTransformBuffer::update(const float pos[3], const float rot[4], const float scale[3]) {
  ModelTransformPod t{};
  for (int i = 0; i < 3; ++i) {
    t.pos[i] = pos[i];
    t.scale[i] = scale[i];
  }
  for (int i = 0; i < 4; ++i) t.rot[i] = rot[i];
  update_with_pod(t);
}

void ModelTransformBuffer::update_with_pod(const ModelTransformPod& pod) {
  pod_ = pod;
  buf_.update(&pod_);
}

GaussianTransformBuffer::GaussianTransformBuffer(Device& dev, int queue)
    : buf_(dev, sizeof(GaussianTransformPod), queue) {
  update(1.0f, kModeSplat, 3, false, 3.0f);
}

void GaussianTransformBuffer::update(float size, int32_t mode, uint32_t sh_deg, bool no_sh0, float max_std_dev) {
  if (mode < kModeSplat || mode > kModePoint) throw Error(GS3D_ERROR_INVALID_ARGUMENT, "unknown display mode");
  if (sh_deg > 3) throw Error(GS3D_ERROR_INVALID_ARGUMENT, "SH degree must be 0..3");
  if (!(max_std_dev > 0.0f)) throw Error(GS3D_ERROR_INVALID_ARGUMENT, "max standard deviation must be positive");
  GaussianTransformPod t{};
  t.size = size;
  t.flags = pack_gaussian_flags(mode, sh_deg, no_sh0);
  t.max_std_dev = max_std_dev;
  update_with_pod(t);
}

void GaussianTransformBuffer::update_with_pod(const GaussianTransformPod& pod) {
  pod_ = pod;
  buf_.update(&pod_);
}

uint16_t float_to_half(float f) {
  const _Float16 h = static_cast<_Float16>(f);
  uint16_t u;
  std::memcpy(&u, &h, 2);
  return u;
}

uint32_t pack_color(const uint8_t c[4]) {
  return uint32_t(c[0]) | (uint32_t(c[1]) << 8) | (uint32_t(c[2]) << 16) | (uint32_t(c[3]) << 24);
}

void gaussian_cov3d(const gs3d_gaussian& g, float out[6]) { cov3d_from_scale_rot(g.scale, g.rot, out); }

uint64_t GaussiansBuffer::bytes_per_gaussian(PodKind pod) {
  return pod == kPodSingle ? (16 + 24 + 192) : (16 + 16 + 96);
}

uint64_t GaussiansBuffer::size_bytes() const { return uint64_t(n_) * bytes_per_gaussian(pod_); }

GaussiansBuffer::GaussiansBuffer(Device& dev, const gs3d_gaussian* g, uint32_t n, PodKind pod) : pod_(pod), n_(n) {
  if (pod != kPodSingle && pod != kPodHalf) throw Error(GS3D_ERROR_INVALID_ARGUMENT, "unknown gaussian pod");
  const uint64_t total = uint64_t(n) * bytes_per_gaussian(pod);
  if (total > dev.buffer_limit())
    throw Error(GS3D_ERROR_MODEL_SIZE_EXCEEDS_DEVICE_LIMIT,
                "model size exceeds the device limit: " + std::to_string(total) + " > " +
                    std::to_string(dev.buffer_limit()) + ", try smaller model or more aggressive compression");
  pc_ = DeviceBuffer(dev, size_t(n) * 16);
  if (pod == kPodSingle) {
    cov_ = DeviceBuffer(dev, size_t(n) * 24);
    sh_ = DeviceBuffer(dev, size_t(n) * 192);
  } else {
    rs_ = DeviceBuffer(dev, size_t(n) * 16);
    sh_ = DeviceBuffer(dev, size_t(n) * 96);
  }
  // pack and upload in chunks to bound host memory
  const uint32_t chunk = 1u << 18;
  std::vector<float> pc, cov, sh32;
  std::vector<uint16_t> rs, sh16;
  std::vector<uint8_t> plane;
  // the SH stream is plane-major (16-byte chunk q of Gaussian i at q * n + i):
  // records [b, b + m) of `planes` chunks each, Gaussian-major in `src`
  auto upload_sh_planes = [&](const void* src, uint32_t m, uint32_t b, int planes) {
    const uint8_t* s8 = static_cast<const uint8_t*>(src);
    plane.resize(size_t(m) * 16 * size_t(planes));
    for (int q = 0; q < planes; ++q)
      for (uint32_t j = 0; j < m; ++j)
        std::memcpy(&plane[(size_t(q) * m + j) * 16], s8 + (size_t(j) * planes + size_t(q)) * 16, 16);
    for (int q = 0; q < planes; ++q)
      sh_.upload(&plane[size_t(q) * m * 16], size_t(m) * 16, (size_t(q) * n + b) * 16);
  };
  for (uint32_t b = 0; b < n; b += chunk) {
    const uint32_t m = std::min(chunk, n - b);
    pc.assign(size_t(m) * 4, 0.0f);
    for (uint32_t j = 0; j < m; ++j) {
      const gs3d_gaussian& q = g[b + j];
      pc[4 * j + 0] = q.pos[0];
      pc[4 * j + 1] = q.pos[1];
      pc[4 * j + 2] = q.pos[2];
      const uint32_t cb = pack_color(q.color);
      std::memcpy(&pc[4 * j + 3], &cb, 4);
    }
    pc_.upload(pc.data(), size_t(m) * 16, size_t(b) * 16);
    if (pod == kPodSingle) {
      cov.assign(size_t(m) * 6, 0.0f);
      sh32.assign(size_t(m) * 48, 0.0f);
      for (uint32_t j = 0; j < m; ++j) {
        const gs3d_gaussian& q = g[b + j];
        gaussian_cov3d(q, &cov[6 * j]);
        std::memcpy(&sh32[48 * size_t(j)], &q.sh[0][0], 45 * sizeof(float));
      }
      cov_.upload(cov.data(), size_t(m) * 24, size_t(b) * 24);
      upload_sh_planes(sh32.data(), m, b, 12);
    } else {
      rs.assign(size_t(m) * 8, 0);
      sh16.assign(size_t(m) * 48, 0);
      for (uint32_t j = 0; j < m; ++j) {
        const gs3d_gaussian& q = g[b + j];
        for (int k = 0; k < 4; ++k) rs[8 * size_t(j) + k] = float_to_half(q.rot[k]);
        for (int k = 0; k < 3; ++k) rs[8 * size_t(j) + 4 + k] = float_to_half(q.scale[k]);
        for (int k = 0; k < 45; ++k) sh16[48 * size_t(j) + k] = float_to_half((&q.sh[0][0])[k]);
      }
      rs_.upload(rs.data(), size_t(m) * 16, size_t(b) * 16);
      upload_sh_planes(sh16.data(), m, b, 6);
    }
    // the staging vectors are reused: wait for the copies of this chunk
    dev.wait();
  }
}

GaussianStreams GaussiansBuffer::streams() const {
  GaussianStreams s{};
  s.pc = pc_.as<float4>();
  s.cov = pod_ == kPodSingle ? cov_.as<float2>() : nullptr;
  s.rs = pod_ == kPodHalf ? rs_.as<uint4>() : nullptr;
  s.sh32 = pod_ == kPodSingle ? sh_.as<float4>() : nullptr;
  s.sh16 = pod_ == kPodHalf ? sh_.as<uint4>() : nullptr;
  s.n = n_;
  return s;
}

IndirectArgsBuffer::IndirectArgsBuffer(Device& dev) : buf_(dev, 16) {
  const uint32_t init[4] = {6, 0, 0, 0};  // src/buffer/indirect_args.rs:16-21
  buf_.upload(init, 16);
  dev.wait();
}

RadixSortIndirectArgsBuffer::RadixSortIndirectArgsBuffer(Device& dev) : buf_(dev, 12) {
  const uint32_t init[3] = {1, 1, 1};  // src/buffer/indirect_args.rs:67
  buf_.upload(init, 12);
  dev.wait();
}

IndirectIndicesBuffer::IndirectIndicesBuffer(Device& dev, uint32_t n) : n_(n), buf_(dev, size_t(n) * 4) {}

uint32_t depth_buffer_len(uint32_t n) { return ((n + kHistoBlockKvs - 1) / kHistoBlockKvs) * kHistoBlockKvs; }

GaussiansDepthBuffer::GaussiansDepthBuffer(Device& dev, uint32_t n)
    : len_(depth_buffer_len(n)), buf_(dev, size_t(depth_buffer_len(n)) * 4) {}

SplatsBuffer::SplatsBuffer(Device& dev, uint32_t n) : dev_(dev), n_(n), buf_(dev, size_t(n) * sizeof(Splat2D)) {}

void SplatsBuffer::enable_depth() {
  if (depth_.size() == 0 && n_ > 0) depth_ = DeviceBuffer(dev_, size_t(n_) * 4);
}

void SplatsBuffer::set_rect_grid(const TileGrid& g) {
  if (rects_.size() == 0 && n_ > 0) rects_ = DeviceBuffer(dev_, size_t(n_) * 8);
  rect_grid_ = g;
}

Texture::Texture(Device& dev, uint32_t w, uint32_t h, TextureFormat fmt)
    : device_(&dev), w_(w), h_(h), fmt_(fmt) {
  if (w == 0 || h == 0 || w > 16384 || h > 16384) throw Error(GS3D_ERROR_INVALID_ARGUMENT, "bad texture size");
  if (fmt != kFormatRgba8Unorm && fmt != kFormatRgba32Float && fmt != kFormatDepth32Float)
    throw Error(GS3D_ERROR_INVALID_ARGUMENT, "unknown texture format");
  buf_ = DeviceBuffer(dev, size_t(w) * h * (fmt == kFormatRgba32Float ? 16 : 4));
}

Texture::~Texture() { device_->forget_texture(this); }

void Texture::write(const void* src, size_t bytes) {
  if (bytes != buf_.size()) throw Error(GS3D_ERROR_INVALID_ARGUMENT, "texture write size mismatch");
  // draws recorded before may still use the texture, or be recorded again
  device_->poll();
  hip_check(hipMemcpy(buf_.get(), src, bytes, hipMemcpyHostToDevice), "hipMemcpy(texture)");
}

void Texture::read(void* dst, size_t bytes) const {
  if (bytes != buf_.size()) throw Error(GS3D_ERROR_INVALID_ARGUMENT, "texture read size mismatch");
  device_->poll();
  buf_.download(dst, bytes);
}

// ---- StageTimer ----------------------------------------------------------------

StageTimer::StageTimer(Device& dev, int queue) : dev_(&dev), st_(dev.stream(queue)) {
  for (int s = 0; s < GS3D_STAGE_COUNT; ++s) {
    hip_check(hipEventCreate(&ev_[s][0]), "hipEventCreate");
    hip_check(hipEventCreate(&ev_[s][1]), "hipEventCreate");
    used_[s] = false;
  }
}

StageTimer::~StageTimer() {
  for (int s = 0; s < GS3D_STAGE_COUNT; ++s) {
    hipEventDestroy(ev_[s][0]);
    hipEventDestroy(ev_[s][1]);
  }
}

void StageTimer::mark(int stage, bool end) {
  if (!enabled_) return;
  hip_check(hipEventRecord(ev_[stage][end ? 1 : 0], st_), "hipEventRecord");
  if (end) used_[stage] = true;
}

hipEvent_t StageTimer::event(int stage, bool end) {
  if (!enabled_) return nullptr;
  if (end) used_[stage] = true;
  return ev_[stage][end ? 1 : 0];
}

void StageTimer::read(float ms[GS3D_STAGE_COUNT]) {
  dev_->wait();
  for (int s = 0; s < GS3D_STAGE_COUNT; ++s) {
    ms[s] = 0.0f;
    if (used_[s]) hip_check(hipEventElapsedTime(&ms[s], ev_[s][0], ev_[s][1]), "hipEventElapsedTime");
  }
}

// ---- Preprocessor ----------------------------------------------------------------

Preprocessor::Preprocessor(Device& dev, const CameraBuffer& camera, const ModelTransformBuffer& model,
                           const GaussianTransformBuffer& gt, const GaussiansBuffer& gaussians,
                           IndirectArgsBuffer& indirect_args, RadixSortIndirectArgsBuffer& rs_args,
                           IndirectIndicesBuffer& indices, GaussiansDepthBuffer& depth, SplatsBuffer& splats,
                           int queue)
    : dev_(dev),
      st_(dev.stream(queue)),
      camera_(camera),
      model_(model),
      gt_(gt),
      gaussians_(gaussians),
      indirect_args_(indirect_args),
      rs_args_(rs_args),
      indices_(indices),
      depth_(depth),
      splats_(splats),
      consts_(dev, sizeof(FrameConsts)),
      status_(dev, size_t(std::max(preprocess_num_tiles(gaussians.len()), 1u)) * 4),
      ticket_(dev, 256),
      thin_idx_(dev, size_t(std::max(gaussians.len(), 1u)) * 4),
      max_tiles_(preprocess_num_tiles(gaussians.len())) {
  // src/preprocessor.rs:239-246
  if (gaussians.size_bytes() > dev.buffer_limit())
    throw Error(GS3D_ERROR_MODEL_SIZE_EXCEEDS_DEVICE_LIMIT, "model size exceeds the device limit");
}

void Preprocessor::frame_pre(uint32_t gaussian_count) {
  ++generation_;
  const uint32_t n = std::min(gaussian_count, gaussians_.len());
  launch_frame_pre(st_, camera_.dev(), model_.dev(), gt_.dev(), consts_.as<FrameConsts>(),
                   indirect_args_.dev(), rs_args_.dev(), status_.as<uint32_t>(), preprocess_num_tiles(n),
                   ticket_.as<uint32_t>(), sort_hist_);
}

void Preprocessor::project(uint32_t gaussian_count) {
  const uint32_t n = std::min(gaussian_count, gaussians_.len());
  PreprocessArgs a{};
  a.fc = consts_.as<FrameConsts>();
  a.g = gaussians_.streams();
  a.g.n = n;
  a.indices = indices_.dev();
  a.depth = depth_.dev();
  a.depth_len = depth_.len();
  a.splats = splats_.dev();
  a.indirect_args = indirect_args_.dev();
  a.dispatch_args = rs_args_.dev();
  a.status = status_.as<uint32_t>();
  a.ticket = ticket_.as<uint32_t>();
  a.num_tiles = preprocess_num_tiles(n);
  a.err = dev_.err();
  a.selection = selection_;
  a.splat_depth = splat_depth_;
  a.rects = splats_.rects_on() ? splats_.rects() : nullptr;
  a.rect_grid = splats_.rect_grid();
  a.sort_hist = sort_hist_;
  a.sort_max = sort_max_;
  a.invert_selection = invert_selection_ ? 1u : 0u;
  a.thin_count = ticket_.as<uint32_t>() + kThinCountWord;
  a.thin_idx = thin_idx_.as<uint32_t>();
  launch_preprocess(st_, gaussians_.pod(), a);
  hip_check(hipGetLastError(), "launch preprocess");
}

void Preprocessor::preprocess(uint32_t gaussian_count) {
  frame_pre(gaussian_count);
  project(gaussian_count);
}

// ---- sort ----------------------------------------------------------------------

SortScratchBuffers::SortScratchBuffers(Device& dev, uint32_t capacity, int passes)
    : dev_(dev),
      capacity_(capacity),
      passes_(passes),
      hist_(dev, size_t(passes) * 256 * 4),
      lookback_(dev, std::max<size_t>(sort_lookback_words(capacity, passes), 1) * 4),
      tickets_(dev, 64) {}

void SortScratchBuffers::set_items(uint32_t items) {
  if (!sort_items_supported(items))
    throw Error(GS3D_ERROR_INVALID_ARGUMENT, "sort keys per thread must be 8, 12, 16, 20 or 24");
  items_ = items;
}

SortScratch SortScratchBuffers::get() const {
  SortScratch s{};
  s.items = items_;
  s.hist = hist_.as<uint32_t>();
  s.lookback = lookback_.as<uint32_t>();
  s.tickets = tickets_.as<uint32_t>();
  s.max_tiles = sort_max_tiles(capacity_);
  s.err = dev_.err();
  return s;
}

RadixSorter::RadixSorter(Device& dev, GaussiansDepthBuffer& depth, IndirectIndicesBuffer& indices, int queue)
    : dev_(dev),
      st_(dev.stream(queue)),
      depth_(depth),
      indices_(indices),
      keys_b_(dev, size_t(indices.len()) * 4),
      vals_b_(dev, size_t(indices.len()) * 4),
      scratch_(dev, indices.len(), 4) {}

void RadixSorter::sort(const RadixSortIndirectArgsBuffer& rs_args, bool hist_ready) {
  // the reference sorts dispatch.x * 3840 entries, payload clamped to N
  // (src/radix_sorter.rs:658-669, 741)
  sort_count(rs_args.dev(), kHistoBlockKvs, hist_ready);
}

void RadixSorter::sort_count(const uint32_t* count_dev, uint32_t count_mul, bool hist_ready) {
  const uint32_t n = indices_.len();
  if (n == 0) return;
  SortCount c{count_dev, count_mul, n};
  uint32_t* keys = reinterpret_cast<uint32_t*>(depth_.dev());
  launch_radix_sort(st_, dev_.num_cus(), c, keys, indices_.dev(), keys_b_.as<uint32_t>(),
                    vals_b_.as<uint32_t>(), 4, scratch_.get(), keys, n, depth_.len(), hist_ready);
  hip_check(hipGetLastError(), "launch sort");
}

// ---- Renderer ----------------------------------------------------------------------

Renderer::Renderer(Device& dev, TextureFormat format, const CameraBuffer& camera, const ModelTransformBuffer& model,
                   const GaussianTransformBuffer& gt, const GaussiansBuffer& gaussians,
                   const IndirectIndicesBuffer& indices, SplatsBuffer& splats, const FrameConsts* shared_consts,
                   int queue)
    : dev_(dev),
      st_(dev.stream(queue)),
      format_(format),
      camera_(camera),
      model_(model),
      gt_(gt),
      gaussians_(gaussians),
      indices_(indices),
      splats_(splats),
      shared_consts_(shared_consts),
      own_consts_(dev, sizeof(FrameConsts)),
      bin_status_(dev, size_t(std::max(bin_max_tiles(indices.len()), 1u)) * 4),
      bin_ticket_(dev, 64),
      pair_count_(dev, 64),
      timer_(dev, queue) {
  if (format != kFormatRgba8Unorm && format != kFormatRgba32Float)
    throw Error(GS3D_ERROR_INVALID_ARGUMENT, "unknown texture format");
  void* mapped = nullptr;
  hip_check(hipHostMalloc(&mapped, 64, hipHostMallocMapped), "hipHostMalloc(mapped)");
  pair_total_host_ = static_cast<uint32_t*>(mapped);
  *pair_total_host_ = 0u;
  void* alias = nullptr;
  hip_check(hipHostGetDevicePointer(&alias, mapped, 0), "hipHostGetDevicePointer");
  pair_total_dev_ = static_cast<uint32_t*>(alias);
  // src/renderer.rs:130-135
  if (gaussians.size_bytes() > dev.buffer_limit())
    throw Error(GS3D_ERROR_MODEL_SIZE_EXCEEDS_DEVICE_LIMIT, "model size exceeds the device limit");
  ensure_capacity(std::max<uint32_t>(1u << 20, std::min<uint64_t>(uint64_t(indices.len()) * 2, kMaxPairs)));
}

Renderer::~Renderer() {
  dev_.forget_renderer(this);
  if (pair_total_host_) hipHostFree(pair_total_host_);
}

void Renderer::set_tile_size(uint32_t w, uint32_t h) {
  const bool ok = (w == 8 && h == 8) || (w == 16 && h == 16) || (w == 32 && h == 8);
  if (!ok) throw Error(GS3D_ERROR_INVALID_ARGUMENT, "tile size must be 8x8, 16x16 or 32x8");
  tile_w_ = w;
  tile_h_ = h;
}

void Renderer::set_sort_items(uint32_t items) {
  if (!sort_items_supported(items))
    throw Error(GS3D_ERROR_INVALID_ARGUMENT, "sort keys per thread must be 8, 12, 16, 20 or 24");
  sort_items_ = items;
  if (scratch_) scratch_->set_items(items);
}

void Renderer::ensure_capacity(uint32_t pairs) {
  if (pairs <= capacity_) return;
  dev_.wait();
  for (int b = 0; b < 2; ++b) {
    pk_[b] = DeviceBuffer(dev_, size_t(pairs) * 4);
    pv_[b] = DeviceBuffer(dev_, size_t(pairs) * 4);
  }
  scratch_ = std::make_unique<SortScratchBuffers>(dev_, pairs, kTileSortPasses);
  scratch_->set_items(sort_items_);
  capacity_ = pairs;
}

void Renderer::set_capacity(uint32_t pairs) {
  if (pairs == 0 || pairs > kMaxPairs) throw Error(GS3D_ERROR_INVALID_ARGUMENT, "pair capacity must be 1..2^30-1");
  dev_.poll();  // nothing recorded before may be recorded again with other lists
  capacity_ = 0;
  ensure_capacity(pairs);
}

bool Renderer::grow_if_overflowed() {
  // the frames have completed: k_ranges left the largest pair total since
  // the last check in mapped memory (and in word 1 of pair_count_)
  const uint32_t total = last_pair_total();
  checked_pair_total_ = total;
  *static_cast<volatile uint32_t*>(pair_total_host_) = 0u;
  const uint32_t zero = 0u;
  hip_check(hipMemcpy(pair_count_.as<uint32_t>() + 1, &zero, 4, hipMemcpyHostToDevice), "reset pair maximum");
  if (total <= capacity_) return false;
  // status words hold 30-bit values: no list may hold more pairs
  if (total > kMaxPairs)
    throw Error(GS3D_ERROR_OUT_OF_MEMORY, "a frame needs " + std::to_string(total) + " (tile, splat) pairs, above " +
                                              std::to_string(kMaxPairs));
  const uint64_t want = std::min<uint64_t>(uint64_t(total) + total / 4 + 1024, kMaxPairs);
  ensure_capacity(uint32_t(want));
  return true;
}

void Renderer::settle() {
  for (int attempt = 0; grow_if_overflowed(); ++attempt) {
    if (attempt >= 3) throw Error(GS3D_ERROR_KERNEL, "tile lists still overflowed after growing them three times");
    rerender();
    hip_check(hipStreamSynchronize(st_), "hipStreamSynchronize");
  }
}

uint64_t Renderer::count_fragments(Texture& target, const IndirectArgsBuffer& indirect_args) {
  if (frag_count_.size() == 0) frag_count_ = DeviceBuffer(dev_, 64);
  const BlendMode old = blend_;
  blend_ = kBlendExactUnorm8;
  counting_ = true;
  try {
    dev_.poll();  // nothing recorded before may be recorded again after this
    // (the count restarts with each attempt of the draw)
    for (int attempt = 0;; ++attempt) {
      hip_check(hipMemsetAsync(frag_count_.get(), 0, 4, st_), "hipMemsetAsync");
      render(target, indirect_args, false);
      dev_.wait();
      if (!grow_if_overflowed()) break;
      if (attempt >= 3) throw Error(GS3D_ERROR_KERNEL, "tile lists still overflowed after growing them three times");
    }
  } catch (...) {
    blend_ = old;
    counting_ = false;
    throw;
  }
  blend_ = old;
  counting_ = false;
  uint32_t n = 0;
  hip_check(hipMemcpyAsync(&n, frag_count_.get(), 4, hipMemcpyDeviceToHost, st_), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(st_), "hipStreamSynchronize");
  return n;
}

void Renderer::rerender() {
  if (last_target_ && last_args_) render(*last_target_, *last_args_, last_standalone_, last_clear_, last_depth_);
}

TileGrid Renderer::grid_for(const Texture& target) const {
  TileGrid g{};
  g.width = target.width();
  g.height = target.height();
  g.tile_w = tile_w_;
  g.tile_h = tile_h_;
  g.tiles_x = (g.width + tile_w_ - 1) / tile_w_;
  g.tiles_y = (g.height + tile_h_ - 1) / tile_h_;
  return g;
}

void Renderer::set_depth_stencil(bool enabled, CompareFn cmp, bool write) {
  if (enabled && (int(cmp) < kCmpNever || int(cmp) > kCmpAlways))
    throw Error(GS3D_ERROR_INVALID_ARGUMENT, "unknown depth compare function");
  depth_enabled_ = enabled;
  depth_cmp_ = cmp;
  depth_write_ = write;
  if (enabled) splats_.enable_depth();
}

static uint32_t ceil_log2(uint32_t x) {
  uint32_t b = 0;
  while ((1u << b) < x && b < 31) ++b;
  return b;
}

void Renderer::render(Texture& target, const IndirectArgsBuffer& indirect_args, bool standalone, bool clear,
                      Texture* depth) {
  if (target.format() != format_)
    throw Error(GS3D_ERROR_INVALID_ARGUMENT, "render target format differs from the renderer's format");
  // a pipeline's depth-stencil state and the pass's depth attachment go
  // together (wgpu validates the same)
  if (depth_enabled_ != (depth != nullptr))
    throw Error(GS3D_ERROR_INVALID_ARGUMENT, depth_enabled_ ? "this renderer needs a depth attachment"
                                                            : "this renderer has no depth-stencil state");
  if (depth && (depth->format() != kFormatDepth32Float || depth->width() != target.width() ||
                depth->height() != target.height()))
    throw Error(GS3D_ERROR_INVALID_ARGUMENT, "the depth attachment must be Depth32Float of the target's size");
  hipStream_t st = st_;
  StageTimer& tm = active_timer();
  const FrameConsts* fc = shared_consts_;
  if (standalone || fc == nullptr) {
    launch_frame_consts(st, camera_.dev(), model_.dev(), gt_.dev(), own_consts_.as<FrameConsts>());
    fc = own_consts_.as<FrameConsts>();
  }
  if (standalone)
    launch_vertex(st, gaussians_.pod(), fc, gaussians_.streams(), indirect_args.dev(), indices_.dev(), splats_.dev(),
                  indices_.len(), depth_enabled_ ? splats_.depth() : nullptr);

  const TileGrid g = grid_for(target);
  // tile_rect packs tile coordinates into 12 bits
  if (g.tiles_x >= 4096u || g.tiles_y >= 4096u) throw Error(GS3D_ERROR_INVALID_ARGUMENT, "tile grid above 4095 tiles");
  const uint32_t num_tiles = g.tiles_x * g.tiles_y;
  if (ranges_words_ < num_tiles * 2) {
    dev_.wait();
    ranges_ = DeviceBuffer(dev_, size_t(num_tiles) * 8);
    ranges_words_ = num_tiles * 2;
  }

  tm.mark(GS3D_STAGE_BIN, false);
  const uint32_t bin_tiles = bin_max_tiles(indices_.len());
  const int passes = int((ceil_log2(num_tiles) + 7) / 8);
  // the tile sort's scratch holds kTileSortPasses passes' histograms and
  // look-back words (16384-pixel targets in 8x8 tiles: 2^22 tiles, 3 passes)
  if (passes > kTileSortPasses) throw Error(GS3D_ERROR_INVALID_ARGUMENT, "tile grid needs more sort passes than allocated");
  const SortScratch tile_scratch = scratch_->get();
  launch_bin_reset(st, ranges_.as<uint32_t>(), num_tiles, bin_status_.as<uint32_t>(), bin_tiles,
                   bin_ticket_.as<uint32_t>(), pair_count_.as<uint32_t>(), tile_scratch.hist,
                   uint32_t(passes) * 256u);
  BinArgs b{};
  b.indirect_args = indirect_args.dev();
  b.indices = indices_.dev();
  b.splats = splats_.dev();
  b.num_gaussians = splats_.len();
  b.fc = fc;
  // rectangles written with this frame's splats for this grid (not for a
  // standalone draw, which recomputes the splats)
  b.rects = (!standalone && splats_.rects_on() && same_grid(splats_.rect_grid(), g)) ? splats_.rects() : nullptr;
  b.grid = g;
  b.pair_keys = pk_[0].as<uint32_t>();
  b.pair_vals = pv_[0].as<uint32_t>();
  b.capacity = capacity_;
  b.pair_count = pair_count_.as<uint32_t>();
  b.status = bin_status_.as<uint32_t>();
  b.ticket = bin_ticket_.as<uint32_t>();
  b.max_tiles = bin_tiles;
  b.err = dev_.err();
  b.sort_hist = tile_scratch.hist;  // the tile sort's first histogram
  launch_bin(st, b);
  tm.mark(GS3D_STAGE_BIN, true);

  tm.mark(GS3D_STAGE_TILE_SORT, false);
  SortCount c{pair_count_.as<uint32_t>(), 1u, capacity_};
  const int which = launch_radix_sort(st, dev_.num_cus(), c, pk_[0].as<uint32_t>(), pv_[0].as<uint32_t>(),
                                      pk_[1].as<uint32_t>(), pv_[1].as<uint32_t>(), passes, tile_scratch, nullptr, 0,
                                      0, true);
  launch_ranges(st, pair_count_.as<uint32_t>(), capacity_, pk_[which].as<uint32_t>(), ranges_.as<uint32_t>(),
                num_tiles, dev_.num_cus(), pair_count_.as<uint32_t>() + 1, pair_total_dev_, dev_.err());
  tm.mark(GS3D_STAGE_TILE_SORT, true);

  tm.mark(GS3D_STAGE_COMPOSITE, false);
  const uint32_t tile_px = tile_w_ * tile_h_;
  const size_t items = composite_max_items(num_tiles, capacity_, seg_len_);
  const size_t heavy = composite_max_heavy_items(capacity_, seg_len_);
  auto grow = [&](DeviceBuffer& b, size_t bytes) {
    if (b.size() < bytes) {
      dev_.wait();
      b = DeviceBuffer(dev_, bytes);
    }
  };
  grow(item_tile_, items * 4);
  grow(item_seg_, items * 4);
  grow(seg_base_, size_t(num_tiles) * 4);
  grow(plan_, kPlanWords * sizeof(uint32_t));
  grow(plan_counts_, size_t(composite_plan_blocks(num_tiles)) * 8);
  grow(partial_, heavy * tile_px * sizeof(float4));
  CompositeArgs ca{};
  ca.fc = fc;
  ca.pair_vals = pv_[which].as<uint32_t>();
  ca.ranges = ranges_.as<uint32_t>();
  ca.splats = splats_.dev();
  ca.num_gaussians = splats_.len();
  ca.err = dev_.err();
  ca.grid = g;
  ca.format = format_;
  // The fast blend tests every fragment against the attachment as it was
  // before the draw, which is what the rasterizer's back-to-front order gives
  // for Less / LessEqual / Equal / Always even with depth writes. With writes,
  // Greater / GreaterEqual / NotEqual depend on what the farther splats drawn
  // before wrote: those draws take the exact (in-order) path.
  const bool order_dependent = depth && depth_write_ &&
                               (depth_cmp_ == kCmpGreater || depth_cmp_ == kCmpGreaterEqual || depth_cmp_ == kCmpNotEqual);
  ca.blend = order_dependent ? kBlendExactUnorm8 : blend_;
  ca.load = clear ? 0 : 1;
  ca.frag_count = counting_ ? frag_count_.as<uint32_t>() : nullptr;
  ca.out = target.dev();
  ca.capacity = capacity_;
  // depth-tested draws keep whole tiles (no split-K segments): a lane's
  // depth write needs its nearest passing fragment in one pass
  ca.seg_len = depth ? 0x7FFFFFFFu : seg_len_;
  ca.item_tile = item_tile_.as<uint32_t>();
  ca.item_seg = item_seg_.as<uint32_t>();
  ca.seg_base = seg_base_.as<uint32_t>();
  ca.pair_count = pair_count_.as<uint32_t>();
  ca.splat_depth = depth ? splats_.depth() : nullptr;
  ca.depth = depth ? static_cast<float*>(depth->dev()) : nullptr;
  ca.depth_cmp = int(depth_cmp_);
  ca.depth_write = depth_write_ ? 1 : 0;
  ca.plan = plan_.as<uint32_t>();
  ca.plan_counts = plan_counts_.as<uint32_t>();
  ca.partial = partial_.as<float4>();
  launch_composite(st, ca, dev_.num_cus(), tm.event(GS3D_STAGE_COMPOSITE_MAIN, false),
                   tm.event(GS3D_STAGE_COMPOSITE_MAIN, true));
  tm.mark(GS3D_STAGE_COMPOSITE, true);
  hip_check(hipGetLastError(), "launch render");

  last_target_ = &target;
  last_args_ = &indirect_args;
  last_standalone_ = standalone;
  last_clear_ = clear;
  last_depth_ = depth;
}

// ---- Viewer --------------------------------------------------------------------------

Viewer::Viewer(Device& dev, TextureFormat format, const gs3d_gaussian* g, uint32_t n, PodKind pod)
    : Viewer(dev, format, std::make_shared<GaussiansBuffer>(dev, g, n, pod), 0) {}

Viewer::Viewer(Device& dev, TextureFormat format, std::shared_ptr<GaussiansBuffer> m, int q)
    : device(dev),
      queue(q),
      model(std::move(m)),
      camera_buffer(dev, q),
      model_transform_buffer(dev, q),
      gaussian_transform_buffer(dev, q),
      gaussians_buffer(*model),
      indirect_args_buffer(dev),
      radix_sort_indirect_args_buffer(dev),
      indirect_indices_buffer(dev, model->len()),
      gaussians_depth_buffer(dev, model->len()),
      splats_buffer(dev, model->len()),
      preprocessor(dev, camera_buffer, model_transform_buffer, gaussian_transform_buffer, gaussians_buffer,
                   indirect_args_buffer, radix_sort_indirect_args_buffer, indirect_indices_buffer,
                   gaussians_depth_buffer, splats_buffer, q),
      radix_sorter(dev, gaussians_depth_buffer, indirect_indices_buffer, q),
      renderer(dev, format, camera_buffer, model_transform_buffer, gaussian_transform_buffer, gaussians_buffer,
               indirect_indices_buffer, splats_buffer, preprocessor.frame_consts(), q),
      timer(dev, q) {
  renderer.set_timer(&timer);
}

void Viewer::update_selection(const uint32_t* words, uint32_t count) {
  const uint32_t need = (gaussians_buffer.len() + 31u) / 32u;
  if (count != need)
    throw Error(GS3D_ERROR_INVALID_ARGUMENT,
                "selection needs " + std::to_string(need) + " words, got " + std::to_string(count));
  if (need == 0) return;
  // the previous frames may still read the bitset, or be recorded again
  device.poll();
  if (selection_buffer.size() == 0) selection_buffer = DeviceBuffer(device, size_t(need) * 4);
  hip_check(hipMemcpy(selection_buffer.get(), words, size_t(need) * 4, hipMemcpyHostToDevice), "hipMemcpy(selection)");
  preprocessor.set_selection(selection_buffer.as<uint32_t>());
}

void Viewer::set_depth_stencil(bool enabled, CompareFn cmp, bool write) {
  renderer.set_depth_stencil(enabled, cmp, write);
  preprocessor.set_splat_depth(enabled ? splats_buffer.depth() : nullptr);
}

Viewer::State Viewer::state() const {
  State s{};
  s.camera = camera_buffer.host();
  s.model = model_transform_buffer.host();
  s.gaussian = gaussian_transform_buffer.host();
  s.blend = renderer.blend();
  s.tile_w = renderer.tile_w();
  s.tile_h = renderer.tile_h();
  return s;
}

void Viewer::restore(const State& s) {
  camera_buffer.update_with_pod(s.camera);
  model_transform_buffer.update_with_pod(s.model);
  gaussian_transform_buffer.update_with_pod(s.gaussian);
  renderer.set_blend(s.blend);
  renderer.set_tile_size(s.tile_w, s.tile_h);
}

void Viewer::draw(Texture& target, bool standalone) {
  renderer.render(target, indirect_args_buffer, standalone);
  // recorded again only over the splats it drew: none of a later frame
  RecordedUnit u;
  const uint64_t gen = preprocessor.generation();
  const State st = state();
  Texture* t = &target;
  u.replay = [this, t, standalone, st] {
    restore(st);
    draw(*t, standalone);
  };
  u.renderers = {&renderer};
  u.textures = {&target};
  u.blocker = [this, gen]() -> const char* {
    return preprocessor.generation() == gen ? nullptr : "the viewer preprocessed another frame after the draw";
  };
  device.record_unit(std::move(u));
}

void Viewer::render_with_depth(Texture& target, Texture& depth, bool load) {
  {
    RecordedUnit u;
    const State st = state();
    Texture *t = &target, *d = &depth;
    u.replay = [this, t, d, load, st] {
      restore(st);
      render_with_depth(*t, *d, load);
    };
    u.renderers = {&renderer};
    u.textures = {&target, &depth};
    const bool writes = renderer.depth_write();
    u.blocker = [load, writes]() -> const char* {
      if (load) return "the draw loads the target's contents (LoadOp::Load)";
      if (writes) return "the draw writes its depth attachment";
      return nullptr;
    };
    device.record_unit(std::move(u));
  }
  splats_buffer.set_rect_grid(renderer.grid_for(target));
  // the projection also writes the depth sort's first histogram
  preprocessor.set_sort_hist(radix_sorter.hist(), radix_sorter.len());
  timer.mark(GS3D_STAGE_PREPROCESS, false);
  preprocessor.frame_pre(gaussians_buffer.len());
  timer.mark(GS3D_STAGE_PREPROCESS_MAIN, false);
  preprocessor.project(gaussians_buffer.len());
  timer.mark(GS3D_STAGE_PREPROCESS_MAIN, true);
  timer.mark(GS3D_STAGE_PREPROCESS, true);
  timer.mark(GS3D_STAGE_SORT, false);
  radix_sorter.sort(radix_sort_indirect_args_buffer, true);
  timer.mark(GS3D_STAGE_SORT, true);
  renderer.render(target, indirect_args_buffer, false, !load, &depth);
}

void Viewer::render(Texture& target, bool record) {
  if (record) {
    RecordedUnit u;
    const State st = state();
    Texture* t = &target;
    u.replay = [this, t, st] {
      restore(st);
      render(*t);
    };
    u.renderers = {&renderer};
    u.textures = {&target};
    device.record_unit(std::move(u));
  }
  splats_buffer.set_rect_grid(renderer.grid_for(target));
  // the projection also writes the depth sort's first histogram
  preprocessor.set_sort_hist(radix_sorter.hist(), radix_sorter.len());
  timer.mark(GS3D_STAGE_PREPROCESS, false);
  preprocessor.frame_pre(gaussians_buffer.len());
  timer.mark(GS3D_STAGE_PREPROCESS_MAIN, false);
  preprocessor.project(gaussians_buffer.len());
  timer.mark(GS3D_STAGE_PREPROCESS_MAIN, true);
  timer.mark(GS3D_STAGE_PREPROCESS, true);
  timer.mark(GS3D_STAGE_SORT, false);
  radix_sorter.sort(radix_sort_indirect_args_buffer, true);
  timer.mark(GS3D_STAGE_SORT, true);
  renderer.render(target, indirect_args_buffer, false);
}

}  // namespace gs3d
