// extern "C" boundary (include/gs3d.h) over the C++ mirror in viewer.hpp.
// Every entry point catches exceptions and returns a gs3d_status; the
// message is kept per thread for gs3d_last_error().
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <string>

#include "../../include/gs3d.h"
#include "batch.hpp"
#include "kernels.hpp"
#include "math.hpp"
#include "viewer.hpp"

// Every handle made from a device shares ownership of it (declared first,
// released last): destroying the device handle while others are alive only
// drops its reference.
using DevicePtr = std::shared_ptr<gs3d::Device>;

struct gs3d_device {
  DevicePtr dev;
  explicit gs3d_device(int ordinal) : dev(std::make_shared<gs3d::Device>(ordinal)) {}
};

struct gs3d_model {
  DevicePtr keep;
  std::shared_ptr<gs3d::GaussiansBuffer> buf;
};

struct gs3d_multi_model_viewer {
  DevicePtr keep;
  gs3d::MultiModelViewer v;
  gs3d_multi_model_viewer(DevicePtr d, gs3d::TextureFormat f) : keep(std::move(d)), v(*keep, f) {}
};

struct gs3d_batch {
  DevicePtr keep;
  gs3d::BatchRenderer b;
  gs3d_batch(DevicePtr d, std::vector<std::shared_ptr<gs3d::GaussiansBuffer>> m, int q)
      : keep(std::move(d)), b(*keep, std::move(m), q) {}
};

struct gs3d_server {
  DevicePtr keep;
  gs3d::Server s;
  gs3d_server(DevicePtr d, std::vector<std::shared_ptr<gs3d::GaussiansBuffer>> m, int q, uint64_t budget)
      : keep(std::move(d)), s(*keep, std::move(m), q, budget) {}
};

struct gs3d_texture {
  DevicePtr keep;
  gs3d::Texture tex;
  gs3d_texture(DevicePtr d, uint32_t w, uint32_t h, gs3d::TextureFormat f) : keep(std::move(d)), tex(*keep, w, h, f) {}
};

struct gs3d_viewer {
  DevicePtr keep;
  gs3d::Viewer v;
  gs3d_viewer(DevicePtr d, gs3d::TextureFormat f, const gs3d_gaussian* g, uint32_t n, gs3d::PodKind p)
      : keep(std::move(d)), v(*keep, f, g, n, p) {}
};

struct gs3d_radix_sorter {
  DevicePtr keep;
  gs3d::Device& dev;
  gs3d::IndirectIndicesBuffer vals;
  gs3d::GaussiansDepthBuffer keys;
  gs3d::DeviceBuffer count;
  gs3d::RadixSorter sorter;
  uint32_t n = 0;
  gs3d_radix_sorter(DevicePtr d, uint32_t cap)
      : keep(std::move(d)), dev(*keep), vals(dev, cap), keys(dev, cap), count(dev, 64), sorter(dev, keys, vals) {}
};

namespace {

thread_local std::string g_last_error;

template <typename F>
gs3d_status guard(F&& f) {
  try {
    f();
    g_last_error.clear();
    return GS3D_OK;
  } catch (const gs3d::Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host out of memory";
    return GS3D_ERROR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return GS3D_ERROR_DEVICE;
  }
}

void need(bool ok, const char* what) {
  if (!ok) throw gs3d::Error(GS3D_ERROR_INVALID_ARGUMENT, what);
}

// a colour format (render targets of viewers)
gs3d::TextureFormat to_format(gs3d_texture_format f) {
  need(f == GS3D_FORMAT_RGBA8_UNORM || f == GS3D_FORMAT_RGBA32_FLOAT, "unknown colour texture format");
  return f == GS3D_FORMAT_RGBA8_UNORM ? gs3d::kFormatRgba8Unorm : gs3d::kFormatRgba32Float;
}

// any texture format (textures may also be depth attachments)
gs3d::TextureFormat to_texture_format(gs3d_texture_format f) {
  if (f == GS3D_FORMAT_DEPTH32_FLOAT) return gs3d::kFormatDepth32Float;
  return to_format(f);
}

}  // namespace

extern "C" {

const char* gs3d_last_error(void) { return g_last_error.c_str(); }

gs3d_status gs3d_device_create(int hip_device, gs3d_device** out) {
  return guard([&] {
    need(out != nullptr, "out is null");
    *out = nullptr;
    *out = new gs3d_device(hip_device);
  });
}

void gs3d_device_destroy(gs3d_device* dev) { delete dev; }

gs3d_status gs3d_device_poll(gs3d_device* dev) {
  return guard([&] {
    need(dev != nullptr, "device is null");
    dev->dev->poll();
  });
}

gs3d_status gs3d_device_set_buffer_limit(gs3d_device* dev, uint64_t bytes) {
  return guard([&] {
    need(dev != nullptr, "device is null");
    dev->dev->set_buffer_limit(bytes);
  });
}

gs3d_status gs3d_device_compute_units(gs3d_device* dev, int* out) {
  return guard([&] {
    need(dev != nullptr && out != nullptr, "null argument");
    *out = dev->dev->num_cus();
  });
}

gs3d_status gs3d_texture_create(gs3d_device* dev, uint32_t width, uint32_t height, gs3d_texture_format format,
                                gs3d_texture** out) {
  return guard([&] {
    need(dev != nullptr && out != nullptr, "null argument");
    *out = nullptr;
    *out = new gs3d_texture(dev->dev, width, height, to_texture_format(format));
  });
}

void gs3d_texture_destroy(gs3d_texture* tex) { delete tex; }

gs3d_status gs3d_texture_read(gs3d_texture* tex, void* dst, size_t bytes) {
  return guard([&] {
    need(tex != nullptr && dst != nullptr, "null argument");
    tex->tex.read(dst, bytes);
  });
}

gs3d_status gs3d_texture_write(gs3d_texture* tex, const void* src, size_t bytes) {
  return guard([&] {
    need(tex != nullptr && src != nullptr, "null argument");
    tex->tex.write(src, bytes);
  });
}

gs3d_status gs3d_viewer_new(gs3d_device* dev, gs3d_texture_format format, const gs3d_gaussian* gaussians,
                            uint32_t count, gs3d_gaussian_pod pod, gs3d_viewer** out) {
  return guard([&] {
    need(dev != nullptr && out != nullptr, "null argument");
    need(count == 0 || gaussians != nullptr, "gaussians is null");
    need(count < (1u << 30), "more than 2^30 Gaussians");
    need(pod == GS3D_POD_SINGLE || pod == GS3D_POD_HALF, "unknown gaussian pod");
    *out = nullptr;
    *out = new gs3d_viewer(dev->dev, to_format(format), gaussians, count,
                           pod == GS3D_POD_SINGLE ? gs3d::kPodSingle : gs3d::kPodHalf);
  });
}

void gs3d_viewer_destroy(gs3d_viewer* v) {
  if (v) {
    try {
      v->v.device.wait();
    } catch (...) {
    }
  }
  delete v;
}

gs3d_status gs3d_viewer_update_camera(gs3d_viewer* v, const gs3d_camera* camera, uint32_t width, uint32_t height) {
  return guard([&] {
    need(v != nullptr && camera != nullptr, "null argument");
    need(width > 0 && height > 0, "empty target size");
    gs3d::Camera c;
    for (int i = 0; i < 3; ++i) c.pos[i] = camera->pos[i];
    c.z_start = camera->z_start;
    c.z_end = camera->z_end;
    c.vertical_fov = camera->vertical_fov;
    c.pitch = camera->pitch;
    c.yaw = camera->yaw;
    v->v.update_camera(c, width, height);
  });
}

gs3d_status gs3d_viewer_update_camera_with_pod(gs3d_viewer* v, const gs3d_camera_pod* pod) {
  return guard([&] {
    need(v != nullptr && pod != nullptr, "null argument");
    gs3d::CameraPod p;
    std::memcpy(&p, pod, sizeof p);
    v->v.update_camera_with_pod(p);
  });
}

gs3d_status gs3d_viewer_update_model_transform(gs3d_viewer* v, const float pos[3], const float rot[4],
                                               const float scale[3]) {
  return guard([&] {
    need(v != nullptr && pos && rot && scale, "null argument");
    v->v.update_model_transform(pos, rot, scale);
  });
}

gs3d_status gs3d_viewer_update_gaussian_transform(gs3d_viewer* v, float size, gs3d_display_mode display_mode,
                                                  uint32_t sh_deg, int no_sh0, float max_std_dev) {
  return guard([&] {
    need(v != nullptr, "viewer is null");
    v->v.update_gaussian_transform(size, int32_t(display_mode), sh_deg, no_sh0 != 0, max_std_dev);
  });
}

gs3d_status gs3d_viewer_update_selection(gs3d_viewer* v, const uint32_t* words, uint32_t word_count) {
  return guard([&] {
    need(v != nullptr && (words != nullptr || word_count == 0), "null argument");
    v->v.update_selection(words, word_count);
  });
}

gs3d_status gs3d_viewer_update_invert_selection(gs3d_viewer* v, int invert) {
  return guard([&] {
    need(v != nullptr, "viewer is null");
    v->v.update_invert_selection(invert != 0);
  });
}

gs3d_status gs3d_viewer_render(gs3d_viewer* v, gs3d_texture* target) {
  return guard([&] {
    need(v != nullptr && target != nullptr, "null argument");
    v->v.render(target->tex);
  });
}

gs3d_status gs3d_viewer_preprocess(gs3d_viewer* v) {
  return guard([&] {
    need(v != nullptr, "viewer is null");
    v->v.preprocessor.preprocess(v->v.gaussians_buffer.len());
  });
}

gs3d_status gs3d_viewer_sort(gs3d_viewer* v) {
  return guard([&] {
    need(v != nullptr, "viewer is null");
    v->v.radix_sorter.sort(v->v.radix_sort_indirect_args_buffer);
  });
}

gs3d_status gs3d_viewer_set_depth_stencil(gs3d_viewer* v, int enabled, gs3d_compare compare, int depth_write) {
  return guard([&] {
    need(v != nullptr, "viewer is null");
    v->v.set_depth_stencil(enabled != 0, gs3d::CompareFn(int(compare)), depth_write != 0);
  });
}

gs3d_status gs3d_viewer_render_with_depth(gs3d_viewer* v, gs3d_texture* target, gs3d_texture* depth, int load) {
  return guard([&] {
    need(v != nullptr && target != nullptr && depth != nullptr, "null argument");
    v->v.render_with_depth(target->tex, depth->tex, load != 0);
  });
}

gs3d_status gs3d_viewer_draw(gs3d_viewer* v, gs3d_texture* target) {
  return guard([&] {
    need(v != nullptr && target != nullptr, "null argument");
    v->v.draw(target->tex, false);
  });
}

gs3d_status gs3d_viewer_draw_standalone(gs3d_viewer* v, gs3d_texture* target) {
  return guard([&] {
    need(v != nullptr && target != nullptr, "null argument");
    v->v.draw(target->tex, true);
  });
}

gs3d_status gs3d_viewer_set_blend_mode(gs3d_viewer* v, gs3d_blend_mode mode) {
  return guard([&] {
    need(v != nullptr, "viewer is null");
    need(mode == GS3D_BLEND_FAST || mode == GS3D_BLEND_EXACT, "unknown blend mode");
    v->v.renderer.set_blend(mode == GS3D_BLEND_FAST ? gs3d::kBlendFast : gs3d::kBlendExactUnorm8);
  });
}

gs3d_status gs3d_viewer_set_pair_capacity(gs3d_viewer* v, uint32_t pairs) {
  return guard([&] {
    need(v != nullptr, "viewer is null");
    v->v.renderer.set_capacity(pairs);
  });
}

gs3d_status gs3d_viewer_set_tile_size(gs3d_viewer* v, uint32_t tile_w, uint32_t tile_h) {
  return guard([&] {
    need(v != nullptr, "viewer is null");
    v->v.renderer.set_tile_size(tile_w, tile_h);
  });
}

gs3d_status gs3d_viewer_set_sort_items(gs3d_viewer* v, uint32_t items) {
  return guard([&] {
    need(v != nullptr, "viewer is null");
    v->v.set_sort_items(items);
  });
}

gs3d_status gs3d_viewer_set_profiling(gs3d_viewer* v, int enabled) {
  return guard([&] {
    need(v != nullptr, "viewer is null");
    v->v.timer.set_enabled(enabled != 0);
  });
}

gs3d_status gs3d_viewer_stage_times(gs3d_viewer* v, float ms[GS3D_STAGE_COUNT]) {
  return guard([&] {
    need(v != nullptr && ms != nullptr, "null argument");
    v->v.timer.read(ms);
  });
}

gs3d_status gs3d_viewer_count_fragments(gs3d_viewer* v, gs3d_texture* target, uint64_t* out) {
  return guard([&] {
    need(v != nullptr && target != nullptr && out != nullptr, "null argument");
    v->v.preprocessor.preprocess(v->v.gaussians_buffer.len());
    v->v.radix_sorter.sort(v->v.radix_sort_indirect_args_buffer);
    *out = v->v.renderer.count_fragments(target->tex, v->v.indirect_args_buffer);
  });
}

gs3d_status gs3d_viewer_pair_count(gs3d_viewer* v, uint32_t* out) {
  return guard([&] {
    need(v != nullptr && out != nullptr, "null argument");
    v->v.device.poll();
    *out = v->v.renderer.checked_pair_total();
  });
}

gs3d_status gs3d_viewer_gaussian_count(gs3d_viewer* v, uint32_t* out) {
  return guard([&] {
    need(v != nullptr && out != nullptr, "null argument");
    *out = v->v.gaussians_buffer.len();
  });
}

gs3d_status gs3d_viewer_read_indirect_args(gs3d_viewer* v, uint32_t out[4]) {
  return guard([&] {
    need(v != nullptr && out != nullptr, "null argument");
    v->v.device.poll();
    v->v.indirect_args_buffer.read(out);
  });
}

gs3d_status gs3d_viewer_read_radix_sort_indirect_args(gs3d_viewer* v, uint32_t out[3]) {
  return guard([&] {
    need(v != nullptr && out != nullptr, "null argument");
    v->v.device.poll();
    v->v.radix_sort_indirect_args_buffer.read(out);
  });
}

gs3d_status gs3d_viewer_read_indirect_indices(gs3d_viewer* v, uint32_t* out, uint32_t count) {
  return guard([&] {
    need(v != nullptr && (out != nullptr || count == 0), "null argument");
    need(count <= v->v.indirect_indices_buffer.len(), "count exceeds the buffer");
    v->v.device.poll();
    v->v.indirect_indices_buffer.read(out, count);
  });
}

gs3d_status gs3d_viewer_read_gaussians_depth(gs3d_viewer* v, float* out, uint32_t count) {
  return guard([&] {
    need(v != nullptr && (out != nullptr || count == 0), "null argument");
    need(count <= v->v.gaussians_depth_buffer.len(), "count exceeds the buffer");
    v->v.device.poll();
    v->v.gaussians_depth_buffer.read(out, count);
  });
}

uint32_t gs3d_depth_buffer_len(uint32_t gaussian_count) { return gs3d::depth_buffer_len(gaussian_count); }

gs3d_status gs3d_radix_sorter_new(gs3d_device* dev, uint32_t capacity, gs3d_radix_sorter** out) {
  return guard([&] {
    need(dev != nullptr && out != nullptr, "null argument");
    need(capacity < (1u << 30), "capacity above 2^30");
    *out = nullptr;
    *out = new gs3d_radix_sorter(dev->dev, capacity);
  });
}

void gs3d_radix_sorter_destroy(gs3d_radix_sorter* s) {
  if (s) {
    try {
      s->dev.wait();
    } catch (...) {
    }
  }
  delete s;
}

gs3d_status gs3d_radix_sorter_write(gs3d_radix_sorter* s, const uint32_t* keys, const uint32_t* vals, uint32_t n) {
  return guard([&] {
    need(s != nullptr && (n == 0 || (keys && vals)), "null argument");
    need(n <= s->vals.len(), "n exceeds the sorter's capacity");
    s->dev.wait();
    gs3d::hip_check(hipMemcpy(s->keys.dev(), keys, size_t(n) * 4, hipMemcpyHostToDevice), "upload keys");
    gs3d::hip_check(hipMemcpy(s->vals.dev(), vals, size_t(n) * 4, hipMemcpyHostToDevice), "upload values");
    gs3d::hip_check(hipMemcpy(s->count.get(), &n, 4, hipMemcpyHostToDevice), "upload count");
    s->n = n;
  });
}

gs3d_status gs3d_radix_sorter_sort(gs3d_radix_sorter* s) {
  return guard([&] {
    need(s != nullptr, "sorter is null");
    s->sorter.sort_count(s->count.as<uint32_t>(), 1u);
  });
}

gs3d_status gs3d_radix_sorter_set_items(gs3d_radix_sorter* s, uint32_t items) {
  return guard([&] {
    need(s != nullptr, "sorter is null");
    s->sorter.set_items(items);
  });
}

gs3d_status gs3d_radix_sorter_read(gs3d_radix_sorter* s, uint32_t* keys, uint32_t* vals, uint32_t n) {
  return guard([&] {
    need(s != nullptr && (n == 0 || (keys && vals)), "null argument");
    need(n <= s->n, "n exceeds the written count");
    s->dev.poll();
    gs3d::hip_check(hipMemcpy(keys, s->keys.dev(), size_t(n) * 4, hipMemcpyDeviceToHost), "download keys");
    gs3d::hip_check(hipMemcpy(vals, s->vals.dev(), size_t(n) * 4, hipMemcpyDeviceToHost), "download values");
  });
}

gs3d_status gs3d_radix_sort_noalloc(gs3d_device* dev, const uint32_t* keys_in, const uint32_t* vals_in,
                                    uint32_t* keys_out, uint32_t* vals_out, uint32_t n) {
  return guard([&] {
    need(dev != nullptr, "device is null");
    if (n > GS3D_RADIX_SORT_NOALLOC_MAX_KEYS)
      throw gs3d::Error(GS3D_ERROR_INVALID_ARGUMENT,
                        "n exceeds GS3D_RADIX_SORT_NOALLOC_MAX_KEYS (4194304 pairs), the size of the "
                        "allocation-free sort's fixed scratch");
    if (n == 0) return;
    need(keys_in && vals_in && keys_out && vals_out, "null buffer");
    need(keys_out != vals_out && keys_in != vals_in, "keys and values share a buffer");
    gs3d::Device& d = *dev->dev;
    d.make_current();
    const int r = gs3d::launch_radix_sort_noalloc(d.stream(0), d.num_cus(), d.ordinal(), keys_in, vals_in,
                                                  keys_out, vals_out, n, d.err());
    if (r == 2) throw gs3d::Error(GS3D_ERROR_DEVICE, "allocation-free sort scratch not initialised on this device");
    gs3d::hip_check(hipGetLastError(), "launch allocation-free sort");
  });
}

gs3d_status gs3d_camera_pod_new(const gs3d_camera* camera, uint32_t width, uint32_t height, gs3d_camera_pod* out) {
  return guard([&] {
    need(camera != nullptr && out != nullptr, "null argument");
    need(width > 0 && height > 0, "empty target size");
    gs3d::Camera c;
    for (int i = 0; i < 3; ++i) c.pos[i] = camera->pos[i];
    c.z_start = camera->z_start;
    c.z_end = camera->z_end;
    c.vertical_fov = camera->vertical_fov;
    c.pitch = camera->pitch;
    c.yaw = camera->yaw;
    const gs3d::CameraPod p = gs3d::camera_pod_new(c, width, height);
    std::memcpy(out, &p, sizeof p);
  });
}

gs3d_status gs3d_camera_pod_look_at(const float eye[3], const float target[3], float vertical_fov, float z_near,
                                    float z_far, uint32_t width, uint32_t height, gs3d_camera_pod* out) {
  return guard([&] {
    need(eye && target && out, "null argument");
    need(width > 0 && height > 0, "empty target size");
    float d[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
    const float l = std::sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    need(l > 0.0f, "eye equals target");
    for (float& x : d) x = x / l;
    gs3d::CameraPod p{};
    const float up[3] = {0.0f, 1.0f, 0.0f};
    gs3d::look_to_rh(eye, d, up, p.view);
    gs3d::perspective_rh(vertical_fov, float(width) / float(height), z_near, z_far, p.proj);
    p.size[0] = float(width);
    p.size[1] = float(height);
    std::memcpy(out, &p, sizeof p);
  });
}

gs3d_status gs3d_gaussian_cov3d(const gs3d_gaussian* g, float out[6]) {
  return guard([&] {
    need(g && out, "null argument");
    gs3d::gaussian_cov3d(*g, out);
  });
}

gs3d_status gs3d_pack_half(const gs3d_gaussian* g, uint32_t count, uint16_t* rot_scale, uint16_t* sh) {
  return guard([&] {
    need(count == 0 || (g && rot_scale && sh), "null argument");
    for (uint32_t j = 0; j < count; ++j) {
      for (int k = 0; k < 4; ++k) rot_scale[8 * size_t(j) + k] = gs3d::float_to_half(g[j].rot[k]);
      for (int k = 0; k < 3; ++k) rot_scale[8 * size_t(j) + 4 + k] = gs3d::float_to_half(g[j].scale[k]);
      rot_scale[8 * size_t(j) + 7] = 0;
      for (int k = 0; k < 45; ++k) sh[48 * size_t(j) + k] = gs3d::float_to_half((&g[j].sh[0][0])[k]);
      for (int k = 45; k < 48; ++k) sh[48 * size_t(j) + k] = 0;
    }
  });
}

gs3d_status gs3d_ply_read_file(const char* path, gs3d_gaussian* out, uint64_t capacity, uint64_t* count) {
  return guard([&] {
    need(path != nullptr && count != nullptr, "null argument");
    *count = gs3d::ply_read_file(path, out, capacity);
  });
}

gs3d_status gs3d_ply_read_memory(const void* data, uint64_t bytes, gs3d_gaussian* out, uint64_t capacity,
                                 uint64_t* count) {
  return guard([&] {
    need(data != nullptr && count != nullptr, "null argument");
    *count = gs3d::ply_read(static_cast<const uint8_t*>(data), size_t(bytes), out, capacity);
  });
}

gs3d_status gs3d_spz_read_file(const char* path, gs3d_gaussian* out, uint64_t capacity, uint64_t* count) {
  return guard([&] {
    need(path != nullptr && count != nullptr, "null argument");
    *count = gs3d::spz_read_file(path, out, capacity);
  });
}

gs3d_status gs3d_spz_read_memory(const void* data, uint64_t bytes, gs3d_gaussian* out, uint64_t capacity,
                                 uint64_t* count) {
  return guard([&] {
    need(data != nullptr && count != nullptr, "null argument");
    *count = gs3d::spz_read(static_cast<const uint8_t*>(data), size_t(bytes), out, capacity);
  });
}

gs3d_status gs3d_synth_gaussians(int config, uint64_t seed, uint32_t first, uint32_t count, gs3d_gaussian* out) {
  return guard([&] {
    need(config >= 0 && config <= 2, "unknown synthetic config");
    need(count == 0 || out != nullptr, "out is null");
    gs3d::synth_gaussians(config, seed, first, count, out);
  });
}

gs3d_status gs3d_model_new(gs3d_device* dev, const gs3d_gaussian* gaussians, uint32_t count, gs3d_gaussian_pod pod,
                           gs3d_model** out) {
  return guard([&] {
    need(dev != nullptr && out != nullptr, "null argument");
    need(count == 0 || gaussians != nullptr, "gaussians is null");
    need(count < (1u << 30), "more than 2^30 Gaussians");
    need(pod == GS3D_POD_SINGLE || pod == GS3D_POD_HALF, "unknown gaussian pod");
    *out = nullptr;
    auto m = std::make_unique<gs3d_model>();
    m->keep = dev->dev;
    m->buf = std::make_shared<gs3d::GaussiansBuffer>(*dev->dev, gaussians, count,
                                                     pod == GS3D_POD_SINGLE ? gs3d::kPodSingle : gs3d::kPodHalf);
    *out = m.release();
  });
}

void gs3d_model_destroy(gs3d_model* m) { delete m; }

static std::vector<std::shared_ptr<gs3d::GaussiansBuffer>> model_list(gs3d_model* const* models, uint32_t n) {
  need(n == 0 || models != nullptr, "models is null");
  std::vector<std::shared_ptr<gs3d::GaussiansBuffer>> v;
  for (uint32_t i = 0; i < n; ++i) {
    need(models[i] != nullptr, "null model");
    v.push_back(models[i]->buf);
  }
  return v;
}

static gs3d::BlendMode to_blend(gs3d_blend_mode m) {
  need(m == GS3D_BLEND_FAST || m == GS3D_BLEND_EXACT, "unknown blend mode");
  return m == GS3D_BLEND_FAST ? gs3d::kBlendFast : gs3d::kBlendExactUnorm8;
}

static gs3d::Camera to_camera(const gs3d_camera* camera) {
  gs3d::Camera c;
  for (int i = 0; i < 3; ++i) c.pos[i] = camera->pos[i];
  c.z_start = camera->z_start;
  c.z_end = camera->z_end;
  c.vertical_fov = camera->vertical_fov;
  c.pitch = camera->pitch;
  c.yaw = camera->yaw;
  return c;
}

gs3d_status gs3d_multi_model_viewer_new(gs3d_device* dev, gs3d_texture_format format, gs3d_multi_model_viewer** out) {
  return guard([&] {
    need(dev != nullptr && out != nullptr, "null argument");
    *out = nullptr;
    *out = new gs3d_multi_model_viewer(dev->dev, to_format(format));
  });
}

void gs3d_multi_model_viewer_destroy(gs3d_multi_model_viewer* v) { delete v; }

gs3d_status gs3d_multi_model_viewer_insert_model(gs3d_multi_model_viewer* v, const char* key,
                                                 const gs3d_gaussian* gaussians, uint32_t count,
                                                 gs3d_gaussian_pod pod) {
  return guard([&] {
    need(v != nullptr && key != nullptr, "null argument");
    need(count == 0 || gaussians != nullptr, "gaussians is null");
    need(pod == GS3D_POD_SINGLE || pod == GS3D_POD_HALF, "unknown gaussian pod");
    v->v.insert_model(key, gaussians, count, pod == GS3D_POD_SINGLE ? gs3d::kPodSingle : gs3d::kPodHalf);
  });
}

gs3d_status gs3d_multi_model_viewer_remove_model(gs3d_multi_model_viewer* v, const char* key) {
  return guard([&] {
    need(v != nullptr && key != nullptr, "null argument");
    if (!v->v.remove_model(key))
      throw gs3d::Error(GS3D_ERROR_MODEL_NOT_FOUND, "model with the given key does not exist");
  });
}

gs3d_status gs3d_multi_model_viewer_update_camera(gs3d_multi_model_viewer* v, const gs3d_camera* camera,
                                                  uint32_t width, uint32_t height) {
  return guard([&] {
    need(v != nullptr && camera != nullptr, "null argument");
    need(width > 0 && height > 0, "empty target size");
    v->v.update_camera(to_camera(camera), width, height);
  });
}

gs3d_status gs3d_multi_model_viewer_update_camera_with_pod(gs3d_multi_model_viewer* v, const gs3d_camera_pod* pod) {
  return guard([&] {
    need(v != nullptr && pod != nullptr, "null argument");
    gs3d::CameraPod p;
    std::memcpy(&p, pod, sizeof p);
    v->v.update_camera_with_pod(p);
  });
}

gs3d_status gs3d_multi_model_viewer_update_model_transform(gs3d_multi_model_viewer* v, const char* key,
                                                           const float pos[3], const float rot[4],
                                                           const float scale[3]) {
  return guard([&] {
    need(v != nullptr && key != nullptr && pos && rot && scale, "null argument");
    v->v.update_model_transform(key, pos, rot, scale);
  });
}

gs3d_status gs3d_multi_model_viewer_update_selection(gs3d_multi_model_viewer* v, const char* key,
                                                     const uint32_t* words, uint32_t word_count) {
  return guard([&] {
    need(v != nullptr && key != nullptr && (words != nullptr || word_count == 0), "null argument");
    v->v.model(key).update_selection(words, word_count);
  });
}

gs3d_status gs3d_multi_model_viewer_update_invert_selection(gs3d_multi_model_viewer* v, const char* key,
                                                            int invert) {
  return guard([&] {
    need(v != nullptr && key != nullptr, "null argument");
    v->v.model(key).update_invert_selection(invert != 0);
  });
}

gs3d_status gs3d_multi_model_viewer_update_gaussian_transform(gs3d_multi_model_viewer* v, float size,
                                                              gs3d_display_mode display_mode, uint32_t sh_deg,
                                                              int no_sh0, float max_std_dev) {
  return guard([&] {
    need(v != nullptr, "viewer is null");
    need(display_mode >= GS3D_DISPLAY_SPLAT && display_mode <= GS3D_DISPLAY_POINT, "unknown display mode");
    need(sh_deg <= 3, "SH degree must be 0..3");
    need(max_std_dev > 0.0f, "max standard deviation must be positive");
    v->v.update_gaussian_transform(size, int32_t(display_mode), sh_deg, no_sh0 != 0, max_std_dev);
  });
}

gs3d_status gs3d_multi_model_viewer_set_blend_mode(gs3d_multi_model_viewer* v, gs3d_blend_mode mode) {
  return guard([&] {
    need(v != nullptr, "viewer is null");
    v->v.set_blend(to_blend(mode));
  });
}

gs3d_status gs3d_multi_model_viewer_render(gs3d_multi_model_viewer* v, gs3d_texture* target,
                                           const char* const* keys, uint32_t count) {
  return guard([&] {
    need(v != nullptr && target != nullptr, "null argument");
    need(count == 0 || keys != nullptr, "keys is null");
    std::vector<std::string> k;
    for (uint32_t i = 0; i < count; ++i) {
      need(keys[i] != nullptr, "null key");
      k.emplace_back(keys[i]);
    }
    v->v.render(target->tex, k);
  });
}

gs3d_status gs3d_batch_new(gs3d_device* dev, gs3d_model* const* models, uint32_t model_count, int queues,
                           gs3d_batch** out) {
  return guard([&] {
    need(dev != nullptr && out != nullptr, "null argument");
    need(queues >= 1 && queues <= 16, "queues must be 1..16");
    *out = nullptr;
    *out = new gs3d_batch(dev->dev, model_list(models, model_count), queues);
  });
}

void gs3d_batch_destroy(gs3d_batch* b) { delete b; }

gs3d_status gs3d_batch_set_layout(gs3d_batch* b, const gs3d_view_spec* views, uint32_t count) {
  return guard([&] {
    need(b != nullptr && (count == 0 || views != nullptr), "null argument");
    std::vector<gs3d::ViewSpec> v;
    for (uint32_t i = 0; i < count; ++i) {
      need(views[i].width > 0 && views[i].height > 0, "empty view size");
      v.push_back(gs3d::ViewSpec{views[i].model, views[i].width, views[i].height});
    }
    b->b.set_layout(v);
  });
}

gs3d_status gs3d_batch_autotune(gs3d_batch* b, const gs3d_camera_pod* pods, int reps) {
  return guard([&] {
    need(b != nullptr && (b->b.views() == 0 || pods != nullptr), "null argument");
    b->b.autotune(reinterpret_cast<const gs3d::CameraPod*>(pods), reps);
  });
}

gs3d_status gs3d_batch_render(gs3d_batch* b, const gs3d_camera_pod* pods, int use_graph) {
  return guard([&] {
    need(b != nullptr && (b->b.views() == 0 || pods != nullptr), "null argument");
    b->b.render(reinterpret_cast<const gs3d::CameraPod*>(pods), use_graph != 0);
  });
}

gs3d_status gs3d_batch_finish(gs3d_batch* b) {
  return guard([&] {
    need(b != nullptr, "batch is null");
    b->b.finish();
  });
}

gs3d_status gs3d_batch_read_target(gs3d_batch* b, uint32_t view, void* dst, size_t bytes) {
  return guard([&] {
    need(b != nullptr && dst != nullptr, "null argument");
    need(view < b->b.views(), "no such view");
    b->b.finish();
    b->b.target(view).read(dst, bytes);
  });
}

gs3d_status gs3d_batch_tile_size(gs3d_batch* b, uint32_t view, uint32_t* tile_w, uint32_t* tile_h) {
  return guard([&] {
    need(b != nullptr && tile_w && tile_h, "null argument");
    need(view < b->b.views(), "no such view");
    *tile_w = b->b.viewer(view).renderer.tile_w();
    *tile_h = b->b.viewer(view).renderer.tile_h();
  });
}

gs3d_status gs3d_batch_sort_items(gs3d_batch* b, uint32_t view, uint32_t* items) {
  return guard([&] {
    need(b != nullptr && items, "null argument");
    need(view < b->b.views(), "no such view");
    *items = b->b.viewer(view).sort_items();
  });
}

static gs3d::ServeRequest to_request(const gs3d_request& r) {
  gs3d::ServeRequest q;
  q.model = r.model;
  q.width = r.width;
  q.height = r.height;
  q.arrival_s = r.arrival_s;
  std::memcpy(&q.pod, &r.pod, sizeof q.pod);
  return q;
}

gs3d_status gs3d_server_new(gs3d_device* dev, gs3d_model* const* models, uint32_t model_count, int queues,
                            uint64_t hbm_budget, gs3d_server** out) {
  return guard([&] {
    need(dev != nullptr && out != nullptr, "null argument");
    need(queues >= 1 && queues <= 16, "queues must be 1..16");
    *out = nullptr;
    *out = new gs3d_server(dev->dev, model_list(models, model_count), queues, hbm_budget);
  });
}

void gs3d_server_destroy(gs3d_server* s) { delete s; }

gs3d_status gs3d_server_run(gs3d_server* s, const gs3d_request* requests, uint32_t count, uint8_t* const* images,
                            gs3d_serve_stats* out) {
  return guard([&] {
    need(s != nullptr && out != nullptr && (count == 0 || requests != nullptr), "null argument");
    std::vector<gs3d::ServeRequest> r;
    r.reserve(count);
    for (uint32_t i = 0; i < count; ++i) {
      need(requests[i].width > 0 && requests[i].height > 0, "empty request size");
      need(i == 0 || requests[i].arrival_s >= requests[i - 1].arrival_s, "requests must be sorted by arrival");
      r.push_back(to_request(requests[i]));
    }
    std::vector<uint8_t*> img;
    if (images) img.assign(images, images + count);
    const gs3d::ServeStats st = s->s.run(r, images ? &img : nullptr);
    out->requests = st.requests;
    out->completed = st.completed;
    out->seconds = st.seconds;
    out->frames_per_s = st.frames_per_s;
    out->p50_ms = st.p50_ms;
    out->p99_ms = st.p99_ms;
    out->mean_ms = st.mean_ms;
    out->max_ms = st.max_ms;
    out->mean_in_flight = st.mean_in_flight;
    out->peak_hbm_bytes = st.peak_hbm_bytes;
    out->deferred_for_memory = st.deferred_for_memory;
  });
}

gs3d_status gs3d_server_render_alone(gs3d_server* s, const gs3d_request* request, uint8_t* out) {
  return guard([&] {
    need(s != nullptr && request != nullptr && out != nullptr, "null argument");
    s->s.render_alone(to_request(*request), out);
  });
}

}  // extern "C"
