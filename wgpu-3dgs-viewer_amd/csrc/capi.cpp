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
#include "math.hpp"
#include "viewer.hpp"

struct gs3d_device {
  gs3d::Device dev;
  explicit gs3d_device(int ordinal) : dev(ordinal) {}
};

struct gs3d_texture {
  gs3d::Texture tex;
  gs3d_texture(gs3d::Device& d, uint32_t w, uint32_t h, gs3d::TextureFormat f) : tex(d, w, h, f) {}
};

struct gs3d_viewer {
  gs3d::Viewer v;
  gs3d_viewer(gs3d::Device& d, gs3d::TextureFormat f, const gs3d_gaussian* g, uint32_t n, gs3d::PodKind p)
      : v(d, f, g, n, p) {}
};

struct gs3d_radix_sorter {
  gs3d::Device& dev;
  gs3d::IndirectIndicesBuffer vals;
  gs3d::GaussiansDepthBuffer keys;
  gs3d::DeviceBuffer count;
  gs3d::RadixSorter sorter;
  uint32_t n = 0;
  gs3d_radix_sorter(gs3d::Device& d, uint32_t cap)
      : dev(d), vals(d, cap), keys(d, cap), count(d, 64), sorter(d, keys, vals) {}
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

gs3d::TextureFormat to_format(gs3d_texture_format f) {
  need(f == GS3D_FORMAT_RGBA8_UNORM || f == GS3D_FORMAT_RGBA32_FLOAT, "unknown texture format");
  return f == GS3D_FORMAT_RGBA8_UNORM ? gs3d::kFormatRgba8Unorm : gs3d::kFormatRgba32Float;
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
    dev->dev.poll();
  });
}

gs3d_status gs3d_device_set_buffer_limit(gs3d_device* dev, uint64_t bytes) {
  return guard([&] {
    need(dev != nullptr, "device is null");
    dev->dev.set_buffer_limit(bytes);
  });
}

gs3d_status gs3d_device_compute_units(gs3d_device* dev, int* out) {
  return guard([&] {
    need(dev != nullptr && out != nullptr, "null argument");
    *out = dev->dev.num_cus();
  });
}

gs3d_status gs3d_texture_create(gs3d_device* dev, uint32_t width, uint32_t height, gs3d_texture_format format,
                                gs3d_texture** out) {
  return guard([&] {
    need(dev != nullptr && out != nullptr, "null argument");
    *out = nullptr;
    *out = new gs3d_texture(dev->dev, width, height, to_format(format));
  });
}

void gs3d_texture_destroy(gs3d_texture* tex) { delete tex; }

gs3d_status gs3d_texture_read(gs3d_texture* tex, void* dst, size_t bytes) {
  return guard([&] {
    need(tex != nullptr && dst != nullptr, "null argument");
    tex->tex.read(dst, bytes);
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

gs3d_status gs3d_viewer_draw(gs3d_viewer* v, gs3d_texture* target) {
  return guard([&] {
    need(v != nullptr && target != nullptr, "null argument");
    v->v.renderer.render(target->tex, v->v.indirect_args_buffer, false);
  });
}

gs3d_status gs3d_viewer_draw_standalone(gs3d_viewer* v, gs3d_texture* target) {
  return guard([&] {
    need(v != nullptr && target != nullptr, "null argument");
    v->v.renderer.render(target->tex, v->v.indirect_args_buffer, true);
  });
}

gs3d_status gs3d_viewer_set_blend_mode(gs3d_viewer* v, gs3d_blend_mode mode) {
  return guard([&] {
    need(v != nullptr, "viewer is null");
    need(mode == GS3D_BLEND_FAST || mode == GS3D_BLEND_EXACT, "unknown blend mode");
    v->v.renderer.set_blend(mode == GS3D_BLEND_FAST ? gs3d::kBlendFast : gs3d::kBlendExactUnorm8);
  });
}

gs3d_status gs3d_viewer_set_tile_size(gs3d_viewer* v, uint32_t tile_w, uint32_t tile_h) {
  return guard([&] {
    need(v != nullptr, "viewer is null");
    v->v.renderer.set_tile_size(tile_w, tile_h);
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

gs3d_status gs3d_radix_sorter_read(gs3d_radix_sorter* s, uint32_t* keys, uint32_t* vals, uint32_t n) {
  return guard([&] {
    need(s != nullptr && (n == 0 || (keys && vals)), "null argument");
    need(n <= s->n, "n exceeds the written count");
    s->dev.poll();
    gs3d::hip_check(hipMemcpy(keys, s->keys.dev(), size_t(n) * 4, hipMemcpyDeviceToHost), "download keys");
    gs3d::hip_check(hipMemcpy(vals, s->vals.dev(), size_t(n) * 4, hipMemcpyDeviceToHost), "download values");
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

gs3d_status gs3d_synth_gaussians(int config, uint64_t seed, uint32_t first, uint32_t count, gs3d_gaussian* out) {
  return guard([&] {
    need(config >= 0 && config <= 2, "unknown synthetic config");
    need(count == 0 || out != nullptr, "out is null");
    gs3d::synth_gaussians(config, seed, first, count, out);
  });
}

}  // extern "C"
