// SPZ ingestion: Niantic's compressed 3D Gaussian Splatting format, the
// second source the reference's examples try (GaussiansSource::Spz,
// examples/simple.rs:157-160), read through wgpu-3dgs-core, which is not part
// of the reference tree.
//
// Layout (a gzip stream): a 16-byte header {magic "NGSP" = 0x5053474e,
// version 2 or 3, point count, SH degree, fractional bits, flags, reserved},
// then, each as one array over all points: positions (3 x 24-bit signed
// fixed point with `fractional bits`), alphas (u8, sigmoid(opacity) * 255),
// colours (3 x u8, dc = (c / 255 - 0.5) / 0.15), scales (3 x u8, log scale =
// s / 16 - 10), rotations (version 2: 3 x u8 xyz = r / 127.5 - 1 with w >= 0
// rebuilt; version 3: the largest component's index in 2 bits and the other
// three as 9-bit magnitudes with a sign bit, scaled by 1 / sqrt(2)), and SH
// coefficients (u8, (s - 128) / 128, per point: coefficient-major, RGB
// innermost). No coordinate-system conversion is applied.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gs3d.h"
#include "viewer.hpp"

namespace gs3d {
namespace {

[[noreturn]] void bad(const std::string& what) { throw Error(GS3D_ERROR_INVALID_ARGUMENT, "SPZ: " + what); }

constexpr uint32_t kMagic = 0x5053474eu;
constexpr float kColorScale = 0.15f;
constexpr double kC0 = 0.28209479177387814;

// Inflates a gzip stream into `out` until it holds `want` bytes or the stream
// ends (no more than `want` bytes are kept, so a header cannot make the
// reader inflate more than its point count needs). The caller may call again
// with a larger `want` on the same stream.
class Gunzip {
 public:
  Gunzip(const uint8_t* data, size_t size) {
    if (inflateInit2(&zs_, 16 + MAX_WBITS) != Z_OK) bad("cannot initialise zlib");
    zs_.next_in = const_cast<Bytef*>(data);
    zs_.avail_in = uInt(size);
    if (size > size_t(uInt(-1))) bad("compressed stream above 4 GiB");
  }
  ~Gunzip() { inflateEnd(&zs_); }
  Gunzip(const Gunzip&) = delete;
  Gunzip& operator=(const Gunzip&) = delete;

  const std::vector<uint8_t>& fill(size_t want) {
    uint8_t chunk[1 << 16];
    while (!ended_ && out_.size() < want) {
      zs_.next_out = chunk;
      zs_.avail_out = uInt(std::min(sizeof(chunk), want - out_.size()));
      const int rc = inflate(&zs_, Z_NO_FLUSH);
      if (rc != Z_OK && rc != Z_STREAM_END) bad("corrupt gzip stream");
      const size_t got = size_t(zs_.next_out - chunk);
      if (rc == Z_OK && got == 0 && zs_.avail_in == 0) bad("truncated gzip stream");
      out_.insert(out_.end(), chunk, chunk + got);
      ended_ = rc == Z_STREAM_END;
    }
    return out_;
  }

 private:
  z_stream zs_{};
  std::vector<uint8_t> out_;
  bool ended_ = false;
};

struct Header {
  uint32_t magic, version, points;
  uint8_t sh_degree, fractional_bits, flags, reserved;
};
static_assert(sizeof(Header) == 16, "SPZ header is 16 bytes");

int sh_dim(int degree) { return degree == 0 ? 0 : degree == 1 ? 3 : degree == 2 ? 8 : 15; }

uint8_t unit_to_u8(double x) {
  if (!(x > 0.0)) return 0;
  if (x >= 1.0) return 255;
  return uint8_t(std::floor(x * 255.0 + 0.5));
}

}  // namespace

uint64_t spz_read(const uint8_t* data, size_t size, gs3d_gaussian* out, uint64_t capacity) {
  if (size < 2 || data[0] != 0x1f || data[1] != 0x8b) bad("not a gzip stream");
  Gunzip gz(data, size);
  if (gz.fill(sizeof(Header)).size() < sizeof(Header)) bad("file is shorter than its header");
  Header h;
  std::memcpy(&h, gz.fill(sizeof(Header)).data(), sizeof(h));
  if (h.magic != kMagic) bad("bad magic");
  if (h.version != 2 && h.version != 3) bad("unsupported version " + std::to_string(h.version));
  if (h.sh_degree > 3) bad("SH degree above 3");
  // positions are 24-bit fixed point: more fractional bits than that is not
  // a position format (and a shift by 32 or more is undefined)
  if (h.fractional_bits > 24) bad("fractional bits above 24");
  const uint64_t n = h.points;
  if (out == nullptr) return n;
  if (capacity < n) bad("output capacity " + std::to_string(capacity) + " < " + std::to_string(n));
  const int dim = sh_dim(h.sh_degree);
  const size_t rot_bytes = h.version == 2 ? 3 : 4;
  const size_t need = sizeof(Header) + n * (9 + 1 + 3 + 3 + rot_bytes + size_t(dim) * 3);
  // inflate exactly what the point count needs, no more
  const std::vector<uint8_t>& raw = gz.fill(need);
  if (raw.size() < need) bad("data is shorter than its point count");
  const uint8_t* pos = raw.data() + sizeof(Header);
  const uint8_t* alpha = pos + n * 9;
  const uint8_t* color = alpha + n;
  const uint8_t* scale = color + n * 3;
  const uint8_t* rot = scale + n * 3;
  const uint8_t* sh = rot + n * rot_bytes;
  const float inv_fixed = 1.0f / float(1u << h.fractional_bits);
  for (uint64_t i = 0; i < n; ++i) {
    gs3d_gaussian& g = out[i];
    for (int c = 0; c < 3; ++c) {
      const uint8_t* p = pos + i * 9 + c * 3;
      int32_t v = int32_t(uint32_t(p[0]) | (uint32_t(p[1]) << 8) | (uint32_t(p[2]) << 16));
      if (v & 0x800000) v |= int32_t(0xFF000000u);  // sign-extend 24 bits
      g.pos[c] = float(v) * inv_fixed;
      g.scale[c] = std::exp(float(scale[i * 3 + c]) / 16.0f - 10.0f);
      const double dc = (double(color[i * 3 + c]) / 255.0 - 0.5) / kColorScale;
      g.color[c] = unit_to_u8(0.5 + kC0 * dc);
    }
    g.color[3] = alpha[i];
    float q[4];  // x, y, z, w
    if (h.version == 2) {
      for (int c = 0; c < 3; ++c) q[c] = float(rot[i * 3 + c]) / 127.5f - 1.0f;
      q[3] = std::sqrt(std::fmax(0.0f, 1.0f - (q[0] * q[0] + q[1] * q[1] + q[2] * q[2])));
    } else {
      const uint8_t* r = rot + i * 4;
      uint32_t comp = uint32_t(r[0]) | (uint32_t(r[1]) << 8) | (uint32_t(r[2]) << 16) | (uint32_t(r[3]) << 24);
      const uint32_t largest = comp >> 30;
      constexpr uint32_t kMask = (1u << 9) - 1u;
      float sum = 0.0f;
      for (int c = 3; c >= 0; --c) {
        if (uint32_t(c) == largest) continue;
        const uint32_t mag = comp & kMask;
        const bool neg = ((comp >> 9) & 1u) != 0u;
        comp >>= 10;
        float v = 0.70710678118654752f * float(mag) / float(kMask);
        if (neg) v = -v;
        q[c] = v;
        sum += v * v;
      }
      q[largest] = std::sqrt(std::fmax(0.0f, 1.0f - sum));
    }
    const float len = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int c = 0; c < 4; ++c) g.rot[c] = len > 0.0f ? q[c] / len : (c == 3 ? 1.0f : 0.0f);
    for (int k = 0; k < 15; ++k)
      for (int c = 0; c < 3; ++c)
        g.sh[k][c] = k < dim ? (float(sh[(i * size_t(dim) + size_t(k)) * 3 + c]) - 128.0f) / 128.0f : 0.0f;
  }
  return n;
}

uint64_t spz_read_file(const char* path, gs3d_gaussian* out, uint64_t capacity) {
  std::FILE* f = std::fopen(path, "rb");
  if (!f) throw Error(GS3D_ERROR_INVALID_ARGUMENT, std::string("cannot open ") + path);
  std::vector<uint8_t> buf;
  if (std::fseek(f, 0, SEEK_END) == 0) {
    const long n = std::ftell(f);
    if (n > 0) buf.resize(size_t(n));
    std::fseek(f, 0, SEEK_SET);
  }
  const size_t got = buf.empty() ? 0 : std::fread(buf.data(), 1, buf.size(), f);
  std::fclose(f);
  if (got != buf.size()) throw Error(GS3D_ERROR_INVALID_ARGUMENT, std::string("cannot read ") + path);
  return spz_read(buf.data(), buf.size(), out, capacity);
}

}  // namespace gs3d
