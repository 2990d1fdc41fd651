// PLY ingestion: the standard 3D Gaussian Splatting vertex layout (x y z,
// optional normals, f_dc_0..2, f_rest_0..(3m-1), opacity, scale_0..2,
// rot_0..3), binary little-endian or ASCII, into gs3d_gaussian records.
//
// The reference reads files through wgpu-3dgs-core 0.6.0's
// Gaussians::read_from_file (examples/simple.rs:157-160, multi_model.rs:133);
// that crate is not part of the reference tree. The conversion below follows
// the published 3DGS conventions: rot_0 is the quaternion's w (normalised),
// scales are stored as logarithms, opacity as a logit, colour as the degree-0
// SH coefficient (rgb = 0.5 + C0 * f_dc), and f_rest is channel-major
// (all coefficients of red, then green, then blue).
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

enum class Scalar { I8, U8, I16, U16, I32, U32, F32, F64 };

size_t scalar_size(Scalar s) {
  switch (s) {
    case Scalar::I8: case Scalar::U8: return 1;
    case Scalar::I16: case Scalar::U16: return 2;
    case Scalar::I32: case Scalar::U32: case Scalar::F32: return 4;
    case Scalar::F64: return 8;
  }
  return 0;
}

bool parse_scalar(const std::string& t, Scalar& s) {
  static const struct { const char* name; Scalar s; } kNames[] = {
      {"char", Scalar::I8},    {"int8", Scalar::I8},     {"uchar", Scalar::U8},  {"uint8", Scalar::U8},
      {"short", Scalar::I16},  {"int16", Scalar::I16},   {"ushort", Scalar::U16}, {"uint16", Scalar::U16},
      {"int", Scalar::I32},    {"int32", Scalar::I32},   {"uint", Scalar::U32},  {"uint32", Scalar::U32},
      {"float", Scalar::F32},  {"float32", Scalar::F32}, {"double", Scalar::F64}, {"float64", Scalar::F64}};
  for (const auto& n : kNames)
    if (t == n.name) { s = n.s; return true; }
  return false;
}

double read_scalar(const uint8_t* p, Scalar s) {
  switch (s) {
    case Scalar::I8: return double(*reinterpret_cast<const int8_t*>(p));
    case Scalar::U8: return double(*p);
    case Scalar::I16: { int16_t v; std::memcpy(&v, p, 2); return v; }
    case Scalar::U16: { uint16_t v; std::memcpy(&v, p, 2); return v; }
    case Scalar::I32: { int32_t v; std::memcpy(&v, p, 4); return v; }
    case Scalar::U32: { uint32_t v; std::memcpy(&v, p, 4); return v; }
    case Scalar::F32: { float v; std::memcpy(&v, p, 4); return v; }
    case Scalar::F64: { double v; std::memcpy(&v, p, 8); return v; }
  }
  return 0.0;
}

struct Property {
  std::string name;
  Scalar type;
  size_t offset;  // within a binary vertex row
};

struct Header {
  bool ascii = false;
  uint64_t vertices = 0;
  std::vector<Property> props;  // of the vertex element
  size_t row_bytes = 0;
  size_t body = 0;  // byte offset of the body
  // bytes of elements declared before "vertex" (skipped; must be fixed-size)
  size_t pre_bytes = 0;
};

[[noreturn]] void bad(const std::string& what) { throw Error(GS3D_ERROR_INVALID_ARGUMENT, "PLY: " + what); }

Header parse_header(const uint8_t* data, size_t size) {
  Header h;
  size_t pos = 0;
  auto next_line = [&](std::string& line) {
    if (pos >= size) bad("unexpected end of header");
    size_t e = pos;
    while (e < size && data[e] != '\n') ++e;
    if (e >= size) bad("unterminated header line");
    line.assign(reinterpret_cast<const char*>(data + pos), e - pos);
    if (!line.empty() && line.back() == '\r') line.pop_back();
    pos = e + 1;
  };
  std::string line;
  next_line(line);
  if (line != "ply") bad("missing 'ply' magic");
  bool in_vertex = false, seen_vertex = false, fmt = false;
  uint64_t other_count = 0;
  size_t other_row = 0;
  bool other_fixed = true;
  auto close_other = [&] {
    if (other_count) {
      if (!other_fixed) bad("variable-size element before 'vertex' is not supported");
      h.pre_bytes += size_t(other_count) * other_row;
    }
    other_count = 0;
    other_row = 0;
    other_fixed = true;
  };
  for (;;) {
    next_line(line);
    char a[64] = {0}, b[64] = {0}, c[64] = {0}, d[64] = {0};
    const int n = std::sscanf(line.c_str(), "%63s %63s %63s %63s", a, b, c, d);
    if (n <= 0) continue;
    const std::string k = a;
    if (k == "end_header") break;
    if (k == "comment" || k == "obj_info") continue;
    if (k == "format") {
      const std::string f = b;
      if (f == "ascii") h.ascii = true;
      else if (f != "binary_little_endian") bad("unsupported format " + f);
      fmt = true;
    } else if (k == "element") {
      if (n < 3) bad("malformed element line");
      if (in_vertex) in_vertex = false;
      if (!seen_vertex) close_other();
      if (std::string(b) == "vertex") {
        in_vertex = true;
        seen_vertex = true;
        h.vertices = std::strtoull(c, nullptr, 10);
      } else if (!seen_vertex) {
        other_count = std::strtoull(c, nullptr, 10);
      }
    } else if (k == "property") {
      if (std::string(b) == "list") {
        if (in_vertex) bad("list property in the vertex element");
        if (!seen_vertex) other_fixed = false;
        continue;
      }
      Scalar s;
      if (!parse_scalar(b, s)) bad(std::string("unknown property type ") + b);
      if (in_vertex) {
        h.props.push_back({c, s, h.row_bytes});
        h.row_bytes += scalar_size(s);
      } else if (!seen_vertex) {
        other_row += scalar_size(s);
      }
    } else {
      bad("unknown header keyword " + k);
    }
  }
  if (!fmt) bad("missing format line");
  if (!seen_vertex) bad("no vertex element");
  if (h.ascii && h.pre_bytes) bad("ASCII files with elements before 'vertex' are not supported");
  h.body = pos;
  return h;
}

struct Columns {
  int pos[3], dc[3], opacity, scale[3], rot[4];
  std::vector<int> rest;  // f_rest_0.. in index order
};

Columns find_columns(const Header& h) {
  Columns c;
  auto find = [&](const std::string& name, bool required) {
    for (size_t i = 0; i < h.props.size(); ++i)
      if (h.props[i].name == name) return int(i);
    if (required) bad("missing vertex property " + name);
    return -1;
  };
  const char* xyz[3] = {"x", "y", "z"};
  for (int i = 0; i < 3; ++i) {
    c.pos[i] = find(xyz[i], true);
    c.dc[i] = find("f_dc_" + std::to_string(i), true);
    c.scale[i] = find("scale_" + std::to_string(i), true);
  }
  for (int i = 0; i < 4; ++i) c.rot[i] = find("rot_" + std::to_string(i), true);
  c.opacity = find("opacity", true);
  for (int i = 0;; ++i) {
    const int k = find("f_rest_" + std::to_string(i), false);
    if (k < 0) break;
    c.rest.push_back(k);
  }
  if (c.rest.size() % 3 != 0 || c.rest.size() > 45) bad("f_rest count must be 3 x (0, 3, 8 or 15)");
  const size_t m = c.rest.size() / 3;
  if (m != 0 && m != 3 && m != 8 && m != 15) bad("f_rest count must be 3 x (0, 3, 8 or 15)");
  return c;
}

uint8_t unit_to_u8(double x) {
  if (!(x > 0.0)) return 0;
  if (x >= 1.0) return 255;
  return uint8_t(std::floor(x * 255.0 + 0.5));
}

void convert(const Columns& c, const double* v, gs3d_gaussian& g) {
  constexpr double kC0 = 0.28209479177387814;
  for (int i = 0; i < 3; ++i) g.pos[i] = float(v[c.pos[i]]);
  // rot_0 = w; stored as (x, y, z, w), normalised
  double q[4] = {v[c.rot[1]], v[c.rot[2]], v[c.rot[3]], v[c.rot[0]]};
  const double len = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  for (int i = 0; i < 4; ++i) g.rot[i] = len > 0.0 ? float(q[i] / len) : (i == 3 ? 1.0f : 0.0f);
  for (int i = 0; i < 3; ++i) g.scale[i] = float(std::exp(v[c.scale[i]]));
  for (int i = 0; i < 3; ++i) g.color[i] = unit_to_u8(0.5 + kC0 * v[c.dc[i]]);
  g.color[3] = unit_to_u8(1.0 / (1.0 + std::exp(-v[c.opacity])));
  const size_t m = c.rest.size() / 3;
  for (int k = 0; k < 15; ++k)
    for (int ch = 0; ch < 3; ++ch) g.sh[k][ch] = size_t(k) < m ? float(v[c.rest[ch * m + k]]) : 0.0f;
}

}  // namespace

uint64_t ply_read(const uint8_t* data, size_t size, gs3d_gaussian* out, uint64_t capacity) {
  const Header h = parse_header(data, size);
  if (out == nullptr) return h.vertices;
  if (capacity < h.vertices) bad("output capacity " + std::to_string(capacity) + " < " + std::to_string(h.vertices));
  const Columns c = find_columns(h);
  std::vector<double> v(h.props.size());
  if (!h.ascii) {
    const size_t start = h.body + h.pre_bytes;
    if (h.row_bytes == 0 || start > size || (size - start) / h.row_bytes < h.vertices)
      bad("file is shorter than its vertex element");
    const uint8_t* p = data + start;
    for (uint64_t i = 0; i < h.vertices; ++i, p += h.row_bytes) {
      for (size_t k = 0; k < h.props.size(); ++k) v[k] = read_scalar(p + h.props[k].offset, h.props[k].type);
      convert(c, v.data(), out[i]);
    }
  } else {
    std::string body(reinterpret_cast<const char*>(data + h.body), size - h.body);
    const char* s = body.c_str();
    for (uint64_t i = 0; i < h.vertices; ++i) {
      for (size_t k = 0; k < h.props.size(); ++k) {
        char* end = nullptr;
        v[k] = std::strtod(s, &end);
        if (end == s) bad("ASCII body ends early at vertex " + std::to_string(i));
        s = end;
      }
      convert(c, v.data(), out[i]);
    }
  }
  return h.vertices;
}

uint64_t ply_read_file(const char* path, gs3d_gaussian* out, uint64_t capacity) {
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
  return ply_read(buf.data(), buf.size(), out, capacity);
}

}  // namespace gs3d
