// Synthetic scenes for tests and benchmarks. Counter-based: every random
// number is splitmix64 of (seed, Gaussian index, field), so any slice of a
// scene can be generated independently and identically on any host.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "viewer.hpp"

namespace gs3d {
namespace {

constexpr double kPi = 3.14159265358979323846;

inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Rng {
  uint64_t key;
  uint64_t i;
  double uniform(uint32_t field) const {
    const uint64_t h = splitmix64(key ^ splitmix64(i * 256u + field));
    return double(h >> 11) * (1.0 / 9007199254740992.0);
  }
  // Box-Muller on fields (f, f+1)
  double normal(uint32_t field) const {
    const double u1 = uniform(field), u2 = uniform(field + 1);
    return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(2.0 * kPi * u2);
  }
};

inline uint8_t to_u8(double x) {
  if (!(x > 0.0)) return 0;
  if (x >= 1.0) return 255;
  return uint8_t(std::floor(x * 255.0 + 0.5));
}

void fill_common(const Rng& r, double log_mean, double log_std, gs3d_gaussian& g) {
  for (int a = 0; a < 3; ++a) g.scale[a] = float(std::exp(log_mean + log_std * r.normal(10 + 2 * a)));
  // Shoemake uniform rotation, stored (x, y, z, w)
  const double u1 = r.uniform(20), u2 = r.uniform(21), u3 = r.uniform(22);
  const double s1 = std::sqrt(1.0 - u1), s2 = std::sqrt(u1);
  g.rot[0] = float(s1 * std::sin(2.0 * kPi * u2));
  g.rot[1] = float(s1 * std::cos(2.0 * kPi * u2));
  g.rot[2] = float(s2 * std::sin(2.0 * kPi * u3));
  g.rot[3] = float(s2 * std::cos(2.0 * kPi * u3));
  // opacity = sigmoid(N(0,1)); colour = 0.5 + SH_C0 * dc with dc ~ U(0,1)
  const double op = 1.0 / (1.0 + std::exp(-r.normal(24)));
  for (int c = 0; c < 3; ++c) g.color[c] = to_u8(0.5 + 0.28209479177387814 * r.uniform(26 + c));
  g.color[3] = to_u8(op);
  // SH bands 1..3 ~ N(0, 0.1)
  for (int k = 0; k < 45; ++k) (&g.sh[0][0])[k] = float(0.1 * r.normal(40 + 2 * k));
}

}  // namespace

static void synth_range(int config, uint64_t seed, uint32_t first, uint32_t count, gs3d_gaussian* out);

void synth_gaussians(int config, uint64_t seed, uint32_t first, uint32_t count, gs3d_gaussian* out) {
  // record i depends only on (seed, i): split the range over host threads
  const uint32_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const uint32_t chunk = 1u << 16;
  if (count < 2 * chunk || hw == 1) {
    synth_range(config, seed, first, count, out);
    return;
  }
  std::atomic<uint32_t> next{0};
  std::vector<std::thread> pool;
  for (uint32_t t = 0; t < hw; ++t)
    pool.emplace_back([&]() {
      for (;;) {
        const uint32_t b = next.fetch_add(chunk);
        if (b >= count) break;
        synth_range(config, seed, first + b, std::min(chunk, count - b), out + b);
      }
    });
  for (auto& th : pool) th.join();
}

static void synth_range(int config, uint64_t seed, uint32_t first, uint32_t count, gs3d_gaussian* out) {
  const uint64_t key = splitmix64(seed * 0x2545F4914F6CDD1Dull + 0x1234567ull);
  // config 2: 64 anisotropic clusters, drawn from a stream of their own
  double centre[64][3], sd[64][3];
  if (config == 2) {
    for (int c = 0; c < 64; ++c) {
      const Rng r{key ^ 0xC1C1C1C1ull, uint64_t(c)};
      for (int a = 0; a < 3; ++a) {
        centre[c][a] = -20.0 + 40.0 * r.uniform(a);
        sd[c][a] = 0.2 + 2.8 * r.uniform(3 + a);
      }
    }
  }
  for (uint32_t j = 0; j < count; ++j) {
    const uint64_t i = uint64_t(first) + j;
    const Rng r{key, i};
    gs3d_gaussian& g = out[j];
    if (config == 0 || config == 1) {
      const double a = config == 0 ? 1.0 : 4.0;
      for (int k = 0; k < 3; ++k) g.pos[k] = float(-a + 2.0 * a * r.uniform(k));
      fill_common(r, -4.0, 0.5, g);
    } else {
      if (r.uniform(0) < 0.2) {
        // jittered ground plane y = -2
        g.pos[0] = float(-20.0 + 40.0 * r.uniform(1));
        g.pos[1] = float(-2.0 + 0.05 * r.normal(4));
        g.pos[2] = float(-20.0 + 40.0 * r.uniform(2));
      } else {
        int c = int(r.uniform(1) * 64.0);
        if (c > 63) c = 63;
        for (int k = 0; k < 3; ++k) g.pos[k] = float(centre[c][k] + sd[c][k] * r.normal(4 + 2 * k));
      }
      fill_common(r, -5.0, 0.7, g);
    }
  }
}

}  // namespace gs3d
