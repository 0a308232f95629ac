#include "inputs.hpp"

#include <cmath>

namespace redwood::inputs {

namespace {
inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
inline double unit_open(uint32_t b) { return (static_cast<double>(b) + 0.5) * (1.0 / 4294967296.0); }
constexpr float kTop = 1023.99993896484375f;  // nextafterf(1024, 0)
}  // namespace

uint32_t bits32(uint64_t seed, uint64_t stream, uint64_t counter) {
  const uint64_t x = seed * 0x9e3779b97f4a7c15ull + stream * 0xd1b54a32d192ed03ull + counter;
  return static_cast<uint32_t>(mix64(x) >> 32);
}

float coord(uint64_t seed, uint64_t stream, uint64_t counter) {
  return static_cast<float>(bits32(seed, stream, counter) >> 8) * (1.0f / 16384.0f);
}

void uniform(uint64_t seed, size_t n, float* xyzw) {
  for (size_t i = 0; i < n; ++i) {
    xyzw[4 * i + 0] = coord(seed, 0, 3 * i + 0);
    xyzw[4 * i + 1] = coord(seed, 0, 3 * i + 1);
    xyzw[4 * i + 2] = coord(seed, 0, 3 * i + 2);
    xyzw[4 * i + 3] = 0.0f;
  }
}

void gaussian_mixture(uint64_t seed, size_t n, uint32_t n_centers, double sigma, float* xyzw) {
  std::vector<double> c(3ull * n_centers);
  for (uint32_t k = 0; k < n_centers; ++k)
    for (int a = 0; a < 3; ++a) c[3 * k + a] = coord(seed, 1, 3ull * k + a);
  const double two_pi = 6.283185307179586476925286766559;
  for (size_t i = 0; i < n; ++i) {
    const uint32_t k = bits32(seed, 2, i) % n_centers;
    const double u1 = unit_open(bits32(seed, 3, 4 * i + 0));
    const double u2 = unit_open(bits32(seed, 3, 4 * i + 1));
    const double u3 = unit_open(bits32(seed, 3, 4 * i + 2));
    const double u4 = unit_open(bits32(seed, 3, 4 * i + 3));
    const double r1 = std::sqrt(-2.0 * std::log(u1));
    const double r2 = std::sqrt(-2.0 * std::log(u3));
    const double z[3] = {r1 * std::cos(two_pi * u2), r1 * std::sin(two_pi * u2),
                         r2 * std::cos(two_pi * u4)};
    for (int a = 0; a < 3; ++a) {
      float v = static_cast<float>(c[3 * k + a] + sigma * z[a]);
      if (!(v >= 0.0f)) v = 0.0f;
      if (v > kTop) v = kTop;
      xyzw[4 * i + a] = v;
    }
    xyzw[4 * i + 3] = 0.0f;
  }
}

std::vector<uint32_t> log_uniform_sizes(uint64_t seed, uint32_t count, double lo_log2,
                                        double hi_log2) {
  std::vector<uint32_t> out(count);
  for (uint32_t s = 0; s < count; ++s) {
    const double u = static_cast<double>(bits32(seed, 4, s)) * (1.0 / 4294967296.0);
    out[s] = static_cast<uint32_t>(std::floor(std::exp2(lo_log2 + (hi_log2 - lo_log2) * u)));
  }
  return out;
}

}  // namespace redwood::inputs
