// Seeded synthetic point clouds for the octree build.
//
// The reference fills its input with std::mt19937(114514) and
// std::uniform_real_distribution (apps/tree/app_data.hpp:32-39), whose output
// is implementation-defined and whose argument evaluation order is
// unspecified.  These generators are counter-based instead: coordinate a of
// point i is a pure function of (seed, stream, counter), so every consumer
// (CPU oracle, GPU path, C++ and Python) sees identical bits.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace redwood::inputs {

// 32 random bits for (seed, stream, counter): the splitmix64 finaliser over a
// Weyl-style combination of the three.
uint32_t bits32(uint64_t seed, uint64_t stream, uint64_t counter);

// A float in [0, 1024): the top 24 bits of bits32 scaled by 2^-14 (exact).
float coord(uint64_t seed, uint64_t stream, uint64_t counter);

// n points uniform in [0,1024)^3, w = 0 (xyzw, 4 floats per point).
void uniform(uint64_t seed, size_t n, float* xyzw);

// n points from an isotropic Gaussian mixture: n_centers centres uniform in
// [0,1024)^3, standard deviation sigma, samples clamped to
// [0, nextafterf(1024, 0)]; w = 0.
void gaussian_mixture(uint64_t seed, size_t n, uint32_t n_centers, double sigma, float* xyzw);

// Cloud sizes of a batch: floor(2^(lo_log2 + (hi_log2 - lo_log2) * u)), u uniform in [0,1).
std::vector<uint32_t> log_uniform_sizes(uint64_t seed, uint32_t count, double lo_log2,
                                        double hi_log2);

}  // namespace redwood::inputs
