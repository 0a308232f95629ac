// Autotuner for the sort's tile shape.  The digit passes are compiled in a
// few shapes (workgroup size x keys per thread, see rwh_sort_variant_shape);
// which one is fastest depends on the size.  The tuner times every variant on
// synthetic 30-bit keys of the size bucket (2^ceil(log2 n)), with HIP events,
// and keeps the winner per bucket in memory and in a small JSON file
// (RWH_TUNE_CACHE, default $HOME/.cache/redwood_hip/tune.json) so later
// processes skip the measurement.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace redwood::tune {

struct SortTiming {
  int variant;
  double ms;
};

// Best variant for n keys (measured once per bucket and process, or read
// from the cache file).  Returns 0 when no GPU is usable.
int best_sort_variant(uint32_t n);

// Measures every variant at size n (reps timed sorts each, after one warm-up).
std::vector<SortTiming> measure_sort_variants(uint32_t n, int reps);

uint32_t bucket_of(uint32_t n);
std::string cache_path();

}  // namespace redwood::tune
