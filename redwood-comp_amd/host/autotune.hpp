// Autotuner for the sort.  The digit passes are compiled in a few tile shapes
// (workgroup size x keys per thread, rwh_sort_variant_shape) and run one of a
// few digit plans (the widths of the four passes over 30-bit keys,
// rwh_sort_plan_widths); which pair is fastest depends on the size.  The tuner
// times every pair on synthetic 30-bit keys of the size bucket
// (2^ceil(log2 n)) with HIP events and keeps the winner per bucket in memory
// and in a small JSON file (RWH_TUNE_CACHE, default
// $HOME/.cache/redwood_hip/tune.json), so later processes skip the
// measurement.  Batches sort on tiles of at most one 8192-point chunk, so
// they are tuned over those shapes only, under their own cache entries.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace redwood::tune {

struct SortConfig {
  int variant = 0;  // tile shape
  int plan = 0;     // digit plan
};

struct SortTiming {
  SortConfig config;
  double ms;
};

// Best configuration for n keys (chunk: a batch's sort, chunk-sized tiles
// only), measured once per bucket and process or read from the cache file.
// The default configuration when no GPU is usable.
SortConfig best_sort_config(uint32_t n, bool chunk = false);
int best_sort_variant(uint32_t n);

// Times every configuration at size n (reps timed sorts each, after one
// warm-up); chunk: chunk-sized tiles only.
std::vector<SortTiming> measure_sort_configs(uint32_t n, int reps, bool chunk = false);

// The cached entry of n's bucket, without measuring (false when none), and
// storing one (written to the cache file at once).
bool lookup(uint32_t n, bool chunk, SortConfig* out);
void store(uint32_t n, bool chunk, SortConfig c);

uint32_t bucket_of(uint32_t n);
std::string cache_path();

}  // namespace redwood::tune
