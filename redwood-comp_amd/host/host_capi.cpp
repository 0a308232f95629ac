// extern "C" access to the host layer for the Python tests and benchmark:
// the seeded input generators, and one whole build through the drop-in path
// (Engine + AppData + run_stage1..7 futures) with every output copied back.
#include <cstring>
#include <string>

#include "autotune.hpp"
#include "build_server.hpp"
#include "hip_engine.hpp"
#include "inputs.hpp"

namespace {
thread_local std::string g_err;
}

extern "C" {

const char* rwt_last_error(void) { return g_err.c_str(); }

uint32_t rwt_bits32(uint64_t seed, uint64_t stream, uint64_t counter) {
  return redwood::inputs::bits32(seed, stream, counter);
}

void rwt_generate_uniform(uint64_t seed, uint64_t n, float* xyzw) {
  redwood::inputs::uniform(seed, n, xyzw);
}

void rwt_generate_gmm(uint64_t seed, uint64_t n, uint32_t n_centers, double sigma, float* xyzw) {
  redwood::inputs::gaussian_mixture(seed, n, n_centers, sigma, xyzw);
}

void rwt_log_uniform_sizes(uint64_t seed, uint32_t count, double lo_log2, double hi_log2,
                           uint32_t* out) {
  const auto v = redwood::inputs::log_uniform_sizes(seed, count, lo_log2, hi_log2);
  std::memcpy(out, v.data(), sizeof(uint32_t) * count);
}

// One build through redwood::hip::run_stage1..7 on a fresh Engine.  Output
// arrays are host memory sized as the AppData buffers (n entries, octree
// arrays oct_cap = floor(0.6 n) entries, children 8 x, corner 4 x);
// counts3 = {n_unique, n_brt, n_octree_nodes}.  fail_stage (1..7, 0 none)
// injects a failure into that stage to exercise the error path.  Returns 0,
// or the failing stage's number with the message in rwt_last_error().
int rwt_dropin_build(const float* xyzw, uint32_t n, int fail_stage, uint32_t* sorted,
                     uint32_t* unique_keys, uint8_t* prefix_n, uint8_t* has_leaf_left,
                     uint8_t* has_leaf_right, int32_t* left_child, int32_t* parents,
                     int32_t* edge_count, int32_t* edge_offset, int32_t* children, float* corner,
                     float* cell_size, int32_t* child_node_mask, int32_t* child_leaf_mask,
                     uint32_t* counts3) {
  using namespace redwood::hip;
  int stage_no = 0;
  try {
    auto engine = Engine::create();
    AppData app(engine->get_mr(), xyzw, n);
    std::future<void> (*stages[7])(AppData&) = {run_stage1, run_stage2, run_stage3, run_stage4,
                                                 run_stage5, run_stage6, run_stage7};
    for (stage_no = 1; stage_no <= 7; ++stage_no) {
      if (fail_stage == stage_no) rwh_debug_fail_next(stage_no);
      stages[stage_no - 1](app).get();
    }
    stage_no = 8;
    const size_t cap = app.oct.u_cell_size.size();
    const auto cp = [](auto& v, auto* dst, size_t count) {
      const auto h = v.to_host(count);
      std::memcpy(dst, h.data(), sizeof(h[0]) * count);
    };
    cp(app.u_morton_keys_alt, sorted, n);
    cp(app.u_morton_keys, unique_keys, n);
    cp(app.brt.u_prefix_n, prefix_n, n);
    cp(app.brt.u_has_leaf_left, has_leaf_left, n);
    cp(app.brt.u_has_leaf_right, has_leaf_right, n);
    cp(app.brt.u_left_child, left_child, n);
    cp(app.brt.u_parents, parents, n);
    cp(app.u_edge_count, edge_count, n);
    cp(app.u_edge_offset, edge_offset, n);
    cp(app.oct.u_children, children, 8 * cap);
    {
      const auto h = app.oct.u_corner.to_host(cap);
      std::memcpy(corner, h.data(), sizeof(float) * 4 * cap);
    }
    cp(app.oct.u_cell_size, cell_size, cap);
    cp(app.oct.u_child_node_mask, child_node_mask, cap);
    cp(app.oct.u_child_leaf_mask, child_leaf_mask, cap);
    counts3[0] = app.get_n_unique();
    counts3[1] = app.get_n_brt_nodes();
    counts3[2] = app.get_n_octree_nodes();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return stage_no == 0 ? -1 : stage_no;
  }
}

// One build through run_stage1..7 on a fresh Engine, read back with
// redwood::hip::read_result: the packed result (sections) goes to out (host,
// cap bytes), its size to *bytes.  Returns 0, or -1 with the message in
// rwt_last_error().
int rwt_dropin_packed(const float* xyzw, uint32_t n, uint32_t sections, void* out, uint64_t cap,
                      uint64_t* bytes) {
  using namespace redwood::hip;
  try {
    auto engine = Engine::create();
    AppData app(engine->get_mr(), xyzw, n);
    for (auto run : {run_stage1, run_stage2, run_stage3, run_stage4, run_stage5, run_stage6, run_stage7})
      run(app).get();
    const auto packed = read_result(app, sections);
    if (packed.bytes().size() > cap) throw std::runtime_error("packed result larger than its buffer");
    std::memcpy(out, packed.bytes().data(), packed.bytes().size());
    *bytes = packed.bytes().size();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// ---- autotuning ----------------------------------------------------------------
// Best sort tile variant for n keys (measured or read from the tune cache).
int rwt_best_sort_variant(uint32_t n) { return redwood::tune::best_sort_variant(n); }

// The tuned (tile shape, digit plan) for n keys; chunk != 0: a batch's sort.
void rwt_best_sort_config(uint32_t n, int chunk, int* variant, int* plan) {
  const auto c = redwood::tune::best_sort_config(n, chunk != 0);
  *variant = c.variant;
  *plan = c.plan;
}

// ms per sort of every configuration at size n (chunk != 0: chunk-sized
// tiles only); returns the number of entries written.
int rwt_measure_sort_configs(uint32_t n, int reps, int chunk, double* ms_out, int* variant_out,
                             int* plan_out, int cap) {
  const auto t = redwood::tune::measure_sort_configs(n, reps, chunk != 0);
  int k = 0;
  for (const auto& x : t)
    if (k < cap) {
      ms_out[k] = x.ms;
      variant_out[k] = x.config.variant;
      plan_out[k] = x.config.plan;
      ++k;
    }
  return k;
}

// The tuning cache without measuring: lookup returns 1 and the cached entry
// of n's size bucket, or 0; store records an entry and writes the file.
int rwt_tune_lookup(uint32_t n, int chunk, int* variant, int* plan) {
  redwood::tune::SortConfig c;
  if (!redwood::tune::lookup(n, chunk != 0, &c)) return 0;
  *variant = c.variant;
  *plan = c.plan;
  return 1;
}

void rwt_tune_store(uint32_t n, int chunk, int variant, int plan) {
  redwood::tune::store(n, chunk != 0, redwood::tune::SortConfig{variant, plan});
}

// ---- serving ------------------------------------------------------------------
// Pure planning (no GPU): how many of the queued sizes the next batch of a
// slot with slot_points index-space points and max_clouds clouds takes.
uint32_t rwt_plan_batch(uint64_t slot_points, uint32_t max_clouds, const uint32_t* sizes,
                        uint32_t n_sizes) {
  redwood::serve::Planner p{slot_points, static_cast<uint64_t>(slot_points * 0.6) + 64ull * max_clouds,
                            max_clouds};
  return static_cast<uint32_t>(p.plan(std::vector<uint32_t>(sizes, sizes + n_sizes)));
}

void* rwt_server_create(uint64_t pool_bytes, uint32_t n_slots, uint32_t slot_points,
                        uint32_t max_clouds, int plain) {
  try {
    redwood::serve::ServerConfig c;
    c.pool_bytes = pool_bytes;
    c.n_slots = n_slots;
    c.slot_points = slot_points;
    c.max_clouds = max_clouds;
    c.plain = plain != 0;
    return new redwood::serve::BuildServer(c);
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void rwt_server_destroy(void* server) { delete static_cast<redwood::serve::BuildServer*>(server); }

uint32_t rwt_server_slots(void* server) {
  return static_cast<redwood::serve::BuildServer*>(server)->n_slots();
}

// Open-loop Poisson load over clouds[k] (host points, sizes[k]);
// stats[12] = rate, n, makespan_s, rps, Mpts/s, p50, p99, mean, max (ms),
// mean batch, rejected, batches launched.
int rwt_serve_load(void* server, uint32_t n_clouds, const float* const* clouds,
                   const uint32_t* sizes, double rate, uint32_t n_requests, uint64_t seed,
                   double* stats) {
  try {
    auto& s = *static_cast<redwood::serve::BuildServer*>(server);
    const uint64_t b0 = s.batches_launched();
    const auto st = redwood::serve::run_open_loop(
        s, std::vector<const float*>(clouds, clouds + n_clouds),
        std::vector<uint32_t>(sizes, sizes + n_clouds), rate, n_requests, seed);
    const double v[12] = {st.rate, static_cast<double>(st.n_requests), st.makespan_s,
                          st.throughput_rps, st.mpts_per_s, st.p50_ms, st.p99_ms, st.mean_ms,
                          st.max_ms, st.mean_batch, static_cast<double>(st.rejected),
                          static_cast<double>(s.batches_launched() - b0)};
    for (int i = 0; i < 12; ++i) stats[i] = v[i];
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// Submits every cloud at once, each asking for the given packed result
// sections, waits, and copies request r's packed result into outs[r] (host,
// out_caps[r] bytes; out_bytes[r] gets its size).  batch_sizes[r] = clouds
// in the batch request r ran in.  Returns 0, or 1 + r for the first failed
// request (the message in rwt_last_error()).
int rwt_serve_mix(void* server, uint32_t n_clouds, const float* const* clouds, const uint32_t* sizes,
                  uint32_t sections, void* const* outs, const uint64_t* out_caps, uint64_t* out_bytes,
                  uint32_t* batch_sizes) {
  auto& s = *static_cast<redwood::serve::BuildServer*>(server);
  std::vector<std::future<redwood::serve::BuildResult>> f;
  for (uint32_t r = 0; r < n_clouds; ++r) f.push_back(s.submit(clouds[r], sizes[r], sections));
  int failed = 0;
  for (uint32_t r = 0; r < n_clouds; ++r) {
    try {
      const auto res = f[r].get();
      const auto& bytes = res.packed.bytes();
      if (bytes.size() > out_caps[r]) throw std::runtime_error("packed result larger than its buffer");
      if (!bytes.empty()) std::memcpy(outs[r], bytes.data(), bytes.size());
      out_bytes[r] = bytes.size();
      batch_sizes[r] = res.batch_size;
    } catch (const std::exception& e) {
      if (failed == 0) {
        g_err = e.what();
        failed = 1 + static_cast<int>(r);
      }
    }
  }
  return failed;
}

}  // extern "C"
