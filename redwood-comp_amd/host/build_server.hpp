// Serving: a request loop that builds octrees for a stream of point clouds
// with continuous batching.  The reference has no serving path; it builds
// one cloud per AppData (apps/tree/main.cpp:18-38), which is what the
// server's "plain" mode reproduces per request.
//
// Design:
//   * Resident slots.  The device memory pool (capped, 64 GB by default) is
//     split into slots.  Each slot owns an index space of points with every
//     AppData buffer, a workspace and a stream, allocated once.
//   * Admission by footprint.  A cloud of n points occupies ceil(n/8192)
//     chunks of 8192 points of a slot's index space (and 0.6 n octree
//     slots).  A request that cannot fit an empty slot is rejected.
//   * Coalescing.  When a slot is free the scheduler takes queued requests in
//     arrival order while they fit (up to max_clouds) and launches them as one
//     batched build (rwh_build_octree_batch).
//   * Continuous batching.  Slots run concurrently on their own streams, so
//     requests that arrive while a batch is on the GPU are admitted into the
//     next free slot and overlap with it on the device.
//   * Completion.  Each slot has a completion thread; results (counts, and
//     on request the build's arrays in the packed result format, packed on
//     the device right after the batch and copied in one transfer per
//     request) are delivered through futures.
#pragma once

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <future>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "result.hpp"
#include "rwhip.h"

namespace redwood::serve {

constexpr uint32_t kChunk = 8192;

// Footprint of one cloud inside a slot.
inline uint64_t padded_points(uint32_t n) {
  const uint64_t c = (static_cast<uint64_t>(n) + kChunk - 1) / kChunk;
  return (c == 0 ? 1 : c) * kChunk;
}
inline uint64_t oct_slots(uint32_t n) {
  const uint64_t cap = static_cast<uint64_t>(static_cast<float>(n) * 0.6f);
  return (cap + 63) / 64 * 64;
}
// Device bytes per index-space point of a slot: points 16, keys 8, radix tree
// 11, edges 8, octree 0.6 x 60, workspace ~5.
constexpr double kBytesPerPoint = 84.0;

// Pure host planning (testable without a GPU): which queued requests form
// the next batch of a slot with the given capacity.
struct PlanItem {
  uint64_t id;
  uint32_t n;
};
struct Planner {
  uint64_t slot_points;
  uint64_t slot_oct;
  uint32_t max_clouds;
  bool fits_empty(uint32_t n) const {
    return padded_points(n) <= slot_points && oct_slots(n) <= slot_oct;
  }
  // Count of requests taken from the front of queue (arrival order).
  size_t plan(const std::vector<uint32_t>& queued_sizes) const;
};

struct ServerConfig {
  uint64_t pool_bytes = 64ull << 30;
  uint32_t n_slots = 3;
  uint32_t slot_points = 1u << 26;  // index-space points per slot
  uint32_t max_clouds = 512;
  bool plain = false;               // one rwh_build_octree per request, one slot
};

struct BuildResult {
  uint32_t n_unique = 0;
  uint32_t n_octree_nodes = 0;
  double t_arrival = 0;  // seconds, server clock
  double t_done = 0;
  uint32_t batch_size = 0;
  // the selected arrays of the build in the packed result format
  // (include/rwhip.h), when the request asked for sections; empty otherwise
  redwood::result::PackedResult packed;
};

class BuildServer {
 public:
  explicit BuildServer(const ServerConfig& cfg);
  ~BuildServer();
  BuildServer(const BuildServer&) = delete;
  BuildServer& operator=(const BuildServer&) = delete;

  // Queues a build of n points (x,y,z,w floats; the memory must stay valid
  // until the future is ready; pinned memory uploads fastest).  sections:
  // the arrays the result carries (RWH_SEC_* bits, e.g. RWH_SECTIONS_ALL;
  // 0: counts only), packed on the device and copied to the host in one
  // transfer.  t_arrival: arrival time on now()'s clock (< 0: now).
  std::future<BuildResult> submit(const float* xyzw, uint32_t n, uint32_t sections = 0,
                                  double t_arrival = -1.0);
  double now() const;
  void drain();  // waits until every submitted request is finished
  uint32_t n_slots() const { return static_cast<uint32_t>(slots_.size()); }
  uint64_t batches_launched() const { return batches_.load(); }
  uint64_t requests_launched() const { return launched_.load(); }
  uint64_t slot_points() const { return planner_.slot_points; }

 private:
  struct Request;
  struct Slot;
  void scheduler_loop();
  void completion_loop(Slot* s);
  void launch(Slot* s, std::vector<std::unique_ptr<Request>> batch);
  int pack_results(Slot* s, const std::vector<std::unique_ptr<Request>>& batch);

  ServerConfig cfg_;
  Planner planner_;
  std::vector<std::unique_ptr<Slot>> slots_;
  std::mutex mu_;
  std::condition_variable cv_, idle_cv_;
  std::deque<std::unique_ptr<Request>> queue_;
  uint64_t next_id_ = 0;
  uint64_t in_flight_ = 0;
  bool stop_ = false;
  std::thread scheduler_;
  std::atomic<uint64_t> batches_{0}, launched_{0};
  std::chrono::steady_clock::time_point t0_;
};

// Open-loop Poisson load: n_requests arrivals at `rate` per second; request r
// builds cloud r % clouds.size().  Latency = done - scheduled arrival.
struct LoadStats {
  double rate = 0;
  uint32_t n_requests = 0;
  double makespan_s = 0;
  double throughput_rps = 0;
  double mpts_per_s = 0;
  double p50_ms = 0, p99_ms = 0, mean_ms = 0, max_ms = 0;
  double mean_batch = 0;
  uint32_t rejected = 0;
};
LoadStats run_open_loop(BuildServer& server, const std::vector<const float*>& clouds,
                        const std::vector<uint32_t>& sizes, double rate, uint32_t n_requests,
                        uint64_t seed);

}  // namespace redwood::serve
