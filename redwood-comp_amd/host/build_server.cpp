#include "build_server.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

#include "autotune.hpp"
#include "hip_engine.hpp"
#include "inputs.hpp"

namespace redwood::serve {

using redwood::hip::check;
using redwood::hip::Error;

size_t Planner::plan(const std::vector<uint32_t>& sizes) const {
  uint64_t pts = 0, oct = 0;
  size_t k = 0;
  for (; k < sizes.size() && k < max_clouds; ++k) {
    const uint64_t p = padded_points(sizes[k]), o = oct_slots(sizes[k]);
    if (pts + p > slot_points || oct + o > slot_oct) break;
    pts += p;
    oct += o;
  }
  return k;
}

struct BuildServer::Request {
  uint64_t id;
  const float* pts;
  uint32_t n;
  uint32_t sections;  // packed result sections (0: counts only)
  double t_arrival;
  std::promise<BuildResult> done;
};

struct BuildServer::Slot {
  uint32_t id = 0;
  rwh_stream_t stream = nullptr;
  rwh_workspace* ws = nullptr;
  rwh_event_t done_ev = nullptr;
  std::vector<void*> allocs;
  rwh_tree_buffers bufs{};
  uint32_t* h_counts = nullptr;  // pinned: n_unique[max_clouds], n_oct[max_clouds], status[max_clouds]
  uint32_t max_clouds = 0;
  // the batch in flight
  std::vector<std::unique_ptr<Request>> batch;
  std::vector<uint32_t> n, pb, ob, oc;
  // packed results of the batch in flight: device arena (grown on demand)
  // and each request's offset in it
  void* arena = nullptr;
  uint64_t arena_bytes = 0;
  std::vector<uint64_t> arena_off;
  bool busy = false;
  bool has_work = false;
  bool stop = false;
  std::mutex m;
  std::condition_variable cv;
  std::thread completion;
};

namespace {
void* dev_alloc(std::vector<void*>& allocs, size_t bytes) {
  void* p = nullptr;
  check(rwh_malloc(&p, bytes < 16 ? 16 : bytes), "rwh_malloc");
  allocs.push_back(p);
  return p;
}
}  // namespace

BuildServer::BuildServer(const ServerConfig& cfg) : cfg_(cfg), t0_(std::chrono::steady_clock::now()) {
  if (cfg_.plain) {
    cfg_.n_slots = 1;
    cfg_.max_clouds = 1;
  }
  if (cfg_.n_slots == 0 || cfg_.max_clouds == 0 || cfg_.slot_points < kChunk)
    throw std::invalid_argument("BuildServer: bad configuration");
  cfg_.slot_points = cfg_.slot_points / kChunk * kChunk;
  // the pool cap bounds the number of resident slots
  const double slot_bytes = kBytesPerPoint * cfg_.slot_points;
  const uint32_t fit = static_cast<uint32_t>(static_cast<double>(cfg_.pool_bytes) / slot_bytes);
  if (fit == 0) throw std::invalid_argument("BuildServer: one slot does not fit the pool cap");
  cfg_.n_slots = std::min(cfg_.n_slots, fit);
  planner_ = Planner{cfg_.slot_points, static_cast<uint64_t>(cfg_.slot_points * 0.6) + 64ull * cfg_.max_clouds,
                     cfg_.max_clouds};

  int ndev = 0;
  check(rwh_device_count(&ndev), "rwh_device_count");
  if (ndev == 0) throw Error(RWH_ERR_NO_DEVICE, "BuildServer: no GPU device");
  const uint64_t P = cfg_.slot_points, O = planner_.slot_oct;
  // one tile shape for every slot, tuned at the size of a typical batch
  const auto tuned = redwood::tune::best_sort_config(std::min<uint32_t>(cfg_.slot_points, 1u << 22), !cfg_.plain);
  for (uint32_t i = 0; i < cfg_.n_slots; ++i) {
    auto s = std::make_unique<Slot>();
    s->id = i;
    s->max_clouds = cfg_.max_clouds;
    check(rwh_stream_create(&s->stream), "rwh_stream_create");
    check(rwh_workspace_create(static_cast<uint32_t>(P), cfg_.max_clouds, &s->ws), "rwh_workspace_create");
    check(rwh_workspace_set_sort_variant(s->ws, tuned.variant), "rwh_workspace_set_sort_variant");
    check(rwh_workspace_set_sort_plan(s->ws, tuned.plan), "rwh_workspace_set_sort_plan");
    check(rwh_event_create(&s->done_ev), "rwh_event_create");
    void* h = nullptr;
    check(rwh_host_alloc(&h, 3 * sizeof(uint32_t) * cfg_.max_clouds), "rwh_host_alloc");
    s->h_counts = static_cast<uint32_t*>(h);
    auto& a = s->allocs;
    rwh_tree_buffers& b = s->bufs;
    b.points = static_cast<float*>(dev_alloc(a, 16 * P));
    b.morton_keys = static_cast<uint32_t*>(dev_alloc(a, 4 * P));
    b.morton_keys_alt = static_cast<uint32_t*>(dev_alloc(a, 4 * P));
    b.prefix_n = static_cast<uint8_t*>(dev_alloc(a, P));
    b.has_leaf_left = static_cast<uint8_t*>(dev_alloc(a, P));
    b.has_leaf_right = static_cast<uint8_t*>(dev_alloc(a, P));
    b.left_child = static_cast<int32_t*>(dev_alloc(a, 4 * P));
    b.parents = static_cast<int32_t*>(dev_alloc(a, 4 * P));
    b.edge_count = static_cast<int32_t*>(dev_alloc(a, 4 * P));
    b.edge_offset = static_cast<int32_t*>(dev_alloc(a, 4 * P));
    b.children = static_cast<int32_t*>(dev_alloc(a, 32 * O));
    b.corner = static_cast<float*>(dev_alloc(a, 16 * O));
    b.cell_size = static_cast<float*>(dev_alloc(a, 4 * O));
    b.child_node_mask = static_cast<int32_t*>(dev_alloc(a, 4 * O));
    b.child_leaf_mask = static_cast<int32_t*>(dev_alloc(a, 4 * O));
    b.oct_capacity = static_cast<uint32_t>(O);
    Slot* sp = s.get();
    s->completion = std::thread([this, sp] { completion_loop(sp); });
    slots_.push_back(std::move(s));
  }
  scheduler_ = std::thread([this] { scheduler_loop(); });
}

BuildServer::~BuildServer() {
  drain();
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  if (scheduler_.joinable()) scheduler_.join();
  for (auto& s : slots_) {
    {
      std::lock_guard<std::mutex> g(s->m);
      s->stop = true;
    }
    s->cv.notify_all();
    if (s->completion.joinable()) s->completion.join();
    (void)rwh_stream_sync(s->stream);
    for (void* p : s->allocs) (void)rwh_free(p);
    if (s->arena) (void)rwh_free(s->arena);
    (void)rwh_workspace_destroy(s->ws);
    (void)rwh_event_destroy(s->done_ev);
    (void)rwh_host_free(s->h_counts);
    (void)rwh_stream_destroy(s->stream);
  }
}

double BuildServer::now() const {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

std::future<BuildResult> BuildServer::submit(const float* xyzw, uint32_t n, uint32_t sections,
                                             double t_arrival) {
  auto r = std::make_unique<Request>();
  r->pts = xyzw;
  r->n = n;
  r->sections = sections;
  r->t_arrival = t_arrival < 0 ? now() : t_arrival;
  std::future<BuildResult> f = r->done.get_future();
  if ((sections & ~RWH_SECTIONS_ALL) != 0) {
    r->done.set_exception(std::make_exception_ptr(
        Error(RWH_ERR_INVALID_ARG, "BuildServer: sections outside RWH_SECTIONS_ALL")));
    return f;
  }
  if (!planner_.fits_empty(n) || (n > 0 && !xyzw)) {
    r->done.set_exception(std::make_exception_ptr(
        Error(RWH_ERR_CAPACITY, "BuildServer: a cloud of " + std::to_string(n) +
                                    " points does not fit one slot of " +
                                    std::to_string(planner_.slot_points) + " points")));
    return f;
  }
  {
    std::lock_guard<std::mutex> g(mu_);
    r->id = next_id_++;
    ++in_flight_;
    queue_.push_back(std::move(r));
  }
  cv_.notify_all();
  return f;
}

void BuildServer::drain() {
  std::unique_lock<std::mutex> lk(mu_);
  idle_cv_.wait(lk, [&] { return in_flight_ == 0; });
}

void BuildServer::scheduler_loop() {
  std::unique_lock<std::mutex> lk(mu_);
  for (;;) {
    Slot* free_slot = nullptr;
    cv_.wait(lk, [&] {
      if (stop_) return true;
      if (queue_.empty()) return false;
      for (auto& s : slots_)
        if (!s->busy) {
          free_slot = s.get();
          return true;
        }
      return false;
    });
    if (stop_) return;
    std::vector<uint32_t> sizes;
    const size_t look = std::min<size_t>(queue_.size(), cfg_.max_clouds);
    sizes.reserve(look);
    for (size_t i = 0; i < look; ++i) sizes.push_back(queue_[i]->n);
    size_t k = planner_.plan(sizes);
    if (k == 0) k = 1;  // admitted at submit, so it fits an empty slot
    std::vector<std::unique_ptr<Request>> batch;
    for (size_t i = 0; i < k; ++i) {
      batch.push_back(std::move(queue_.front()));
      queue_.pop_front();
    }
    free_slot->busy = true;
    lk.unlock();
    launch(free_slot, std::move(batch));
    lk.lock();
  }
}

// Enqueues the packing of every request of the batch that asked for
// sections, into the slot's arena (grown when too small: the slot's stream is
// idle here, its last batch finished before the slot was freed).
int BuildServer::pack_results(Slot* s, const std::vector<std::unique_ptr<Request>>& batch) {
  const uint32_t k = static_cast<uint32_t>(batch.size());
  s->arena_off.assign(k, 0);
  uint64_t need = 0;
  for (uint32_t i = 0; i < k; ++i) {
    if (batch[i]->sections == 0) continue;
    s->arena_off[i] = need;
    need += (rwh_result_bytes(s->n[i], s->n[i], s->oc[i], s->oc[i], batch[i]->sections) + 255) / 256 * 256;
  }
  if (need == 0) return RWH_OK;
  if (need > s->arena_bytes) {
    if (s->arena) (void)rwh_free(s->arena);
    s->arena = nullptr;
    s->arena_bytes = 0;
    const int st = rwh_malloc(&s->arena, need);
    if (st != RWH_OK) return st;
    s->arena_bytes = need;
  }
  const uint32_t* d_counts = rwh_workspace_counts(s->ws);
  for (uint32_t i = 0; i < k; ++i) {
    const uint32_t sections = batch[i]->sections;
    if (sections == 0) continue;
    rwh_tree_buffers b = s->bufs;  // the cloud's arrays
    const size_t p = s->pb[i], o = s->ob[i];
    b.points += 4 * p;
    b.morton_keys += p;
    b.morton_keys_alt += p;
    b.prefix_n += p;
    b.has_leaf_left += p;
    b.has_leaf_right += p;
    b.left_child += p;
    b.parents += p;
    b.edge_count += p;
    b.edge_offset += p;
    b.children += 8 * o;
    b.corner += 4 * o;
    b.cell_size += o;
    b.child_node_mask += o;
    b.child_leaf_mask += o;
    b.oct_capacity = s->oc[i];
    const int st = rwh_pack_result(&b, d_counts + i, s->max_clouds, s->n[i], sections,
                                   static_cast<char*>(s->arena) + s->arena_off[i],
                                   s->arena_bytes - s->arena_off[i], s->stream);
    if (st != RWH_OK) return st;
  }
  return RWH_OK;
}

void BuildServer::launch(Slot* s, std::vector<std::unique_ptr<Request>> batch) {
  const uint32_t k = static_cast<uint32_t>(batch.size());
  s->n.assign(k, 0);
  s->pb.assign(k, 0);
  s->ob.assign(k, 0);
  s->oc.assign(k, 0);
  for (uint32_t i = 0; i < k; ++i) s->n[i] = batch[i]->n;
  int st = RWH_OK;
  uint32_t space = 0, ospace = 0;
  st = rwh_batch_layout(s->n.data(), k, s->pb.data(), s->ob.data(), s->oc.data(), &space, &ospace);
  for (uint32_t i = 0; st == RWH_OK && i < k; ++i)
    st = rwh_memcpy_h2d(const_cast<float*>(s->bufs.points) + 4ull * s->pb[i], batch[i]->pts,
                        16ull * batch[i]->n, s->stream);
  uint32_t* d_counts = rwh_workspace_counts(s->ws);
  if (st == RWH_OK) {
    if (cfg_.plain)
      st = rwh_build_octree(s->ws, &s->bufs, s->n[0], 0.0f, 1024.0f, s->stream);
    else
      st = rwh_build_octree_batch(s->ws, &s->bufs, s->n.data(), s->pb.data(), s->ob.data(),
                                  s->oc.data(), k, 0.0f, 1024.0f, s->stream);
  }
  if (st == RWH_OK)
    st = rwh_memcpy_d2h(s->h_counts, d_counts, sizeof(uint32_t) * k, s->stream);
  if (st == RWH_OK)
    st = rwh_memcpy_d2h(s->h_counts + s->max_clouds, d_counts + s->max_clouds,
                        sizeof(uint32_t) * k, s->stream);
  if (st == RWH_OK)
    st = rwh_memcpy_d2h(s->h_counts + 2 * s->max_clouds, d_counts + 2 * s->max_clouds,
                        sizeof(uint32_t) * k, s->stream);
  if (st == RWH_OK) st = pack_results(s, batch);
  if (st == RWH_OK) st = rwh_event_record(s->done_ev, s->stream);
  if (st != RWH_OK) {
    const std::string msg = std::string("BuildServer launch: ") + rwh_error_string(st) + " (" +
                            rwh_last_error() + ")";
    for (auto& r : batch) r->done.set_exception(std::make_exception_ptr(Error(st, msg)));
    std::lock_guard<std::mutex> g(mu_);
    s->busy = false;
    in_flight_ -= k;
    cv_.notify_all();
    idle_cv_.notify_all();
    return;
  }
  batches_.fetch_add(1);
  launched_.fetch_add(k);
  {
    std::lock_guard<std::mutex> g(s->m);
    s->batch = std::move(batch);
    s->has_work = true;
  }
  s->cv.notify_all();
}

void BuildServer::completion_loop(Slot* s) {
  for (;;) {
    std::vector<std::unique_ptr<Request>> batch;
    {
      std::unique_lock<std::mutex> lk(s->m);
      s->cv.wait(lk, [&] { return s->stop || s->has_work; });
      if (!s->has_work) return;
      batch = std::move(s->batch);
      s->has_work = false;
    }
    const int st = rwh_event_sync(s->done_ev);
    const double t_done = now();
    const uint32_t k = static_cast<uint32_t>(batch.size());
    for (uint32_t i = 0; i < k; ++i) {
      Request& r = *batch[i];
      try {
        check(st, "rwh_event_sync");
        BuildResult res;
        res.n_unique = s->h_counts[i];
        res.n_octree_nodes = s->h_counts[s->max_clouds + i];
        res.t_arrival = r.t_arrival;
        res.t_done = t_done;
        res.batch_size = k;
        if (s->h_counts[2 * s->max_clouds + i] != 0)
          throw Error(RWH_ERR_OVERFLOW, "BuildServer: a kernel bound was exceeded (status " +
                                            std::to_string(s->h_counts[2 * s->max_clouds + i]) + ")");
        if (res.n_octree_nodes > s->oc[i])
          throw Error(RWH_ERR_CAPACITY, "BuildServer: octree above the cloud's 0.6 n capacity");
        if (r.sections != 0) {
          const uint64_t bytes =
              rwh_result_bytes(r.n, res.n_unique, res.n_octree_nodes, s->oc[i], r.sections);
          std::vector<uint8_t> blob(bytes);
          check(rwh_memcpy_d2h(blob.data(), static_cast<char*>(s->arena) + s->arena_off[i], bytes, s->stream),
                "rwh_memcpy_d2h");
          check(rwh_stream_sync(s->stream), "rwh_stream_sync");
          res.packed = redwood::result::PackedResult(std::move(blob));
        }
        r.done.set_value(std::move(res));
      } catch (...) {
        r.done.set_exception(std::current_exception());
      }
    }
    std::lock_guard<std::mutex> g(mu_);
    s->busy = false;
    in_flight_ -= k;
    cv_.notify_all();
    idle_cv_.notify_all();
  }
}

LoadStats run_open_loop(BuildServer& server, const std::vector<const float*>& clouds,
                        const std::vector<uint32_t>& sizes, double rate, uint32_t n_requests,
                        uint64_t seed) {
  LoadStats st;
  st.rate = rate;
  st.n_requests = n_requests;
  std::vector<double> arrival(n_requests);
  std::vector<uint32_t> which(n_requests);
  double t = 0;
  for (uint32_t r = 0; r < n_requests; ++r) {
    const double u = (static_cast<double>(redwood::inputs::bits32(seed, 5, r)) + 0.5) / 4294967296.0;
    t += -std::log(u) / rate;
    arrival[r] = t;
    which[r] = redwood::inputs::bits32(seed, 6, r) % static_cast<uint32_t>(clouds.size());
  }
  std::vector<std::future<BuildResult>> fut;
  fut.reserve(n_requests);
  const double t_start = server.now() + 0.01;
  for (uint32_t r = 0; r < n_requests; ++r) {
    const double ta = t_start + arrival[r];
    const double wait = ta - server.now();
    if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
    fut.push_back(server.submit(clouds[which[r]], sizes[which[r]], 0, ta));
  }
  std::vector<double> lat;
  lat.reserve(n_requests);
  double last_done = t_start, total_pts = 0, batch_sum = 0;
  for (uint32_t r = 0; r < n_requests; ++r) {
    try {
      const BuildResult res = fut[r].get();
      lat.push_back((res.t_done - res.t_arrival) * 1e3);
      last_done = std::max(last_done, res.t_done);
      total_pts += sizes[which[r]];
      batch_sum += res.batch_size;
    } catch (...) {
      ++st.rejected;
    }
  }
  st.makespan_s = last_done - (t_start + arrival[0]);
  if (!lat.empty()) {
    std::vector<double> s = lat;
    std::sort(s.begin(), s.end());
    const auto pct = [&](double p) {
      const size_t i = static_cast<size_t>(std::ceil(p * s.size())) - 1;
      return s[std::min(i, s.size() - 1)];
    };
    st.p50_ms = pct(0.50);
    st.p99_ms = pct(0.99);
    st.max_ms = s.back();
    double sum = 0;
    for (double v : s) sum += v;
    st.mean_ms = sum / s.size();
    st.mean_batch = batch_sum / lat.size();
  }
  st.throughput_rps = lat.size() / st.makespan_s;
  st.mpts_per_s = total_pts / st.makespan_s / 1e6;
  return st;
}

}  // namespace redwood::serve
