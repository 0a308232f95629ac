#include "hip_engine.hpp"

#include <new>
#include <vector>

#include "autotune.hpp"
#include "inputs.hpp"

namespace redwood::tree {

AppData::AppData(std::pmr::memory_resource* mr, size_t n, uint64_t seed)
    : BaseAppData(mr),
      n_input(static_cast<uint32_t>(n)),
      u_input_points(n, mr),
      u_morton_keys(n, mr),
      u_morton_keys_alt(n, mr),
      u_edge_count(n, mr),
      u_edge_offset(n, mr),
      brt(n, mr),
      oct(octree_capacity(n), mr) {
  std::vector<float> pts(4 * n);
  redwood::inputs::uniform(seed, n, pts.data());
  u_input_points.from_host(reinterpret_cast<const Vec4*>(pts.data()), n);
}

AppData::AppData(std::pmr::memory_resource* mr, const float* xyzw, size_t n)
    : BaseAppData(mr),
      n_input(static_cast<uint32_t>(n)),
      u_input_points(n, mr),
      u_morton_keys(n, mr),
      u_morton_keys_alt(n, mr),
      u_edge_count(n, mr),
      u_edge_offset(n, mr),
      brt(n, mr),
      oct(octree_capacity(n), mr) {
  u_input_points.from_host(reinterpret_cast<const Vec4*>(xyzw), n);
}

}  // namespace redwood::tree

namespace redwood::hip {

void check(int status, const char* what) {
  if (status != RWH_OK)
    throw Error(status, std::string(what) + ": " + rwh_error_string(status) + " (" +
                            rwh_last_error() + ")");
}

// ---- memory resource ------------------------------------------------------------
void* HipMemoryResource::do_allocate(size_t bytes, size_t alignment) {
  if (alignment > 256) throw std::bad_alloc();
  void* p = nullptr;
  if (rwh_malloc(&p, bytes == 0 ? 1 : bytes) != RWH_OK) throw std::bad_alloc();
  return p;
}
void HipMemoryResource::do_deallocate(void* p, size_t, size_t) {
  // a buffer may still be in use by enqueued work
  engine_->sync();
  (void)rwh_free(p);
}
void HipMemoryResource::zero(void* p, size_t bytes) {
  check(rwh_memset(p, 0, bytes, engine_->stream()), "rwh_memset");
}
void HipMemoryResource::upload(void* dst, const void* src, size_t bytes) {
  check(rwh_memcpy_h2d(dst, src, bytes, engine_->stream()), "rwh_memcpy_h2d");
  engine_->sync();  // src is pageable caller memory
}
void HipMemoryResource::download(void* dst, const void* src, size_t bytes) {
  check(rwh_memcpy_d2h(dst, src, bytes, engine_->stream()), "rwh_memcpy_d2h");
  engine_->sync();
}

// ---- engine ---------------------------------------------------------------------
Engine::Engine() : mr_(this) {}

std::unique_ptr<Engine> Engine::create() {
  int n = 0;
  check(rwh_device_count(&n), "rwh_device_count");
  if (n == 0) throw Error(RWH_ERR_NO_DEVICE, "redwood::hip::Engine: no GPU device");
  std::unique_ptr<Engine> e(new Engine());
  check(rwh_stream_create(&e->stream_), "rwh_stream_create");
  e->worker_ = std::thread([p = e.get()] { p->completion_loop(); });
  return e;
}

Engine::~Engine() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  if (worker_.joinable()) worker_.join();
  if (stream_) {
    (void)rwh_stream_sync(stream_);
    (void)rwh_stream_destroy(stream_);
  }
}

void Engine::sync() { check(rwh_stream_sync(stream_), "rwh_stream_sync"); }

std::string Engine::device_name() const {
  char buf[256];
  check(rwh_device_name(buf, sizeof buf), "rwh_device_name");
  return buf;
}

std::future<void> Engine::when_done(std::function<void()> then) {
  rwh_event_t ev = nullptr;
  check(rwh_event_create(&ev), "rwh_event_create");
  const int st = rwh_event_record(ev, stream_);
  if (st != RWH_OK) {
    (void)rwh_event_destroy(ev);
    check(st, "rwh_event_record");
  }
  Pending p{ev, std::move(then), {}};
  std::future<void> f = p.done.get_future();
  {
    std::lock_guard<std::mutex> g(mu_);
    queue_.push_back(std::move(p));
  }
  cv_.notify_one();
  return f;
}

void Engine::completion_loop() {
  for (;;) {
    Pending p;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
      if (queue_.empty()) return;
      p = std::move(queue_.front());
      queue_.pop_front();
    }
    try {
      check(rwh_event_sync(p.event), "rwh_event_sync");
      if (p.then) p.then();
      p.done.set_value();
    } catch (...) {
      p.done.set_exception(std::current_exception());
    }
    (void)rwh_event_destroy(p.event);
  }
}

// ---- per-build device state ---------------------------------------------------------
namespace {

struct DeviceState {
  rwh_workspace* ws = nullptr;
  uint32_t* h_counts = nullptr;  // pinned: n_unique, n_octree_nodes, status word
  ~DeviceState() {
    if (ws) (void)rwh_workspace_destroy(ws);
    if (h_counts) (void)rwh_host_free(h_counts);
  }
};

Engine& engine_of(AppData& a) {
  auto* mr = dynamic_cast<HipMemoryResource*>(a.mr_);
  if (!mr)
    throw std::invalid_argument("redwood::hip dispatchers need an AppData built on an Engine's get_mr()");
  return mr->engine();
}

DeviceState& state_of(AppData& a) {
  if (!a.backend_state) {
    auto s = std::make_shared<DeviceState>();
    check(rwh_workspace_create(a.n_input, 1, &s->ws), "rwh_workspace_create");
    const auto tuned = redwood::tune::best_sort_config(a.n_input);
    check(rwh_workspace_set_sort_variant(s->ws, tuned.variant), "rwh_workspace_set_sort_variant");
    check(rwh_workspace_set_sort_plan(s->ws, tuned.plan), "rwh_workspace_set_sort_plan");
    void* h = nullptr;
    check(rwh_host_alloc(&h, 3 * sizeof(uint32_t)), "rwh_host_alloc");
    s->h_counts = static_cast<uint32_t*>(h);
    a.backend_state = s;
  }
  return *static_cast<DeviceState*>(a.backend_state.get());
}

rwh_tree_buffers buffers_of(AppData& a) {
  rwh_tree_buffers b{};
  b.points = reinterpret_cast<const float*>(a.u_input_points.data());
  b.morton_keys = a.u_morton_keys.data();
  b.morton_keys_alt = a.u_morton_keys_alt.data();
  b.prefix_n = a.brt.u_prefix_n.data();
  b.has_leaf_left = a.brt.u_has_leaf_left.data();
  b.has_leaf_right = a.brt.u_has_leaf_right.data();
  b.left_child = a.brt.u_left_child.data();
  b.parents = a.brt.u_parents.data();
  b.edge_count = a.u_edge_count.data();
  b.edge_offset = a.u_edge_offset.data();
  b.children = a.oct.u_children.data();
  b.corner = reinterpret_cast<float*>(a.oct.u_corner.data());
  b.cell_size = a.oct.u_cell_size.data();
  b.child_node_mask = a.oct.u_child_node_mask.data();
  b.child_leaf_mask = a.oct.u_child_leaf_mask.data();
  b.oct_capacity = static_cast<uint32_t>(a.oct.u_cell_size.size());
  return b;
}

uint32_t* d_n_unique(DeviceState& s) { return rwh_workspace_counts(s.ws); }
uint32_t* d_n_oct(DeviceState& s) { return rwh_workspace_counts(s.ws) + 1; }  // max_segments = 1
uint32_t* d_status(DeviceState& s) { return rwh_workspace_counts(s.ws) + 2; }

// Enqueues the copy of the status word; the stage's completion raises on it.
int fetch_status(Engine& e, DeviceState& s) {
  return rwh_memcpy_d2h(s.h_counts + 2, d_status(s), sizeof(uint32_t), e.stream());
}

void raise_status(const DeviceState& s, const char* stage) {
  if (s.h_counts[2] != 0)
    throw Error(RWH_ERR_OVERFLOW, std::string(stage) + ": a kernel bound was exceeded (status " +
                                      std::to_string(s.h_counts[2]) + ", RWH_STATUS_* bits)");
}

// Runs enqueue(); a failure becomes the future's exception.  then(state)
// runs on the completion thread once the stage's work is done; the AppData
// must outlive the returned future.
template <typename F, typename G>
std::future<void> stage(AppData& a, const char* name, F&& enqueue, G then) {
  try {
    Engine& e = engine_of(a);
    DeviceState& s = state_of(a);
    std::shared_ptr<void> keep = a.backend_state;
    check(enqueue(e, s), name);
    DeviceState* sp = &s;
    return e.when_done([keep, sp, then] { then(*sp); });
  } catch (...) {
    std::promise<void> p;
    p.set_exception(std::current_exception());
    return p.get_future();
  }
}

template <typename F>
std::future<void> stage(AppData& a, const char* name, F&& enqueue) {
  return stage(a, name, std::forward<F>(enqueue), [](DeviceState&) {});
}

}  // namespace

std::future<void> run_stage1(AppData& a) {
  return stage(a, "run_stage1", [&](Engine& e, DeviceState&) {
    return rwh_morton(reinterpret_cast<const float*>(a.u_input_points.data()), a.n_input,
                      AppData::min_coord, AppData::range, a.u_morton_keys.data(), e.stream());
  });
}

std::future<void> run_stage2(AppData& a) {
  return stage(a, "run_stage2", [&](Engine& e, DeviceState& s) {
    return rwh_sort_u32(s.ws, a.u_morton_keys.data(), a.u_morton_keys_alt.data(), a.n_input, 30,
                        e.stream());
  });
}

std::future<void> run_stage3(AppData& a) {
  return stage(
      a, "run_stage3",
      [&](Engine& e, DeviceState& s) {
        int st = rwh_unique_u32(s.ws, a.get_sorted_morton_keys(), a.n_input,
                                a.get_unique_morton_keys(), d_n_unique(s), e.stream());
        if (st == RWH_OK) st = rwh_memcpy_d2h(s.h_counts, d_n_unique(s), sizeof(uint32_t), e.stream());
        return st;
      },
      [&a](DeviceState& s) {
        const uint32_t u = s.h_counts[0];
        a.set_n_unique(u);
        a.set_n_brt_nodes(u > 0 ? u - 1 : 0);
      });
}

std::future<void> run_stage4(AppData& a) {
  return stage(
      a, "run_stage4",
      [&](Engine& e, DeviceState& s) {
        int st = rwh_radix_tree(s.ws, a.get_unique_morton_keys(), d_n_unique(s), a.n_input,
                                a.brt.u_prefix_n.data(), a.brt.u_has_leaf_left.data(),
                                a.brt.u_has_leaf_right.data(), a.brt.u_left_child.data(),
                                a.brt.u_parents.data(), e.stream());
        if (st == RWH_OK) st = fetch_status(e, s);
        return st;
      },
      [](DeviceState& s) { raise_status(s, "run_stage4"); });
}

std::future<void> run_stage5(AppData& a) {
  return stage(a, "run_stage5", [&](Engine& e, DeviceState& s) {
    return rwh_edge_count(a.brt.u_prefix_n.data(), a.brt.u_parents.data(), d_n_unique(s), a.n_input,
                          a.u_edge_count.data(), e.stream());
  });
}

std::future<void> run_stage6(AppData& a) {
  const uint32_t cap = static_cast<uint32_t>(a.oct.u_cell_size.size());
  return stage(
      a, "run_stage6",
      [&](Engine& e, DeviceState& s) {
        int st = rwh_prefix_sum(s.ws, a.u_edge_count.data(), d_n_unique(s), a.n_input,
                                a.u_edge_offset.data(), d_n_oct(s), e.stream());
        if (st == RWH_OK) st = rwh_memcpy_d2h(s.h_counts + 1, d_n_oct(s), sizeof(uint32_t), e.stream());
        return st;
      },
      [&a, cap](DeviceState& s) {
        const uint32_t n_oct = s.h_counts[1];
        a.set_n_octree_nodes(n_oct);
        if (n_oct > cap)
          throw Error(RWH_ERR_CAPACITY, "run_stage6: n_octree_nodes = " + std::to_string(n_oct) +
                                            " exceeds the octree capacity " + std::to_string(cap));
      });
}

std::future<void> run_stage7(AppData& a) {
  return stage(
      a, "run_stage7",
      [&](Engine& e, DeviceState& s) {
        const rwh_tree_buffers b = buffers_of(a);
        int st = rwh_octree(s.ws, &b, d_n_unique(s), d_n_oct(s), a.n_input, AppData::min_coord,
                            AppData::range, d_status(s), e.stream());
        if (st == RWH_OK) st = fetch_status(e, s);
        return st;
      },
      [](DeviceState& s) { raise_status(s, "run_stage7"); });
}

std::future<void> run_link_leaf(AppData& a) {
  return stage(
      a, "run_link_leaf",
      [&](Engine& e, DeviceState& s) {
        const rwh_tree_buffers b = buffers_of(a);
        int st = rwh_link_leaf(&b, d_n_unique(s), d_n_oct(s), a.n_input, d_status(s), e.stream());
        if (st == RWH_OK) st = fetch_status(e, s);
        return st;
      },
      [](DeviceState& s) { raise_status(s, "run_link_leaf"); });
}

redwood::result::PackedResult read_result(AppData& a, uint32_t sections) {
  Engine& e = engine_of(a);
  DeviceState& s = state_of(a);
  const rwh_tree_buffers b = buffers_of(a);
  const uint64_t max_bytes = rwh_result_bytes(a.n_input, a.n_input, b.oct_capacity, b.oct_capacity, sections);
  if (max_bytes == 0) throw Error(RWH_ERR_INVALID_ARG, "read_result: sections outside RWH_SECTIONS_ALL");
  void* scratch = nullptr;
  check(rwh_malloc(&scratch, max_bytes), "rwh_malloc");
  try {
    auto out = redwood::result::fetch(b, d_n_unique(s), 1, a.n_input, sections, scratch, max_bytes, e.stream());
    check(rwh_free(scratch), "rwh_free");
    return out;
  } catch (...) {
    (void)rwh_stream_sync(e.stream());
    (void)rwh_free(scratch);
    throw;
  }
}

}  // namespace redwood::hip
