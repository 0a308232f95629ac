// The redwood HIP backend: an Engine (device, stream, memory resource,
// typed buffers) and the seven stage dispatchers of the tree app returning
// std::future<void>, the shape the reference documents for every backend
// (README.md:13-16, 24-36, 133-147) and that drops in beside its CPU and CUDA
// dispatchers (apps/tree/host/host_dispatchers.hpp:8-41,
// apps/tree/cuda/cu_dispatcher.cuh:8-19).
//
// Only the C ABI of include/rwhip.h is called; no HIP header is included.
#pragma once

#include <condition_variable>
#include <deque>
#include <functional>
#include <future>
#include <memory>
#include <mutex>
#include <string>
#include <thread>

#include "app_data.hpp"
#include "redwood_base.hpp"
#include "result.hpp"
#include "rwhip.h"

namespace redwood::hip {

// Thrown for a non-zero status of the C ABI.
class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

// Throws Error when status != RWH_OK (the reference's CUDA_CHECK,
// redwood/cuda/helpers.cuh:12-20).
void check(int status, const char* what);

class Engine;

// Device memory resource of an engine: allocations are hipMalloc'd HBM.
class HipMemoryResource final : public redwood::DeviceResource {
 public:
  explicit HipMemoryResource(Engine* engine) : engine_(engine) {}
  Engine& engine() const { return *engine_; }
  void zero(void* p, size_t bytes) override;
  void upload(void* dst, const void* src, size_t bytes) override;
  void download(void* dst, const void* src, size_t bytes) override;

 private:
  void* do_allocate(size_t bytes, size_t alignment) override;
  void do_deallocate(void* p, size_t bytes, size_t alignment) override;
  bool do_is_equal(const std::pmr::memory_resource& other) const noexcept override {
    return this == &other;
  }
  Engine* engine_;
};

template <typename T>
class TypedBuffer {
 public:
  TypedBuffer(size_t n, HipMemoryResource* mr) : v_(n, mr) {}
  T* data() { return v_.data(); }
  const T* data() const { return v_.data(); }
  size_t size() const { return v_.size(); }
  std::vector<T> to_host() const { return v_.to_host(); }
  void from_host(const T* src, size_t n) { v_.from_host(src, n); }

 private:
  redwood::UsmVector<T> v_;
};

class Engine {
 public:
  // Uses the current HIP device; throws Error when there is none.
  static std::unique_ptr<Engine> create();
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  std::pmr::memory_resource* get_mr() { return &mr_; }
  template <typename T>
  std::shared_ptr<TypedBuffer<T>> typed_buffer(size_t n) {
    return std::make_shared<TypedBuffer<T>>(n, &mr_);
  }
  rwh_stream_t stream() const { return stream_; }
  void sync();
  std::string device_name() const;

  // A future that is ready once everything enqueued on the engine's stream
  // so far has finished and then() has run on the engine's completion
  // thread; an exception from then() is stored in the future.
  std::future<void> when_done(std::function<void()> then);

 private:
  Engine();
  void completion_loop();

  HipMemoryResource mr_;
  rwh_stream_t stream_ = nullptr;
  struct Pending {
    rwh_event_t event;
    std::function<void()> then;
    std::promise<void> done;
  };
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<Pending> queue_;
  bool stop_ = false;
  std::thread worker_;
};

using redwood::tree::AppData;

// Stage dispatchers.  Each enqueues its kernels on the engine of
// app_data.mr_ (which must be an engine's get_mr()) and returns at once; the
// future completes when the stage's device work is done.  Stage 3's future
// also sets n_unique / n_brt_nodes, stage 6's sets n_octree_nodes (and fails
// with RWH_ERR_CAPACITY when it exceeds the octree's 0.6 x n_input slots).
std::future<void> run_stage1(AppData& app_data);  // points -> Morton codes
std::future<void> run_stage2(AppData& app_data);  // codes -> sorted codes (u_morton_keys_alt)
std::future<void> run_stage3(AppData& app_data);  // sorted -> unique (u_morton_keys)
std::future<void> run_stage4(AppData& app_data);  // unique -> radix tree
std::future<void> run_stage5(AppData& app_data);  // radix tree -> edge counts
std::future<void> run_stage6(AppData& app_data);  // edge counts -> offsets
std::future<void> run_stage7(AppData& app_data);  // everything -> octree nodes
// Opt-in leaf linking (the reference's never-launched k_LinkLeafNodes).
std::future<void> run_link_leaf(AppData& app_data);

// The build's arrays in the packed result format (include/rwhip.h), after
// the stages run on app_data have finished: packed on the device and copied
// to the host in one transfer (waits for the engine's stream).  sections:
// RWH_SEC_* bits.
redwood::result::PackedResult read_result(AppData& app_data, uint32_t sections = RWH_SECTIONS_ALL);

}  // namespace redwood::hip
