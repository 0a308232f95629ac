// The redwood backend-core types the tree app is written against
// (reference: redwood/base_appdata.hpp:6-21): BaseAppData holding the memory
// resource, and UsmVector, the buffer type of every AppData field.
//
// The reference's UsmVector is a std::pmr::vector over unified memory, which
// the host may index directly.  On a discrete MI355X the buffers live in HBM,
// so UsmVector here is a fixed-size buffer allocated through the resource
// that never touches its elements from the host unless the resource is
// host-visible: a DeviceResource clears and copies through the GPU; any
// other resource (std::pmr::new_delete_resource() for the CPU path) gives
// ordinary host memory with vector-like element access.
#pragma once

#include <cassert>
#include <cstddef>
#include <cstring>
#include <memory_resource>
#include <stdexcept>
#include <utility>
#include <vector>

namespace redwood {

// A memory resource whose memory the host cannot dereference.
class DeviceResource : public std::pmr::memory_resource {
 public:
  virtual void zero(void* p, size_t bytes) = 0;
  virtual void upload(void* dst, const void* src, size_t bytes) = 0;
  virtual void download(void* dst, const void* src, size_t bytes) = 0;
};

template <typename T>
class UsmVector {
 public:
  UsmVector(size_t n, std::pmr::memory_resource* mr)
      : mr_(mr), dev_(dynamic_cast<DeviceResource*>(mr)), n_(n) {
    if (n_ == 0) return;
    p_ = static_cast<T*>(mr_->allocate(n_ * sizeof(T), alignof(T) < 16 ? 16 : alignof(T)));
    if (dev_)
      dev_->zero(p_, n_ * sizeof(T));
    else
      std::memset(static_cast<void*>(p_), 0, n_ * sizeof(T));  // value-initialised, as pmr::vector
  }
  ~UsmVector() { release(); }
  UsmVector(const UsmVector&) = delete;
  UsmVector& operator=(const UsmVector&) = delete;
  UsmVector(UsmVector&& o) noexcept : mr_(o.mr_), dev_(o.dev_), p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }

  T* data() { return p_; }
  const T* data() const { return p_; }
  size_t size() const { return n_; }
  bool on_device() const { return dev_ != nullptr; }

  // Host element access: only for host-visible memory.
  T& operator[](size_t i) {
    host_only();
    return p_[i];
  }
  const T& operator[](size_t i) const {
    host_only();
    return p_[i];
  }
  T* begin() {
    host_only();
    return p_;
  }
  T* end() {
    host_only();
    return p_ + n_;
  }

  // Copies the first count elements out to / in from host memory.
  std::vector<T> to_host(size_t count) const {
    if (count > n_) throw std::out_of_range("UsmVector::to_host");
    std::vector<T> out(count);
    if (count == 0) return out;
    if (dev_)
      dev_->download(out.data(), p_, count * sizeof(T));
    else
      std::memcpy(static_cast<void*>(out.data()), p_, count * sizeof(T));
    return out;
  }
  std::vector<T> to_host() const { return to_host(n_); }
  void from_host(const T* src, size_t count) {
    if (count > n_) throw std::out_of_range("UsmVector::from_host");
    if (count == 0) return;
    if (dev_)
      dev_->upload(p_, src, count * sizeof(T));
    else
      std::memcpy(static_cast<void*>(p_), src, count * sizeof(T));
  }

 private:
  void host_only() const {
    if (dev_) throw std::logic_error("UsmVector: device memory is not host-addressable");
  }
  void release() {
    if (p_) mr_->deallocate(p_, n_ * sizeof(T), alignof(T) < 16 ? 16 : alignof(T));
    p_ = nullptr;
  }

  std::pmr::memory_resource* mr_;
  DeviceResource* dev_;
  T* p_ = nullptr;
  size_t n_;
};

struct BaseAppData {
  explicit BaseAppData(std::pmr::memory_resource* mr) : mr_(mr) {}
  virtual ~BaseAppData() = default;
  std::pmr::memory_resource* mr_;
};

}  // namespace redwood
