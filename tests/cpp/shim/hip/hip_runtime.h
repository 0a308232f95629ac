// TEST INFRASTRUCTURE ONLY: a host stand-in for <hip/hip_runtime.h> under
// which the library's kernel sources compile with g++ and run on the CPU, one
// OS thread per work-item (tests/cpp/tree_emul.cpp).  Nothing here is part of
// the product, which is built with hipcc for gfx950 only.
//
// Execution model:
//   * a launch runs its workgroups one after another; a workgroup's
//     work-items are std::threads (created once per launch and reused by
//     every workgroup), __syncthreads is a std::barrier over them, and a
//     __shared__ variable is a static (one copy: one workgroup at a time);
//   * wave collectives (ballot, shuffles, readfirstlane) synchronise a
//     "domain" of lanes: the whole 64-lane wave, or, for kernels whose lane
//     groups loop independently, the group (emu::set_domain);
//   * a work-item that returns leaves every barrier it belonged to;
//   * atomics are std::atomic_ref operations.
// Built with AddressSanitizer, every out-of-range index into a __shared__
// array or a buffer aborts the run.
#pragma once

#include <algorithm>
#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <type_traits>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) __attribute__((aligned(n)))
#define __HIP_MEMORY_SCOPE_AGENT 0

using hipStream_t = void*;
using hipError_t = int;
constexpr hipError_t hipSuccess = 0;
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 0 };

struct dim3 {
  uint32_t x, y, z;
  constexpr dim3(uint32_t x_ = 1, uint32_t y_ = 1, uint32_t z_ = 1) : x(x_), y(y_), z(z_) {}
};
struct uint2 { uint32_t x, y; };
struct int2 { int32_t x, y; };
struct uint4 { uint32_t x, y, z, w; };
struct int4 { int32_t x, y, z, w; };
struct float4 { float x, y, z, w; };
inline uint2 make_uint2(uint32_t x, uint32_t y) { return {x, y}; }
inline int4 make_int4(int32_t x, int32_t y, int32_t z, int32_t w) { return {x, y, z, w}; }
inline uint4 make_uint4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) { return {x, y, z, w}; }
inline float4 make_float4(float x, float y, float z, float w) { return {x, y, z, w}; }

namespace emu {

constexpr uint32_t kWave = 64;

struct Domain {
  std::unique_ptr<std::barrier<>> bar;
};

struct Block {
  uint32_t threads = 0, domain = kWave;
  std::unique_ptr<std::barrier<>> sync;
  std::vector<Domain> domains;
  std::vector<uint64_t> slot;     // one exchange value per work-item
  std::vector<uint8_t> present;   // 1 while the work-item runs
};

inline thread_local dim3 t_thread, t_block;
inline dim3 g_grid, g_blockdim;
inline Block* g_cur = nullptr;
inline std::map<const void*, uint32_t>& domains() {
  static std::map<const void*, uint32_t> m;
  return m;
}
// Kernels whose lane groups of `lanes` lanes run collectives independently.
template <class K>
void set_domain(K kernel, uint32_t lanes) { domains()[reinterpret_cast<const void*>(kernel)] = lanes; }

inline uint32_t tid() { return t_thread.x; }
inline Domain& my_domain() { return g_cur->domains[tid() / g_cur->domain]; }
inline uint32_t domain_first() { return tid() / g_cur->domain * g_cur->domain; }

// Every work-item of the domain publishes v; returns the values by lane.
template <class F>
inline auto exchange(uint64_t v, F&& read) {
  Block& b = *g_cur;
  b.slot[tid()] = v;
  my_domain().bar->arrive_and_wait();
  auto r = read(b, domain_first());
  my_domain().bar->arrive_and_wait();
  return r;
}

inline uint64_t ballot(bool p) {
  return exchange(p ? 1u : 0u, [](Block& b, uint32_t first) {
    uint64_t m = 0;
    const uint32_t wave_first = first / kWave * kWave;
    for (uint32_t t = first; t < first + b.domain && t < b.threads; ++t)
      if (b.present[t] && b.slot[t]) m |= 1ull << (t - wave_first);
    return m;
  });
}

inline uint64_t read_lane(uint64_t v, uint32_t lane_in_wave) {
  return exchange(v, [lane_in_wave](Block& b, uint32_t first) {
    const uint32_t src = first / kWave * kWave + lane_in_wave;
    if (src < first || src >= first + b.domain || src >= b.threads || !b.present[src]) {
      std::fprintf(stderr, "emu: shuffle from lane %u outside the domain or finished\n", lane_in_wave);
      std::abort();
    }
    return b.slot[src];
  });
}

inline uint64_t first_lane(uint64_t v) {
  return exchange(v, [](Block& b, uint32_t first) {
    for (uint32_t t = first; t < first + b.domain && t < b.threads; ++t)
      if (b.present[t]) return b.slot[t];
    return uint64_t{0};
  });
}

inline void sync_block() { g_cur->sync->arrive_and_wait(); }

inline void leave() {
  Block& b = *g_cur;
  b.present[tid()] = 0;
  b.slot[tid()] = 0;
  my_domain().bar->arrive_and_drop();
  b.sync->arrive_and_drop();
}

template <class K, class... A>
void launch(K kernel, dim3 grid, dim3 block, A... args) {
  g_grid = grid;
  g_blockdim = block;
  const uint32_t n = block.x * block.y * block.z;
  uint32_t dom = kWave;
  if (auto it = domains().find(reinterpret_cast<const void*>(kernel)); it != domains().end()) dom = it->second;
  // one thread per work-item for the whole launch; the workgroups run one
  // after another, each between two meetings of the threads with this one
  Block b;
  dim3 cur;
  bool quit = false;
  std::barrier<> start(n + 1), done(n + 1);
  std::vector<std::thread> th;
  th.reserve(n);
  for (uint32_t t = 0; t < n; ++t)
    th.emplace_back([&, t] {
      t_thread = dim3(t, 0, 0);
      for (;;) {
        start.arrive_and_wait();
        if (quit) return;
        t_block = cur;
        kernel(args...);
        leave();
        done.arrive_and_wait();
      }
    });
  for (uint32_t bz = 0; bz < grid.z; ++bz)
    for (uint32_t by = 0; by < grid.y; ++by)
      for (uint32_t bx = 0; bx < grid.x; ++bx) {
        b.threads = n;
        b.domain = dom;
        b.sync = std::make_unique<std::barrier<>>(n);
        b.domains.clear();
        for (uint32_t d = 0; d < (n + dom - 1) / dom; ++d)
          b.domains.push_back(Domain{std::make_unique<std::barrier<>>(std::min(dom, n - d * dom))});
        b.slot.assign(n, 0);
        b.present.assign(n, 1);
        g_cur = &b;
        cur = dim3(bx, by, bz);
        start.arrive_and_wait();
        done.arrive_and_wait();
        g_cur = nullptr;
      }
  quit = true;
  start.arrive_and_wait();
  for (auto& x : th) x.join();
}

}  // namespace emu

#define threadIdx (emu::t_thread)
#define blockIdx (emu::t_block)
#define blockDim (emu::g_blockdim)
#define gridDim (emu::g_grid)
#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) emu::launch(kernel, dim3(grid), dim3(block), __VA_ARGS__)

inline void __syncthreads() { emu::sync_block(); }
inline void __threadfence() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline uint64_t __ballot(bool p) { return emu::ballot(p); }

template <class T>
inline T __shfl(T v, int src, int width = 64) {
  const uint32_t lane = emu::tid() % 64;
  const uint32_t s = (lane / width) * width + static_cast<uint32_t>(src) % width;
  return static_cast<T>(emu::read_lane(static_cast<uint64_t>(v), s));
}
template <class T>
inline T __shfl_up(T v, unsigned d, int width = 64) {
  const uint32_t lane = emu::tid() % 64;
  const uint32_t base = (lane / width) * width;
  const uint32_t s = (lane - base >= d) ? lane - d : lane;
  return static_cast<T>(emu::read_lane(static_cast<uint64_t>(v), s));
}
template <class T>
inline T __shfl_xor(T v, int m, int width = 64) {
  const uint32_t lane = emu::tid() % 64;
  return static_cast<T>(emu::read_lane(static_cast<uint64_t>(v), lane ^ static_cast<uint32_t>(m)));
}

inline int __clz(uint32_t x) { return x ? __builtin_clz(x) : 32; }
inline int __popc(uint32_t x) { return __builtin_popcount(x); }
inline int __popcll(uint64_t x) { return __builtin_popcountll(x); }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline float __fsub_rn(float a, float b) { return a - b; }
inline float __fadd_rn(float a, float b) { return a + b; }
inline float __fmul_rn(float a, float b) { return a * b; }
inline float __fdiv_rn(float a, float b) { return a / b; }

inline uint32_t __builtin_amdgcn_mbcnt_lo(uint32_t m, uint32_t acc) {
  const uint32_t lane = emu::tid() % 64;
  const uint32_t below = lane >= 32 ? 0xffffffffu : ((1u << lane) - 1u);
  return acc + static_cast<uint32_t>(__builtin_popcount(m & below));
}
inline uint32_t __builtin_amdgcn_mbcnt_hi(uint32_t m, uint32_t acc) {
  const uint32_t lane = emu::tid() % 64;
  const uint32_t below = lane < 32 ? 0u : ((1u << (lane - 32)) - 1u);
  return acc + static_cast<uint32_t>(__builtin_popcount(m & below));
}
inline uint32_t __builtin_amdgcn_readfirstlane(uint32_t v) { return static_cast<uint32_t>(emu::first_lane(v)); }
inline int __builtin_amdgcn_readlane(int v, int lane) {
  return static_cast<int>(emu::read_lane(static_cast<uint32_t>(v), static_cast<uint32_t>(lane)));
}
// v_perm_b32: byte i of the result selected from the 8 bytes {a, b} (b low) by sel byte i
inline uint32_t __builtin_amdgcn_perm(uint32_t a, uint32_t b, uint32_t sel) {
  const uint64_t v = (static_cast<uint64_t>(a) << 32) | b;
  uint32_t r = 0;
  for (int i = 0; i < 4; ++i) {
    const uint32_t s = (sel >> (8 * i)) & 0xffu;
    uint32_t byte;
    if (s < 8) byte = static_cast<uint32_t>((v >> (8 * s)) & 0xffu);
    else if (s == 12) byte = 0;
    else if (s > 12) byte = 0xff;
    else { std::fprintf(stderr, "emu: perm selector %u not modelled\n", s); std::abort(); }
    r |= byte << (8 * i);
  }
  return r;
}
inline int __builtin_amdgcn_sbfe(int v, int off, int w) {
  return static_cast<int>(static_cast<int64_t>(static_cast<uint64_t>(static_cast<uint32_t>(v)) << (64 - off - w)) >> (64 - w));
}
inline uint64_t __builtin_amdgcn_uicmp(uint32_t a, uint32_t b, int op) {
  return emu::ballot(op == 33 ? a != b : a == b);
}
inline uint32_t __builtin_amdgcn_bitop3_b32(uint32_t a, uint32_t b, uint32_t c, uint32_t table) {
  uint32_t r = 0;
  for (int i = 0; i < 32; ++i) {
    const uint32_t idx = (((a >> i) & 1u) << 2) | (((b >> i) & 1u) << 1) | ((c >> i) & 1u);
    r |= ((table >> idx) & 1u) << i;
  }
  return r;
}
inline void __builtin_amdgcn_s_waitcnt(int) {}
// on the GPU the lanes of a wave run in lockstep (this builtin only orders the
// compiler); here they are threads, so the domain meets at a barrier
inline void __builtin_amdgcn_wave_barrier() { emu::my_domain().bar->arrive_and_wait(); }
// a cache hint only: a plain load on the CPU
template <class T>
inline T __builtin_nontemporal_load(const T* p) { return *p; }
template <class T>
inline void __builtin_nontemporal_store(T v, T* p) { *p = v; }

template <class T, class U>
inline T atomicAdd(T* p, U v) { return std::atomic_ref<T>(*p).fetch_add(static_cast<T>(v)); }
template <class T, class U>
inline T atomicExch(T* p, U v) { return std::atomic_ref<T>(*p).exchange(static_cast<T>(v)); }
template <class T, class U>
inline T atomicOr(T* p, U v) { return std::atomic_ref<T>(*p).fetch_or(static_cast<T>(v)); }
template <class T, class U>
inline T atomicAnd(T* p, U v) { return std::atomic_ref<T>(*p).fetch_and(static_cast<T>(v)); }
template <class T, class U>
inline T atomicMax(T* p, U v) {
  std::atomic_ref<T> a(*p);
  T cur = a.load();
  while (cur < static_cast<T>(v) && !a.compare_exchange_weak(cur, static_cast<T>(v))) {
  }
  return cur;
}
template <class T, class U>
inline void __hip_atomic_store(T* p, U v, int, int) { std::atomic_ref<T>(*p).store(static_cast<T>(v), std::memory_order_relaxed); }
template <class T>
inline T __hip_atomic_load(T* p, int, int) { return std::atomic_ref<T>(*p).load(std::memory_order_relaxed); }

template <class A, class B>
inline std::common_type_t<A, B> min(A a, B b) { return a < b ? a : b; }
template <class A, class B>
inline std::common_type_t<A, B> max(A a, B b) { return a < b ? b : a; }

inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
inline hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) { *v = 1; return hipSuccess; }
template <class K>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, K, int, size_t) { *n = 1; return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char* hipGetErrorString(hipError_t) { return "emulated"; }
