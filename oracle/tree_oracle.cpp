// TEST INFRASTRUCTURE ONLY — the CPU oracle of the seven-stage octree build.
//
// This file restates the reference CPU path (apps/tree/host/*) in plain C++20
// with a pthread pool, under REF-SPEC:
//   * the per-element arithmetic of every stage follows the reference line for
//     line in meaning (each function cites the lines it follows);
//   * the stage wiring follows the documented intent (app_data.hpp:139-146) and
//     the CUDA dispatcher where the CPU wiring is undefined (D1-D8);
//   * every multi-writer slot is resolved deterministically (atomic OR for masks,
//     atomic max for index slots), so results do not depend on the thread count.
// Only tests, the smoke entry point and the benchmark's CPU-baseline leg use it.
//
// Parity pins (all in tests/): S1 known answers + bit-loop cross-check; S2
// permutation + order + the reference's own func_sort.cpp compiled from
// source; S3 numpy.unique; S4 brute-force top-down construction on tiny inputs
// + the reference's own func_brt.hpp compiled from source; S5 the reference's
// func_edge.hpp compiled from source + closed form; S6 numpy.cumsum; S7
// structural invariants (single owner per slot, mask/child consistency, cell
// size closed form, corner = decoded prefix) + thread-count invariance.

#include "tree_oracle.h"

#include <algorithm>
#include <atomic>
#include <bit>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace oracle {

// ---------------------------------------------------------------------------
// Pool: persistent workers; run(nb, fn) executes fn(0..nb-1) and waits.
// The block partition used by the stages is the reference's submit_blocks
// (redwood/host/thread_pool.hpp:137-162): M blocks of ceil(n/M) items.
// ---------------------------------------------------------------------------
class Pool {
 public:
  explicit Pool(unsigned n) : n_(n ? n : 1) {
    for (unsigned t = 1; t < n_; ++t) workers_.emplace_back([this] { loop(); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& w : workers_) w.join();
  }
  unsigned size() const { return n_; }

  void run(unsigned nb, const std::function<void(unsigned)>& fn) {
    if (nb == 0) return;
    if (n_ == 1 || nb == 1) {
      for (unsigned b = 0; b < nb; ++b) fn(b);
      return;
    }
    {
      std::lock_guard<std::mutex> g(mu_);
      job_ = &fn;
      nb_ = nb;
      next_.store(0);
      done_ = 0;
      ++gen_;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return done_ == nb_; });
    job_ = nullptr;
  }

  // fn(start, end) over the reference's submit_blocks partition.
  template <class F>
  void blocks(int64_t first, int64_t last, F&& fn) {
    if (last <= first) return;
    const int64_t m = n_;
    const int64_t bs = (last - first + m - 1) / m;
    const unsigned nb = static_cast<unsigned>((last - first + bs - 1) / bs);
    run(nb, [&](unsigned b) {
      const int64_t s = first + b * bs;
      fn(s, std::min(s + bs, last));
    });
  }

 private:
  void work() {
    for (;;) {
      const unsigned b = next_.fetch_add(1);
      if (b >= nb_) return;
      (*job_)(b);
      std::lock_guard<std::mutex> g(mu_);
      if (++done_ == nb_) done_cv_.notify_all();
    }
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
      }
      work();
    }
  }

  unsigned n_;
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(unsigned)>* job_ = nullptr;
  unsigned nb_ = 0;
  std::atomic<unsigned> next_{0};
  unsigned done_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// ---------------------------------------------------------------------------
// S1  Morton encode — func_morton.hpp:15-39.
// Quantisation is u32((v - min) / range * 1024) (float sub, div, mul, each
// rounded; the build uses -ffp-contract=off), then the low 10 bits (Q8).
// The interleave is written as a plain bit loop here (x -> bit 3b, y -> 3b+1,
// z -> 3b+2), a second formulation of the reference's magic-number spread.
// ---------------------------------------------------------------------------
// The float -> u32 conversion truncates toward zero.  Outside [0, 2^32) it is
// undefined in the reference; here (D10) it saturates as the GPU's conversion
// does: NaN and negatives give 0, values >= 2^32 give 2^32 - 1.
static inline uint32_t quantize(float v, float min_coord, float range) {
  const float t = (v - min_coord) / range;
  const float s = t * 1024.0f;
  if (!(s > 0.0f)) return 0u;
  if (s >= 4294967296.0f) return 0xffffffffu;
  return static_cast<uint32_t>(static_cast<int64_t>(s));
}

static inline uint32_t interleave3(uint32_t i, uint32_t j, uint32_t k) {
  uint32_t code = 0;
  for (int b = 0; b < 10; ++b) {
    code |= ((i >> b) & 1u) << (3 * b);
    code |= ((j >> b) & 1u) << (3 * b + 1);
    code |= ((k >> b) & 1u) << (3 * b + 2);
  }
  return code;
}

static inline uint32_t encode(const float* p, float min_coord, float range) {
  return interleave3(quantize(p[0], min_coord, range) & 0x3ffu,
                     quantize(p[1], min_coord, range) & 0x3ffu,
                     quantize(p[2], min_coord, range) & 0x3ffu);
}

// Decode — func_morton.hpp:45-80: corner = (float(d) / 1024) * range + min, w = 1 (Q7).
static inline void decode(uint32_t code, float min_coord, float range, float* out) {
  uint32_t d[3] = {0, 0, 0};
  for (int b = 0; b < 10; ++b)
    for (int a = 0; a < 3; ++a) d[a] |= ((code >> (3 * b + a)) & 1u) << b;
  for (int a = 0; a < 3; ++a) {
    const float q = static_cast<float>(d[a]) / 1024.0f;
    const float r = q * range;
    out[a] = r + min_coord;
  }
  out[3] = 1.0f;
}

// ---------------------------------------------------------------------------
// S2  LSD radix sort — func_sort.cpp:117-211, four 8-bit passes (shift 0, 8,
// 16, 24).  Each worker owns one contiguous block of the input; per pass it
// histograms its block, the start of (digit, block) is the exclusive scan in
// digit-major / block-minor order (the ordered hand-off of func_sort.cpp:152-164
// computes exactly this), and every block scatters its keys in input order.
// ---------------------------------------------------------------------------
static void sort_keys(Pool& pool, const uint32_t* in, uint32_t* out, uint32_t* tmp,
                      uint32_t n) {
  if (n == 0) return;
  const unsigned nb = std::min<unsigned>(pool.size(), n);
  // block partition of my_blocks (func_sort.cpp:20-68): sizes n/nb, first n%nb get +1
  std::vector<uint32_t> bstart(nb + 1);
  {
    const uint32_t bs = n / nb, rem = n % nb;
    for (unsigned b = 0; b < nb; ++b) bstart[b] = b * bs + std::min<uint32_t>(b, rem);
    bstart[nb] = n;
  }
  std::vector<uint32_t> hist(static_cast<size_t>(nb) * 256);
  // pass buffers: in -> tmp -> out -> tmp -> out
  const uint32_t* src[4] = {in, tmp, out, tmp};
  uint32_t* dst[4] = {tmp, out, tmp, out};
  for (int pass = 0; pass < 4; ++pass) {
    const int shift = 8 * pass;
    const uint32_t* s = src[pass];
    uint32_t* d = dst[pass];
    pool.run(nb, [&](unsigned b) {
      uint32_t* h = &hist[static_cast<size_t>(b) * 256];
      std::fill(h, h + 256, 0u);
      for (uint32_t i = bstart[b]; i < bstart[b + 1]; ++i) ++h[(s[i] >> shift) & 255u];
    });
    uint32_t run = 0;
    for (int dg = 0; dg < 256; ++dg)
      for (unsigned b = 0; b < nb; ++b) {
        uint32_t& h = hist[static_cast<size_t>(b) * 256 + dg];
        const uint32_t c = h;
        h = run;
        run += c;
      }
    pool.run(nb, [&](unsigned b) {
      uint32_t* h = &hist[static_cast<size_t>(b) * 256];
      for (uint32_t i = bstart[b]; i < bstart[b + 1]; ++i) {
        const uint32_t k = s[i];
        d[h[(k >> shift) & 255u]++] = k;
      }
    });
  }
}

// ---------------------------------------------------------------------------
// S4  Karras binary radix tree — func_brt.hpp:27-138 (CUDA 04_radix_tree.cu:12-119).
// ---------------------------------------------------------------------------
// Q1: delta = clz(a ^ b) - 1 kept as uint8_t (func_brt.hpp:110-116); D3: clz(0) = 32.
static inline int delta8(uint32_t a, uint32_t b) {
  return static_cast<uint8_t>(std::countl_zero(a ^ b) - 1);
}
// func_brt.hpp:118-126; D3: log2_ceil(0) = 0.
static inline int log2_ceil(uint32_t x) {
  if (x == 0) return 0;
  const int fl = 31 - std::countl_zero(x);
  return fl + ((1u << fl) < x ? 1 : 0);
}

static inline void atomic_max_i32(int32_t* p, int32_t v) {
  int32_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
  while (cur < v &&
         !__atomic_compare_exchange_n(p, &cur, v, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
  }
}
static inline void atomic_or_i32(int32_t* p, int32_t v) {
  __atomic_fetch_or(p, v, __ATOMIC_RELAXED);
}

// One node, n = n_brt (func_brt.hpp:45-138).
static void brt_node(int i, int n, const uint32_t* codes, uint8_t* prefix_n,
                     uint8_t* has_leaf_left, uint8_t* has_leaf_right, int32_t* left_child,
                     int32_t* parent) {
  const uint32_t ci = codes[i];
  int d;
  if (i == 0) {
    d = 1;
  } else {
    const int diff = delta8(ci, codes[i + 1]) - delta8(ci, codes[i - 1]);
    d = (diff > 0) - (diff < 0);
  }

  int l = 0;
  if (i == 0) {
    l = n - 1;  // Q2: the root's range stops one key short
  } else {
    const int dmin = delta8(ci, codes[i - d]);
    int64_t lmax = 2;
    while (i + lmax * d >= 0 && i + lmax * d <= n &&
           delta8(ci, codes[i + lmax * d]) > dmin)
      lmax *= 2;
    const int cutoff = (d == -1) ? i : n - i;
    for (int64_t t = lmax / 2, div = 2; t >= 1; div *= 2, t = lmax / div)
      if (l + t <= cutoff && delta8(ci, codes[i + (l + t) * d]) > dmin) l += static_cast<int>(t);
  }
  const int j = i + l * d;

  const int dnode = delta8(codes[i], codes[j]);
  prefix_n[i] = static_cast<uint8_t>(dnode);
  int s = 0;
  const uint32_t max_div = 1u << log2_ceil(static_cast<uint32_t>(l));
  const int scut = (d == -1) ? i - 1 : n - i - 1;
  const auto ceil_div = [](uint32_t a, uint32_t b) { return (a + b - 1) / b; };
  for (uint32_t div = 2, t = ceil_div(static_cast<uint32_t>(l), 2); div <= max_div;
       div <<= 1, t = ceil_div(static_cast<uint32_t>(l), div)) {
    if (static_cast<int64_t>(s) + t <= scut &&
        delta8(ci, codes[i + (s + static_cast<int>(t)) * d]) > dnode)
      s += static_cast<int>(t);
  }
  const int gamma = i + s * d + std::min(d, 0);
  left_child[i] = gamma;
  const bool ll = std::min(i, j) == gamma;
  const bool lr = std::max(i, j) == gamma + 1;
  has_leaf_left[i] = ll ? 1 : 0;
  has_leaf_right[i] = lr ? 1 : 0;
  // D7: a parent slot written by several nodes keeps the largest writer index.
  if (!ll) atomic_max_i32(&parent[gamma], i);
  if (!lr) atomic_max_i32(&parent[gamma + 1], i);
}

// ---------------------------------------------------------------------------
// S7  octree nodes — func_octree.hpp:9-112 (CUDA 07_octree.cu:11-114, 175-226).
// ---------------------------------------------------------------------------
struct OctOut {
  int32_t* children;
  float* corner;
  float* cell;
  int32_t* mask;
  int64_t n_oct;
  std::atomic<int64_t>* dropped;
};

// set_child (func_octree.hpp:9-16) with D6: child slot = max, mask = atomic OR.
static inline void set_child(const OctOut& o, int64_t node, uint32_t which, int32_t oct) {
  if (node < 0 || node >= o.n_oct) {
    o.dropped->fetch_add(1, std::memory_order_relaxed);
    return;
  }
  atomic_max_i32(&o.children[node * 8 + which], oct);
  atomic_or_i32(&o.mask[node], static_cast<int32_t>(1u << which));
}

static inline void set_corner_cell(const OctOut& o, int64_t slot, uint32_t code, float min_coord,
                                   float range, int level_diff) {
  if (slot < 0 || slot >= o.n_oct) {
    o.dropped->fetch_add(1, std::memory_order_relaxed);
    return;
  }
  decode(code, min_coord, range, &o.corner[slot * 4]);
  // cell = range / float(1 << (level - root_level)); the shift count is taken
  // mod 32 as the hardware does (it is in [0, 10] for every reachable input).
  o.cell[slot] = range / static_cast<float>(1u << (static_cast<uint32_t>(level_diff) & 31u));
}

static void oct_node(int i, const OctOut& o, const int32_t* offsets, const int32_t* counts,
                     const uint32_t* codes, const uint8_t* prefix_n, const int32_t* parent,
                     float min_coord, float range) {
  constexpr int kBits = 30;
  int64_t oct = offsets[i];
  const int n_new = counts[i];
  const int root_level = prefix_n[0] / 3;
  for (int j = 0; j < n_new - 1; ++j) {
    const int level = prefix_n[i] / 3 - j;
    const uint32_t sh = static_cast<uint32_t>(kBits - 3 * level) & 31u;
    const uint32_t node_prefix = codes[i] >> sh;
    const uint32_t which = node_prefix & 7u;
    const int64_t up = oct + 1;
    set_child(o, up, which, static_cast<int32_t>(oct));
    set_corner_cell(o, oct, node_prefix << sh, min_coord, range, level - root_level);
    oct = up;
  }
  if (n_new > 0) {
    int rt = parent[i];
    int steps = 0;
    while (counts[rt] == 0) {  // Q5: nearest contributing ancestor, capped walk
      rt = parent[rt];
      if (++steps > 30) break;
    }
    const int64_t oct_parent = offsets[rt];
    const int top_level = prefix_n[i] / 3 - n_new + 1;
    const uint32_t sh = static_cast<uint32_t>(kBits - 3 * top_level) & 31u;
    const uint32_t top_prefix = codes[i] >> sh;
    set_child(o, oct_parent, top_prefix & 7u, static_cast<int32_t>(oct));
    set_corner_cell(o, oct, top_prefix << sh, min_coord, range, top_level - root_level);
  }
}

}  // namespace oracle

using oracle::Pool;

extern "C" {

void* oracle_pool_create(int n_threads) {
  return new Pool(static_cast<unsigned>(n_threads > 0 ? n_threads : 1));
}
void oracle_pool_destroy(void* pool) { delete static_cast<Pool*>(pool); }
int oracle_pool_threads(void* pool) { return static_cast<int>(static_cast<Pool*>(pool)->size()); }

void oracle_morton(void* pool, const float* xyzw, uint32_t n, float min_coord, float range,
                   uint32_t* codes) {
  // host_dispatchers.cpp:17-41
  static_cast<Pool*>(pool)->blocks(0, n, [&](int64_t s, int64_t e) {
    for (int64_t i = s; i < e; ++i) codes[i] = oracle::encode(&xyzw[i * 4], min_coord, range);
  });
}

void oracle_sort(void* pool, const uint32_t* in, uint32_t* out, uint32_t* tmp, uint32_t n) {
  oracle::sort_keys(*static_cast<Pool*>(pool), in, out, tmp, n);
}

uint32_t oracle_unique(const uint32_t* sorted, uint32_t n, uint32_t* out) {
  // host_dispatchers.cpp:183-194 (std::unique_copy: first of every run of equal keys)
  uint32_t m = 0;
  for (uint32_t i = 0; i < n; ++i)
    if (i == 0 || sorted[i] != sorted[i - 1]) out[m++] = sorted[i];
  return m;
}

void oracle_radix_tree(void* pool, const uint32_t* codes, uint32_t n_unique, uint8_t* prefix_n,
                       uint8_t* has_leaf_left, uint8_t* has_leaf_right, int32_t* left_child,
                       int32_t* parent) {
  // D1: consumes unique keys; D2: i in [0, n_brt); D7: parent zeroed, max writer.
  const int64_t n_brt = n_unique > 0 ? static_cast<int64_t>(n_unique) - 1 : 0;
  std::memset(parent, 0, sizeof(int32_t) * std::max<uint32_t>(n_unique, 1));
  static_cast<Pool*>(pool)->blocks(0, n_brt, [&](int64_t s, int64_t e) {
    for (int64_t i = s; i < e; ++i)
      oracle::brt_node(static_cast<int>(i), static_cast<int>(n_brt), codes, prefix_n,
                       has_leaf_left, has_leaf_right, left_child, parent);
  });
}

void oracle_edge_count(void* pool, const uint8_t* prefix_n, const int32_t* parent, uint32_t n_brt,
                       int32_t* edge_count) {
  // func_edge.hpp:9-16; count[0] = 0 as 05_edge_count.cu:26-28 (and parent[0] = 0).
  static_cast<Pool*>(pool)->blocks(0, n_brt, [&](int64_t s, int64_t e) {
    for (int64_t i = s; i < e; ++i)
      edge_count[i] = prefix_n[i] / 3 - prefix_n[parent[i]] / 3;
  });
  if (n_brt > 0) edge_count[0] = 0;
}

int64_t oracle_prefix_sum(const int32_t* edge_count, uint32_t n_brt, int32_t* edge_offset) {
  // D4: one exclusive scan over all nodes (06_prefix_sum.cu:26-37); total = sum.
  int64_t run = 0;
  for (uint32_t i = 0; i < n_brt; ++i) {
    edge_offset[i] = static_cast<int32_t>(run);
    run += edge_count[i];
  }
  return run;
}

int64_t oracle_octree(void* pool, const int32_t* edge_offset, const int32_t* edge_count,
                      const uint32_t* codes, const uint8_t* prefix_n, const int32_t* parent,
                      uint32_t n_brt, uint32_t n_oct, uint32_t oct_cap, float min_coord,
                      float range, int32_t* children, float* corner, float* cell_size,
                      int32_t* child_node_mask, int32_t* child_leaf_mask) {
  const int64_t rows = std::min<int64_t>(n_oct, oct_cap);
  std::memset(children, 0, sizeof(int32_t) * 8 * rows);
  std::memset(child_node_mask, 0, sizeof(int32_t) * rows);
  std::memset(child_leaf_mask, 0, sizeof(int32_t) * rows);
  if (n_brt == 0) return 0;
  // D8: root corner / cell (07_octree.cu:188-198) before any node is built.
  if (oct_cap > 0) {
    const int root_level = prefix_n[0] / 3;
    const uint32_t sh = static_cast<uint32_t>(30 - 3 * root_level) & 31u;
    const uint32_t root_prefix = codes[0] >> sh;
    oracle::decode(root_prefix << sh, min_coord, range, &corner[0]);
    cell_size[0] = range;
  }
  std::atomic<int64_t> dropped{0};
  const oracle::OctOut o{children, corner, cell_size, child_node_mask, rows, &dropped};
  // D5: brt nodes i in [1, n_brt) (07_octree.cu:208-211).
  static_cast<Pool*>(pool)->blocks(1, n_brt, [&](int64_t s, int64_t e) {
    for (int64_t i = s; i < e; ++i)
      oracle::oct_node(static_cast<int>(i), o, edge_offset, edge_count, codes, prefix_n, parent,
                       min_coord, range);
  });
  return dropped.load();
}

int64_t oracle_link_leaf(void* pool, const int32_t* edge_offset, const int32_t* edge_count,
                         const uint32_t* codes, const uint8_t* has_leaf_left,
                         const uint8_t* has_leaf_right, const uint8_t* prefix_n,
                         const int32_t* parent, const int32_t* left_child, uint32_t n_brt,
                         uint32_t n_oct, int32_t* children, int32_t* child_leaf_mask) {
  std::atomic<int64_t> dropped{0};
  const auto link = [&](int i, int leaf) {
    const int leaf_level = prefix_n[i] / 3 + 1;
    const uint32_t sh = static_cast<uint32_t>(30 - 3 * leaf_level) & 31u;
    const uint32_t which = (codes[leaf] >> sh) & 7u;
    int rt = i;
    int steps = 0;
    while (edge_count[rt] == 0) {
      rt = parent[rt];
      if (++steps > 30) break;
    }
    const int64_t bottom = edge_offset[rt];
    if (bottom < 0 || bottom >= static_cast<int64_t>(n_oct)) {
      dropped.fetch_add(1);
      return;
    }
    oracle::atomic_max_i32(&children[bottom * 8 + which], leaf);
    // set_leaf clears the bit (func_octree.hpp:24), kept as written.
    __atomic_fetch_and(&child_leaf_mask[bottom], ~static_cast<int32_t>(1u << which),
                       __ATOMIC_RELAXED);
  };
  static_cast<Pool*>(pool)->blocks(0, n_brt, [&](int64_t s, int64_t e) {
    for (int64_t i = s; i < e; ++i) {
      if (has_leaf_left[i]) link(static_cast<int>(i), left_child[i]);
      if (has_leaf_right[i]) link(static_cast<int>(i), left_child[i] + 1);
    }
  });
  return dropped.load();
}

int oracle_build(void* pool, const float* xyzw, uint32_t n, float min_coord, float range,
                 uint32_t* codes, uint32_t* sorted, uint32_t* unique_keys, uint32_t* tmp,
                 uint8_t* prefix_n, uint8_t* has_leaf_left, uint8_t* has_leaf_right,
                 int32_t* left_child, int32_t* parent, int32_t* edge_count,
                 int32_t* edge_offset, uint32_t oct_cap, int32_t* children, float* corner,
                 float* cell_size, int32_t* child_node_mask, int32_t* child_leaf_mask,
                 uint32_t* counts_out, double* stage_ms) {
  using clk = std::chrono::steady_clock;
  double ms[7] = {0, 0, 0, 0, 0, 0, 0};
  auto t0 = clk::now();
  const auto lap = [&](int s) {
    const auto t1 = clk::now();
    ms[s] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    t0 = t1;
  };
  oracle_morton(pool, xyzw, n, min_coord, range, codes);
  lap(0);
  oracle_sort(pool, codes, sorted, tmp, n);
  lap(1);
  const uint32_t n_unique = oracle_unique(sorted, n, unique_keys);
  lap(2);
  const uint32_t n_brt = n_unique > 0 ? n_unique - 1 : 0;
  oracle_radix_tree(pool, unique_keys, n_unique, prefix_n, has_leaf_left, has_leaf_right,
                    left_child, parent);
  lap(3);
  oracle_edge_count(pool, prefix_n, parent, n_brt, edge_count);
  lap(4);
  const int64_t n_oct = oracle_prefix_sum(edge_count, n_brt, edge_offset);
  lap(5);
  int rc = 0;
  if (n_oct < 0 || n_oct > static_cast<int64_t>(oct_cap)) rc = -1;
  oracle_octree(pool, edge_offset, edge_count, unique_keys, prefix_n, parent, n_brt,
                static_cast<uint32_t>(std::max<int64_t>(0, std::min<int64_t>(n_oct, oct_cap))),
                oct_cap, min_coord, range, children, corner, cell_size, child_node_mask,
                child_leaf_mask);
  lap(6);
  if (counts_out) {
    counts_out[0] = n_unique;
    counts_out[1] = n_brt;
    counts_out[2] = static_cast<uint32_t>(n_oct);
  }
  if (stage_ms)
    for (int s = 0; s < 7; ++s) stage_ms[s] = ms[s];
  return rc;
}

}  // extern "C"
