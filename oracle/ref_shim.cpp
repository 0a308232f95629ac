// TEST INFRASTRUCTURE ONLY — a translation unit that instantiates the
// reference's own header-only CPU kernels, so the oracle can be checked
// against the reference code itself.  It is compiled by oracle/Makefile
// (target "ref") with -I pointing at the reference tree, into oracle/_ref/.
// Only the reference files that build standalone are used:
//   apps/tree/host/func_brt.hpp   (process_radix_tree_i, S4)
//   apps/tree/host/func_edge.hpp  (process_edge_count_i, S5)
//   apps/tree/host/func_sort.cpp  (dispatch_binning_pass, S2; compiled alongside)
//   redwood/host/thread_pool.hpp  (core::thread_pool used by the sort)
// func_morton.hpp / func_octree.hpp need glm, which this image lacks: S1 and S7
// are not built from the reference (no stand-in headers are written for it).
#include <barrier>
#include <cstdint>
#include <cstring>

#include "apps/tree/host/func_brt.hpp"
#include "apps/tree/host/func_edge.hpp"
#include "apps/tree/host/func_sort.hpp"
#include "redwood/host/thread_pool.hpp"

extern "C" {

// Stage 4 of the reference on unique keys: nodes i in [0, n_brt) run in
// ascending order on one thread (so a parent slot written twice keeps the
// larger writer index); parent[] is zeroed first, as the reference's
// value-initialised buffers are on a first build.
void ref_radix_tree(const uint32_t* codes, uint32_t n_unique, uint8_t* prefix_n,
                    uint8_t* has_leaf_left, uint8_t* has_leaf_right, int32_t* left_child,
                    int32_t* parent) {
  const int n_brt = n_unique > 0 ? static_cast<int>(n_unique) - 1 : 0;
  std::memset(parent, 0, sizeof(int32_t) * (n_unique > 0 ? n_unique : 1));
  for (int i = 0; i < n_brt; ++i)
    cpu::kernels::process_radix_tree_i(i, n_brt, codes, prefix_n, has_leaf_left,
                                       has_leaf_right, left_child, parent);
}

void ref_edge_count(const uint8_t* prefix_n, const int32_t* parent, uint32_t n_brt,
                    int32_t* edge_count) {
  for (uint32_t i = 0; i < n_brt; ++i)
    cpu::kernels::process_edge_count_i(static_cast<int>(i), prefix_n, parent, edge_count);
}

// Stage 2 of the reference: four binning passes on a pool of n_threads
// workers (host_dispatchers.cpp:135-177).  The result ends in keys; alt is
// the ping-pong buffer.
void ref_sort(uint32_t* keys, uint32_t* alt, uint32_t n, int n_threads) {
  core::thread_pool pool(n_threads);
  std::barrier bar(n_threads);
  const int shifts[4] = {0, 8, 16, 24};
  for (int p = 0; p < 4; ++p) {
    const uint32_t* src = (p % 2 == 0) ? keys : alt;
    uint32_t* dst = (p % 2 == 0) ? alt : keys;
    cpu::kernels::dispatch_binning_pass(pool, n_threads, bar, static_cast<int>(n), src, dst,
                                        shifts[p])
        .wait();
  }
}

}  // extern "C"
