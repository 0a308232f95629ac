/*
 * TEST INFRASTRUCTURE ONLY — the CPU oracle of the octree build.
 *
 * A C++20 pthread restatement of the reference CPU path of the tree app's
 * seven-stage octree build (apps/tree/host/host_dispatchers.cpp:17-336 and the
 * per-element kernels it calls), under the semantics written down as REF-SPEC
 * (quirks Q1-Q9 replicated, deviations D1-D9 defined).  Only tests/, the smoke
 * entry point and the benchmark's CPU-baseline leg may load this library; the
 * product never links it.
 *
 * Every function takes host pointers.  Outputs are fully written for the index
 * ranges documented per function.
 */
#ifndef REDWOOD_TREE_ORACLE_H
#define REDWOOD_TREE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Thread pool (the reference's core::thread_pool, redwood/host/thread_pool.hpp:42-162). */
void* oracle_pool_create(int n_threads);
void oracle_pool_destroy(void* pool);
int oracle_pool_threads(void* pool);

/* S1 — xyz_to_morton32 (apps/tree/host/func_morton.hpp:31-39) over n float4 points. */
void oracle_morton(void* pool, const float* xyzw, uint32_t n, float min_coord,
                   float range, uint32_t* codes);

/* S2 — 4-pass LSD 8-bit binning sort (apps/tree/host/func_sort.cpp:117-211,
 * pass order host_dispatchers.cpp:135-177).  in is left untouched; out gets the
 * ascending keys; tmp is scratch of n keys. */
void oracle_sort(void* pool, const uint32_t* in, uint32_t* out, uint32_t* tmp,
                 uint32_t n);

/* S3 — std::unique_copy (host_dispatchers.cpp:183-194).  Returns n_unique. */
uint32_t oracle_unique(const uint32_t* sorted, uint32_t n, uint32_t* out);

/* S4 — Karras radix tree, nodes i in [0, n_unique-1) (apps/tree/host/func_brt.hpp:45-138,
 * D2, D3, D7).  parent[] is zeroed over [0, max(n_unique,1)) first and every parent
 * slot ends as the largest writer index (0 if none). */
void oracle_radix_tree(void* pool, const uint32_t* codes, uint32_t n_unique,
                       uint8_t* prefix_n, uint8_t* has_leaf_left,
                       uint8_t* has_leaf_right, int32_t* left_child,
                       int32_t* parent);

/* S5 — edge count (apps/tree/host/func_edge.hpp:9-16, CUDA 05_edge_count.cu:26-28). */
void oracle_edge_count(void* pool, const uint8_t* prefix_n, const int32_t* parent,
                       uint32_t n_brt, int32_t* edge_count);

/* S6 — global exclusive scan (D4).  Returns the total = n_octree_nodes. */
int64_t oracle_prefix_sum(const int32_t* edge_count, uint32_t n_brt,
                          int32_t* edge_offset);

/* S7 — octree nodes for brt i in [1, n_brt) (apps/tree/host/func_octree.hpp:28-112,
 * D5, D6, D8).  Rows [0, n_oct) of children (8 per node), child_node_mask and
 * child_leaf_mask are zeroed first; corner/cell_size slot 0 gets the root init when
 * oct_cap > 0.  Writes to slots outside [0, n_oct) are dropped (returned count). */
int64_t oracle_octree(void* pool, const int32_t* edge_offset,
                      const int32_t* edge_count, const uint32_t* codes,
                      const uint8_t* prefix_n, const int32_t* parent,
                      uint32_t n_brt, uint32_t n_oct, uint32_t oct_cap,
                      float min_coord, float range, int32_t* children,
                      float* corner /* xyzw */, float* cell_size,
                      int32_t* child_node_mask, int32_t* child_leaf_mask);

/* Leaf linking — process_link_leaf (func_octree.hpp:114-163), never run by the
 * reference (Q6); opt-in here.  Uses the capped parent walk of the CUDA copy
 * (07_octree.cu:138-148) on both children.  Child slots resolve as max,
 * leaf-mask bits are cleared with atomic AND. */
int64_t oracle_link_leaf(void* pool, const int32_t* edge_offset,
                         const int32_t* edge_count, const uint32_t* codes,
                         const uint8_t* has_leaf_left, const uint8_t* has_leaf_right,
                         const uint8_t* prefix_n, const int32_t* parent,
                         const int32_t* left_child, uint32_t n_brt, uint32_t n_oct,
                         int32_t* children, int32_t* child_leaf_mask);

/* All seven stages, as the reference's end-to-end loop runs them
 * (apps/tree/pipe_main_baseline.cpp:23-31).  The caller owns every buffer:
 * codes/sorted/unique/tmp n keys; brt arrays and edge arrays n entries;
 * oct arrays oct_cap (children 8*oct_cap, corner 4*oct_cap).
 * counts_out = {n_unique, n_brt, n_octree_nodes}; stage_ms[7] gets per-stage
 * wall time.  Returns 0, or -1 when n_octree_nodes exceeds oct_cap. */
int oracle_build(void* pool, const float* xyzw, uint32_t n, float min_coord,
                 float range, uint32_t* codes, uint32_t* sorted, uint32_t* unique_keys,
                 uint32_t* tmp, uint8_t* prefix_n, uint8_t* has_leaf_left,
                 uint8_t* has_leaf_right, int32_t* left_child, int32_t* parent,
                 int32_t* edge_count, int32_t* edge_offset, uint32_t oct_cap,
                 int32_t* children, float* corner, float* cell_size,
                 int32_t* child_node_mask, int32_t* child_leaf_mask,
                 uint32_t* counts_out, double* stage_ms);

#ifdef __cplusplus
}
#endif

#endif
