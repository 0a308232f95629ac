// Stage 7 of the octree build: octree nodes from the radix tree, edge counts
// and offsets (reference: apps/tree/host/func_octree.hpp:28-112,
// apps/tree/cuda/07_octree.cu:30-114, 175-226), plus the opt-in leaf linking
// the reference defines but never runs (func_octree.hpp:114-163,
// 07_octree.cu:116-173, 228-258).
//
// Brt node i >= 1 owns the oct slots [offset[i], offset[i] + count[i]): slot
// offset[i] + j is the node j levels above the deepest one.  Every slot but
// the chain's first ("bottom") receives exactly one child, from the slot below
// it in the same chain, so its whole row (8 children, node mask, leaf mask)
// is written in one go.  Bottom slots receive children from other chains (the
// top of every chain links to the bottom slot of its nearest contributing
// ancestor, a chain start by construction of the exclusive offsets), so they
// are cleared by their owner in the first kernel and filled with atomic max /
// atomic OR in the second: the result is the same as the reference's
// |= / = updates resolved as OR and max (D6), independent of scheduling.
// Slots outside [0, min(n_oct, capacity)) are never written.
#include "rw_device.hpp"
#include "rw_internal.hpp"

namespace rw {

constexpr int kMortonBits = 30;

struct OctSeg {
  uint32_t base, rows;
};

__device__ __forceinline__ OctSeg oct_seg(const OctArgs& a, const Seg& sg) {
  OctSeg o;
  o.base = a.oct_base ? a.oct_base[sg.id] : 0u;
  const uint32_t cap = a.oct_cap ? a.oct_cap[sg.id] : a.single_cap;
  const uint32_t n = a.n_oct[sg.id];
  o.rows = n < cap ? n : cap;
  return o;
}

__device__ __forceinline__ uint32_t level_shift(int level) {
  return static_cast<uint32_t>(kMortonBits - 3 * level) & 31u;
}

__device__ __forceinline__ float cell_of(float range, int level_diff) {
  return __fdiv_rn(range, static_cast<float>(1u << (static_cast<uint32_t>(level_diff) & 31u)));
}

// Kernel A: corners, cell sizes, chain-internal rows, cleared bottom rows, and
// the root cell when no node contributes (D8).
__global__ __launch_bounds__(kElemBlock) void k_oct_nodes(OctArgs a, SegArgs seg) {
  const uint32_t gi = blockIdx.x * kElemBlock + threadIdx.x;
  const Seg sg = seg_of(seg, blockIdx.x * kElemBlock);
  const int i = static_cast<int>(gi - sg.base);
  if (gi < sg.base || i >= static_cast<int>(sg.count)) return;
  const OctSeg os = oct_seg(a, sg);
  const uint32_t* codes = a.codes + sg.base;
  const uint8_t* prefix_n = a.prefix_n + sg.base;
  const int root_level = prefix_n[0] / 3;
  int4* rows8 = reinterpret_cast<int4*>(a.children + static_cast<size_t>(os.base) * 8);
  float4* corner = a.corner + os.base;
  float* cell = a.cell_size + os.base;
  int32_t* mask = a.child_node_mask + os.base;
  int32_t* leaf_mask = a.child_leaf_mask + os.base;

  if (i == 0) {
    const uint32_t cap = a.oct_cap ? a.oct_cap[sg.id] : a.single_cap;
    if (a.n_oct[sg.id] == 0 && cap > 0) {
      const uint32_t sh = level_shift(root_level);
      corner[0] = morton_decode((codes[0] >> sh) << sh, a.min_coord, a.range);
      cell[0] = a.range;
    }
    return;
  }
  const int count = a.edge_count[sg.base + i];
  if (count <= 0) return;
  const int64_t off = a.edge_offset[sg.base + i];
  const uint32_t ci = codes[i];
  const int lvl = prefix_n[i] / 3;
  const int64_t rows = os.rows;
  for (int j = 0; j < count; ++j) {
    const int64_t slot = off + j;
    const int level = lvl - j;
    const uint32_t sh = level_shift(level);
    const uint32_t node_prefix = ci >> sh;
    if (slot >= 0 && slot < rows) {
      corner[slot] = morton_decode(node_prefix << sh, a.min_coord, a.range);
      cell[slot] = cell_of(a.range, level - root_level);
    }
    // row of this slot: the bottom is cleared; others hold the child below
    if (slot >= 0 && slot < rows) {
      // which = 8 marks "no child" (the bottom slot)
      uint32_t which = 8u;
      if (j > 0) which = (ci >> level_shift(lvl - (j - 1))) & 7u;
      const int v = static_cast<int>(slot - 1);
      int r[8];
#pragma unroll
      for (uint32_t c = 0; c < 8; ++c) r[c] = (which == c) ? v : 0;
      rows8[slot * 2] = make_int4(r[0], r[1], r[2], r[3]);
      rows8[slot * 2 + 1] = make_int4(r[4], r[5], r[6], r[7]);
      mask[slot] = (which < 8u) ? static_cast<int32_t>(1u << which) : 0;
      leaf_mask[slot] = 0;
    }
  }
}

// Kernel B: the top of each chain links to its nearest contributing ancestor
// (walk capped at 31 steps, Q5).
__global__ __launch_bounds__(kElemBlock) void k_oct_links(OctArgs a, SegArgs seg) {
  const uint32_t gi = blockIdx.x * kElemBlock + threadIdx.x;
  const Seg sg = seg_of(seg, blockIdx.x * kElemBlock);
  const int i = static_cast<int>(gi - sg.base);
  if (gi < sg.base || i < 1 || i >= static_cast<int>(sg.count)) return;
  const int32_t* counts = a.edge_count + sg.base;
  const int count = counts[i];
  if (count <= 0) return;
  const OctSeg os = oct_seg(a, sg);
  const int32_t* parent = a.parent + sg.base;
  int rt = parent[i];
  int steps = 0;
  while (counts[rt] == 0) {
    rt = parent[rt];
    if (++steps > 30) break;
  }
  const int64_t oct_parent = a.edge_offset[sg.base + rt];
  const int64_t top = static_cast<int64_t>(a.edge_offset[sg.base + i]) + count - 1;
  const int top_level = a.prefix_n[sg.base + i] / 3 - count + 1;
  const uint32_t which = (a.codes[sg.base + i] >> level_shift(top_level)) & 7u;
  if (oct_parent < 0 || oct_parent >= static_cast<int64_t>(os.rows)) return;
  int32_t* children = a.children + static_cast<size_t>(os.base) * 8;
  atomicMax(&children[oct_parent * 8 + which], static_cast<int32_t>(top));
  atomicOr(&a.child_node_mask[os.base + oct_parent], static_cast<int32_t>(1u << which));
}

void launch_octree(const OctArgs& a, SegArgs seg, uint32_t n_index, hipStream_t stream) {
  const uint32_t grid = (n_index + kElemBlock - 1) / kElemBlock;
  if (grid == 0) return;
  seg.count_sub = 1;
  hipLaunchKernelGGL(k_oct_nodes, dim3(grid), dim3(kElemBlock), 0, stream, a, seg);
  hipLaunchKernelGGL(k_oct_links, dim3(grid), dim3(kElemBlock), 0, stream, a, seg);
}

// Leaf linking (opt-in): each leaf child of brt node i goes into the bottom
// slot of the nearest contributing node at or above i.  Child slot = max;
// the leaf mask bit is cleared (func_octree.hpp:24, kept as written).
__global__ __launch_bounds__(kElemBlock) void k_link_leaf(OctArgs a, SegArgs seg) {
  const uint32_t gi = blockIdx.x * kElemBlock + threadIdx.x;
  const Seg sg = seg_of(seg, blockIdx.x * kElemBlock);
  const int i = static_cast<int>(gi - sg.base);
  if (gi < sg.base || i >= static_cast<int>(sg.count)) return;
  const OctSeg os = oct_seg(a, sg);
  const int32_t* counts = a.edge_count + sg.base;
  const int32_t* parent = a.parent + sg.base;
  int32_t* children = a.children + static_cast<size_t>(os.base) * 8;
  const int lvl = a.prefix_n[sg.base + i] / 3 + 1;
  const uint32_t sh = level_shift(lvl);
  for (int side = 0; side < 2; ++side) {
    const bool is_leaf = side == 0 ? a.has_leaf_left[sg.base + i] : a.has_leaf_right[sg.base + i];
    if (!is_leaf) continue;
    const int leaf = a.left_child[sg.base + i] + side;
    const uint32_t which = (a.codes[sg.base + leaf] >> sh) & 7u;
    int rt = i;
    int steps = 0;
    while (counts[rt] == 0) {
      rt = parent[rt];
      if (++steps > 30) break;
    }
    const int64_t bottom = a.edge_offset[sg.base + rt];
    if (bottom < 0 || bottom >= static_cast<int64_t>(os.rows)) continue;
    atomicMax(&children[bottom * 8 + which], leaf);
    atomicAnd(&a.child_leaf_mask[os.base + bottom], ~static_cast<int32_t>(1u << which));
  }
}

void launch_link_leaf(const OctArgs& a, SegArgs seg, uint32_t n_index, hipStream_t stream) {
  const uint32_t grid = (n_index + kElemBlock - 1) / kElemBlock;
  if (grid == 0) return;
  seg.count_sub = 1;
  hipLaunchKernelGGL(k_link_leaf, dim3(grid), dim3(kElemBlock), 0, stream, a, seg);
}

}  // namespace rw
