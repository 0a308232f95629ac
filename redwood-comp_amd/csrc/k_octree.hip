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
// is written by the owner in one go.
//
// Bottom slots receive the tops of other chains: the reference pushes, from
// every chain c, a link into offset[T(c)], where T(c) is found by walking up
// parent[] while edge_count is 0 (at most 31 steps; Q5).  Here the owner of a
// bottom pulls those links instead, so no slot has two writers and no atomic
// is needed:
//   * parent[] is the complete Karras tree (node 0 spans [0, n_brt]), so the
//     nodes whose parent is p >= 1 are p's internal children (left_child[p]
//     unless has_leaf_left[p], left_child[p] + 1 unless has_leaf_right[p]),
//     and those of node 0 are the internal children of the complete-range
//     root;
//   * a node of edge count 0 has the same octree level as its parent and a
//     longer prefix, so at most two of them follow each other on a path: the
//     31-step cap can only bite at node 0 (count 0, its own parent), where
//     the walk stays.  T(c) is thus the nearest ancestor with a non-zero count,
//     or node 0;
//   * the owner p pulls every chain reached from its children through nodes of
//     count 0, and the owner of slot 0 (the first chain, offset 0 = offset[0])
//     also pulls for node 0.
// Child slots keep the largest top and the mask ORs all links (D6).  Slots
// outside [0, min(n_oct, capacity)) are never written.
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

// One incoming link: chain c's top slot at its top level's octant (the
// reference's set_child arguments, func_octree.hpp:96-105).
__device__ __forceinline__ void link_into(int cnt, int off, uint32_t code, int pfx, int32_t (&row)[8],
                                          int32_t& mask) {
  const int32_t top = off + cnt - 1;
  const int top_level = pfx / 3 - cnt + 1;
  const uint32_t which = (code >> level_shift(top_level)) & 7u;
#pragma unroll
  for (uint32_t w = 0; w < 8; ++w) row[w] = (w == which) ? max(row[w], top) : row[w];
  mask |= static_cast<int32_t>(1u << which);
}

__device__ __forceinline__ void add_link(const OctArgs& a, uint32_t base, int c, int32_t (&row)[8],
                                         int32_t& mask) {
  link_into(a.edge_count[base + c], a.edge_offset[base + c], a.codes[base + c],
            a.prefix_n[base + c], row, mask);
}

// Pulls every chain reached from the given roots through nodes of count 0
// (depth-first, one node per step).
__device__ __forceinline__ void pull_links(const OctArgs& a, uint32_t base, int* stack, int sp,
                                           int32_t (&row)[8], int32_t& mask) {
  constexpr int kStack = 16;
  while (sp > 0) {
    const int q = stack[--sp];
    if (a.edge_count[base + q] > 0) {
      add_link(a, base, q, row, mask);
      continue;
    }
    const int lc = a.left_child[base + q];
    if (!a.has_leaf_left[base + q] && sp < kStack) stack[sp++] = lc;
    if (!a.has_leaf_right[base + q] && sp < kStack) stack[sp++] = lc + 1;
  }
}

// What the pull needs of one radix node, loaded in one round trip.
struct PullNode {
  int cnt, off, lc;
  uint32_t code;
  int pfx;
  bool leaf_l, leaf_r;
};

template <bool kWithChildren>
__device__ __forceinline__ PullNode load_pull(const OctArgs& a, uint32_t base, int q, bool live) {
  PullNode n{0, 0, 0, 0u, 0, true, true};
  if (live) {
    n.cnt = a.edge_count[base + q];
    n.off = a.edge_offset[base + q];
    n.code = a.codes[base + q];
    n.pfx = a.prefix_n[base + q];
    if (kWithChildren) {
      n.lc = a.left_child[base + q];
      n.leaf_l = a.has_leaf_left[base + q] != 0;
      n.leaf_r = a.has_leaf_right[base + q] != 0;
    }
  }
  return n;
}

// The pull of a bottom slot owned by brt node p >= 1, breadth-first: the
// nodes of one level are loaded together, so the walk costs one memory round
// trip per level instead of one per node.  Below p (count > 0) at most two
// nodes of count 0 follow each other on a path (prefix lengths increase
// strictly and a count-0 node has its parent's level), so every internal node
// three levels down has a non-zero count; any that does not is still handled,
// by the depth-first walk.
__device__ __forceinline__ void pull_below(const OctArgs& a, uint32_t base, int lc, bool leaf_l,
                                           bool leaf_r, int32_t (&row)[8], int32_t& mask) {
  PullNode n1[2];
  n1[0] = load_pull<true>(a, base, lc, !leaf_l);
  n1[1] = load_pull<true>(a, base, lc + 1, !leaf_r);
  int q2[4];
  bool v2[4];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const bool live = (k == 0) ? !leaf_l : !leaf_r;
    if (live && n1[k].cnt > 0) link_into(n1[k].cnt, n1[k].off, n1[k].code, n1[k].pfx, row, mask);
    const bool down = live && n1[k].cnt <= 0;
    q2[2 * k] = n1[k].lc;
    v2[2 * k] = down && !n1[k].leaf_l;
    q2[2 * k + 1] = n1[k].lc + 1;
    v2[2 * k + 1] = down && !n1[k].leaf_r;
  }
  PullNode n2[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) n2[k] = load_pull<true>(a, base, q2[k], v2[k]);
  int q3[8];
  bool v3[8];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (v2[k] && n2[k].cnt > 0) link_into(n2[k].cnt, n2[k].off, n2[k].code, n2[k].pfx, row, mask);
    const bool down = v2[k] && n2[k].cnt <= 0;
    q3[2 * k] = n2[k].lc;
    v3[2 * k] = down && !n2[k].leaf_l;
    q3[2 * k + 1] = n2[k].lc + 1;
    v3[2 * k + 1] = down && !n2[k].leaf_r;
  }
  PullNode n3[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) n3[k] = load_pull<false>(a, base, q3[k], v3[k]);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (!v3[k]) continue;
    if (n3[k].cnt > 0) {
      link_into(n3[k].cnt, n3[k].off, n3[k].code, n3[k].pfx, row, mask);
    } else {
      int stack[16];
      stack[0] = q3[k];
      pull_links(a, base, stack, 1, row, mask);
    }
  }
}

__device__ __forceinline__ int delta8o(uint32_t x, uint32_t y) {
  return static_cast<uint8_t>(__clz(x ^ y) - 1);
}

// Split of the complete range [0, n] (the root of the complete tree), by the
// reference's split search (func_brt.hpp:90-104) with l = n, d = +1.
__device__ int full_root_split(const uint32_t* codes, int n) {
  const uint32_t c0 = codes[0];
  const int dnode = delta8o(c0, codes[n]);
  int l2 = 0;  // log2_ceil(n)
  while ((1u << l2) < static_cast<uint32_t>(n)) ++l2;
  const uint32_t max_div = 1u << l2;
  int s = 0;
  uint32_t t = (static_cast<uint32_t>(n) + 1) / 2;
  for (uint32_t div = 2; div <= max_div; div <<= 1) {
    if (s + static_cast<int>(t) <= n - 1 && delta8o(c0, codes[s + static_cast<int>(t)]) > dnode)
      s += static_cast<int>(t);
    t = (static_cast<uint32_t>(n) + (div << 1) - 1) / (div << 1);
  }
  return s;
}

// Corners, cell sizes and every row of the chain; the bottom row pulls its
// links.  Node 0 writes the root cell when no node contributes (D8).
__global__ __launch_bounds__(kElemBlock) void k_oct_nodes(OctArgs a, SegArgs seg) {
  const uint32_t gi = blockIdx.x * kElemBlock + threadIdx.x;
  const Seg sg = seg_of(seg, blockIdx.x * kElemBlock);
  const int i = static_cast<int>(gi - sg.base);
  const int n = static_cast<int>(sg.count);  // n_brt
  if (gi < sg.base || i >= n) return;
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
  const int64_t off = a.edge_offset[sg.base + i];
  const uint32_t ci = codes[i];
  const int lvl = prefix_n[i] / 3;
  // the bottom's own children, loaded with the rest of the node
  const int lc = a.left_child[sg.base + i];
  const bool leaf_l = a.has_leaf_left[sg.base + i] != 0;
  const bool leaf_r = a.has_leaf_right[sg.base + i] != 0;
  if (count <= 0) return;
  const int64_t rows = os.rows;
  for (int j = 0; j < count; ++j) {
    const int64_t slot = off + j;
    if (slot < 0 || slot >= rows) continue;
    const int level = lvl - j;
    const uint32_t sh = level_shift(level);
    corner[slot] = morton_decode((ci >> sh) << sh, a.min_coord, a.range);
    cell[slot] = cell_of(a.range, level - root_level);
    int32_t r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int32_t m = 0;
    if (j > 0) {  // the child below, in the same chain
      const uint32_t which = (ci >> level_shift(lvl - (j - 1))) & 7u;
#pragma unroll
      for (uint32_t w = 0; w < 8; ++w) r[w] = (w == which) ? static_cast<int32_t>(slot - 1) : 0;
      m = static_cast<int32_t>(1u << which);
    } else {  // the bottom: pull the chains that link here
      pull_below(a, sg.base, lc, leaf_l, leaf_r, r, m);
      if (off == 0) {  // slot 0 is also offset[0]: node 0's links land here
        const int g = full_root_split(codes, n);
        int stack[16];
        int sp = 0;
        if (g != 0) stack[sp++] = g;
        if (g + 1 < n) stack[sp++] = g + 1;
        pull_links(a, sg.base, stack, sp, r, m);
      }
    }
    rows8[slot * 2] = make_int4(r[0], r[1], r[2], r[3]);
    rows8[slot * 2 + 1] = make_int4(r[4], r[5], r[6], r[7]);
    mask[slot] = m;
    leaf_mask[slot] = 0;
  }
}

void launch_octree(const OctArgs& a, SegArgs seg, uint32_t n_index, hipStream_t stream) {
  const uint32_t grid = (n_index + kElemBlock - 1) / kElemBlock;
  if (grid == 0) return;
  seg.count_sub = 1;
  hipLaunchKernelGGL(k_oct_nodes, dim3(grid), dim3(kElemBlock), 0, stream, a, seg);
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
