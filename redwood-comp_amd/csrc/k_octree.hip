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
//
// A pull reads its nodes from the tile's LDS window.  The few walks that leave
// the window (about one bottom in 500 on uniform points, at most a dozen per
// tile) would make their whole wave wait on global memory while the
// workgroup holds its LDS, so they are queued per tile instead and a second
// kernel walks them from global memory, merging its links into the rows with
// atomicMax / atomicOr (the same max and OR, in any order).
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

// A node's pull descriptor:
//   count > 0:  x = top slot (offset + count - 1), y = kLinkBit | octant of
//               the chain's top level (the reference's set_child arguments,
//               func_octree.hpp:96-105);
//   otherwise:  x = left_child, y = has_leaf_left | has_leaf_right << 1.
constexpr uint32_t kLinkBit = 0x80000000u;

__device__ __forceinline__ uint2 make_desc(int cnt, int off, uint32_t code, int pfx, int lc,
                                           bool leaf_l, bool leaf_r) {
  uint2 d;
  if (cnt > 0) {
    const int top_level = pfx / 3 - cnt + 1;
    d.x = static_cast<uint32_t>(off + cnt - 1);
    d.y = kLinkBit | ((code >> level_shift(top_level)) & 7u);
  } else {
    d.x = static_cast<uint32_t>(lc);
    d.y = (leaf_l ? 1u : 0u) | (leaf_r ? 2u : 0u);
  }
  return d;
}

// The descriptor of node gq from global memory: its seven fields are loaded
// together, so a walk waits for one round trip per node.
__device__ __forceinline__ uint2 desc_global(const OctArgs& a, uint32_t gq) {
  return make_desc(a.edge_count[gq], a.edge_offset[gq], a.codes[gq], a.prefix_n[gq],
                   a.left_child[gq], a.has_leaf_left[gq] != 0, a.has_leaf_right[gq] != 0);
}

// Steps a pull walk may take.  On a tree built from sorted unique codes a
// walk visits at most two levels of count-0 nodes (a handful of steps); the
// budget only ends walks over malformed trees handed in through the C ABI
// (a cycle of left_child links), which would otherwise never end.
constexpr int kPullBudget = 1024;

// Pulls every chain reached from the given roots through nodes of count 0
// (depth-first, one node per step, from global memory) into a register row:
// child slots keep the largest top, mask ORs the octants.  kCheck: nodes
// outside [0, n), a full stack or the step budget set RWH_STATUS_PULL in
// flags (a tree handed in through the C ABI); without it the walk trusts the
// tree (one built by the radix-tree kernels, whose bounds it provably keeps).
// steps counts the steps of all walks of one slot against kPullBudget (in
// and out).
template <bool kCheck>
__device__ __forceinline__ void pull_links(const OctArgs& a, uint32_t base, int n, int* stack, int sp,
                                           int32_t (&row)[8], int32_t& mask, uint32_t& flags, int& steps) {
  constexpr int kStack = 16;
  for (; sp > 0 && steps < kPullBudget; ++steps) {
    const int q = stack[--sp];
    if (kCheck && (q < 0 || q >= n)) {
      flags |= RWH_STATUS_PULL;
      continue;
    }
    const uint2 d = desc_global(a, base + static_cast<uint32_t>(q));
    if (d.y & kLinkBit) {
      const uint32_t which = d.y & 7u;
      const int32_t top = static_cast<int32_t>(d.x);
#pragma unroll
      for (uint32_t w = 0; w < 8; ++w) row[w] = (w == which) ? max(row[w], top) : row[w];
      mask |= static_cast<int32_t>(1u << which);
      continue;
    }
    const int lc = static_cast<int>(d.x);
    for (int side = 0; side < 2; ++side) {
      if ((d.y >> side) & 1u) continue;  // a leaf
      if (sp < kStack) stack[sp++] = lc + side;
      else flags |= RWH_STATUS_PULL;
    }
  }
  if (sp > 0) flags |= RWH_STATUS_PULL;
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

// ---------------------------------------------------------------------------
// The node kernel.  A workgroup owns kOctSpan consecutive radix nodes (a
// "tile"; loader thread t reads the four nodes starting at 4t) and first builds, in
// LDS, the pull descriptor (make_desc) of every node of a window reaching
// kOctHalo nodes beyond the tile on either side.  All descendants of a node
// lie inside its key range, so the pull of a bottom whose range fits the
// window never leaves LDS; walks that do are queued for k_oct_far (or read
// the same descriptors from global memory).  The node arrays are read with
// 16-byte (int) and 4-byte (byte arrays) loads, four nodes per thread.
// ---------------------------------------------------------------------------
constexpr int kOctBlock = 512;                // threads per workgroup
constexpr int kOctLoaders = 256;              // threads that read four nodes each
constexpr int kSpan = static_cast<int>(kOctSpan);
static_assert(kOctLoaders * 4 == kSpan, "loader threads read four nodes each");
constexpr int kOctHalo = 128;            // window margin on each side (multiple of 4)
constexpr int kOctWin = kSpan + 2 * kOctHalo;

// Four consecutive radix nodes q0 .. q0+3 of a segment (q0 a multiple of 4);
// nodes at or beyond n read as count 0.
struct Nodes4 {
  int cnt[4], off[4], lc[4];
  uint32_t code[4];
  uint32_t pfx, leaf_l, leaf_r;  // one byte per node
  __device__ int prefix(int e) const { return static_cast<int>((pfx >> (8 * e)) & 0xffu); }
  __device__ bool ll(int e) const { return ((leaf_l >> (8 * e)) & 0xffu) != 0; }
  __device__ bool lr(int e) const { return ((leaf_r >> (8 * e)) & 0xffu) != 0; }
  __device__ uint2 desc(int e) const {
    return make_desc(cnt[e], off[e], code[e], prefix(e), lc[e], ll(e), lr(e));
  }
};

__device__ __forceinline__ Nodes4 load_nodes4(const OctArgs& a, uint32_t base, int q0, int n) {
  Nodes4 g;
  const uint32_t gq = base + static_cast<uint32_t>(q0);
  if (q0 + 4 <= n) {
    const int4 c = *reinterpret_cast<const int4*>(a.edge_count + gq);
    const int4 o = *reinterpret_cast<const int4*>(a.edge_offset + gq);
    const int4 l = *reinterpret_cast<const int4*>(a.left_child + gq);
    const uint4 k = *reinterpret_cast<const uint4*>(a.codes + gq);
    g.cnt[0] = c.x; g.cnt[1] = c.y; g.cnt[2] = c.z; g.cnt[3] = c.w;
    g.off[0] = o.x; g.off[1] = o.y; g.off[2] = o.z; g.off[3] = o.w;
    g.lc[0] = l.x; g.lc[1] = l.y; g.lc[2] = l.z; g.lc[3] = l.w;
    g.code[0] = k.x; g.code[1] = k.y; g.code[2] = k.z; g.code[3] = k.w;
    g.pfx = *reinterpret_cast<const uint32_t*>(a.prefix_n + gq);
    g.leaf_l = *reinterpret_cast<const uint32_t*>(a.has_leaf_left + gq);
    g.leaf_r = *reinterpret_cast<const uint32_t*>(a.has_leaf_right + gq);
  } else {
    g.pfx = g.leaf_l = g.leaf_r = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const bool in = q0 + e < n;
      g.cnt[e] = in ? a.edge_count[gq + e] : 0;
      g.off[e] = in ? a.edge_offset[gq + e] : 0;
      g.lc[e] = in ? a.left_child[gq + e] : 0;
      g.code[e] = in ? a.codes[gq + e] : 0u;
      if (in) {
        g.pfx |= static_cast<uint32_t>(a.prefix_n[gq + e]) << (8 * e);
        g.leaf_l |= static_cast<uint32_t>(a.has_leaf_left[gq + e]) << (8 * e);
        g.leaf_r |= static_cast<uint32_t>(a.has_leaf_right[gq + e]) << (8 * e);
      }
    }
  }
  return g;
}

struct Window {
  const uint2* desc;  // LDS, nodes [w0, w1)
  int w0, w1;
};

// ---- rare paths ----------------------------------------------------------------


// Merges a register row into this thread's LDS row (entry w at row[w *
// kOctBlock]); returns the mask.
__device__ __forceinline__ int32_t merge_row(const int32_t (&r)[8], int32_t m, int32_t* row) {
#pragma unroll
  for (int w = 0; w < 8; ++w)
    if ((m >> w) & 1) atomicMax(&row[w * kOctBlock], r[w]);
  return m;
}

// The whole pull of a chain bottom from global memory, for the rare walk
// that did not fit the LDS pull stack or the tile's queue of far walks (links
// merged twice are harmless: max and OR).  Kept out of pull_pair's loop: a
// global walk inside it, even one never taken, slows the common LDS walk.
template <bool kCheck>
__device__ __forceinline__ int32_t pull_spill(const OctArgs& a, uint32_t base, int n, uint32_t pair, int32_t* row,
                                           uint32_t& flags, int& steps) {
  int32_t r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int32_t m = 0;
  int stack[16];
  int sp = 0;
  const int c = static_cast<int>(pair & 0x3fffffffu);
  for (int side = 0; side < 2; ++side)
    if (!((pair >> (30 + side)) & 1u) && c + side < n) stack[sp++] = c + side;
  pull_links<kCheck>(a, base, n, stack, sp, r, m, flags, steps);
  return merge_row(r, m, row);
}

// Node 0's links (slot 0 is also offset[0]): every chain reached from the
// children of the complete-range root through nodes of count 0.
template <bool kCheck>
__device__ __forceinline__ int32_t pull_root(const OctArgs& a, uint32_t base, int n, int32_t* row,
                                             uint32_t& flags) {
  int32_t r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int32_t m = 0;
  const int g = full_root_split(a.codes + base, n);
  int stack[16];
  int sp = 0;
  if (g != 0) stack[sp++] = g;
  if (g + 1 < n) stack[sp++] = g + 1;
  int steps = 0;
  pull_links<kCheck>(a, base, n, stack, sp, r, m, flags, steps);
  return merge_row(r, m, row);
}

// ---- the pull ------------------------------------------------------------------
// Pending sibling pairs of a pull, in registers: a push shifts every entry
// down and puts the new pair on top, a pop shifts them up (no indexed
// access, so the stack stays in four VGPRs and costs no LDS).
constexpr int kPullStack = 4;
struct PullStack {
  uint32_t e0 = 0, e1 = 0, e2 = 0, e3 = 0;  // e0 on top
  int sp = 0;
  __device__ __forceinline__ bool full() const { return sp >= kPullStack; }
  __device__ __forceinline__ bool empty() const { return sp == 0; }
  __device__ __forceinline__ void push(uint32_t v) {
    e3 = e2;
    e2 = e1;
    e1 = e0;
    e0 = v;
    ++sp;
  }
  __device__ __forceinline__ uint32_t pop() {
    const uint32_t v = e0;
    e0 = e1;
    e1 = e2;
    e2 = e3;
    --sp;
    return v;
  }
};

// The tile's queue of far walks: the number taken (LDS) and the tile's
// kOctFarCap items in the workspace (nullptr: none, every walk runs in
// place).  An item is (bottom slot relative to the segment's octree base,
// node relative to the segment); unused items hold kNoFarItem.
struct FarQueue {
  uint32_t* count;  // LDS
  uint2* item;
};
constexpr uint32_t kNoFarItem = 0xffffffffu;

// Pull of a chain bottom whose radix node has the children pair `pair` (left
// child | has_leaf_left << 30 | has_leaf_right << 31): depth-first through
// nodes of count 0, a pair of siblings per step.  Every link is an LDS max
// into this thread's row; returns the mask of octants linked.  Below a node
// of count > 0 at most two nodes of count 0 follow each other on a path
// (prefix lengths increase strictly and a count-0 node has its parent's
// level), so at most three pairs wait on the stack; a deeper walk is still
// handled (pull_spill).  A node outside the window is queued with the
// bottom's slot (slot_rel, relative to the segment's octree base) while the
// tile's queue has room; otherwise the whole walk is redone by pull_spill.
template <bool kCheck>
__device__ __forceinline__ int32_t pull_pair(const OctArgs& a, uint32_t base, const Window& w,
                                             int n, uint32_t pair, int32_t* row,
                                             uint32_t& flags, const FarQueue& far, uint32_t slot_rel) {
  const uint32_t pair0 = pair;
  int32_t m = 0;
  PullStack stk;
  int steps = 0;
  bool spill = false;
  for (; steps < kPullBudget; ++steps) {
    const int c = static_cast<int>(pair & 0x3fffffffu);
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int q = c + side;
      if (((pair >> (30 + side)) & 1u) || q >= n) continue;  // a leaf (or no such node)
      uint2 d;
      if (q >= w.w0 && q < w.w1) {
        d = w.desc[q - w.w0];
      } else {
        const uint32_t idx = far.item ? atomicAdd(far.count, 1u) : kOctFarCap;
        if (idx < kOctFarCap) far.item[idx] = make_uint2(slot_rel, static_cast<uint32_t>(q));
        else spill = true;
        continue;
      }
      if (d.y & kLinkBit) {
        atomicMax(&row[(d.y & 7u) * kOctBlock], static_cast<int32_t>(d.x));
        m |= static_cast<int32_t>(1u << (d.y & 7u));
      } else if (!stk.full()) {
        stk.push((d.x & 0x3fffffffu) | ((d.y & 3u) << 30));
      } else {
        spill = true;
      }
    }
    if (stk.empty()) break;
    pair = stk.pop();
  }
  if (spill) m |= pull_spill<kCheck>(a, base, n, pair0, row, flags, steps);
  if (steps >= kPullBudget) flags |= RWH_STATUS_PULL;  // a malformed tree: the walk was cut short
  return m;
}

// Corners, cell sizes and rows.  A workgroup owns kOctSpan consecutive radix
// nodes (a "tile"; loader thread t reads the four starting at 4t).  It builds the
// window's descriptors in LDS, then walks the contiguous slot range its nodes
// own ([offset of its first node, offset + count of its last node)), one slot
// per thread, so every store is coalesced: the owner of slot s is the last
// node of the tile whose offset is <= s (offsets are an exclusive scan of the
// counts), found by a binary search in LDS.  Slot offset + j is the node j
// levels above the chain's deepest one; j = 0 is the bottom, whose row is
// pulled.  Node 0 writes the root cell when no node contributes (D8).
// kCheck: the tree comes through the C ABI and may be malformed: child
// indices are checked and walks that stop at their bounds are reported in
// a.status (RWH_STATUS_PULL).  The whole builds run the kernel without the
// checks on the tree their radix-tree kernels built.
// LDS (about 40 KiB) and registers (at most 64 VGPRs and 100 SGPRs, which
// waves_per_eu(8) asks of the compiler) fit four workgroups, 32 waves, per
// CU: the kernel waits on memory, and a fourth workgroup per CU took it from
// 244 to 227 us (uniform 2^24) and from 227 to 196 us (clustered).
template <bool kCheck>
__global__ __launch_bounds__(kOctBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_oct_nodes(
    OctArgs a, SegArgs seg, uint32_t n_tiles) {
  __shared__ uint2 s_desc[kOctWin];
  __shared__ __align__(16) int32_t s_off[kSpan];
  __shared__ __align__(16) uint32_t s_code[kSpan];
  __shared__ __align__(16) uint32_t s_lcf[kSpan];  // left_child | leaf_l << 30 | leaf_r << 31
  __shared__ uint32_t s_lvl[kSpan / 4];  // prefix_n / 3, one byte per node
  // each thread's pulled row, entry-major (entry e of thread t at
  // [e * kOctBlock + t]) so that a wave's accesses fall in distinct banks
  __shared__ int32_t s_row[8 * kOctBlock];
  __shared__ float s_cell[32];  // range / 2^k
  __shared__ int64_t s_end;
  __shared__ uint32_t s_nfar;  // far walks queued by this tile

  // neighbouring tiles share window margins: XCD-grouped order
  const uint32_t tile = xcd_tile();
  if (tile >= n_tiles) return;
  const uint32_t first = tile * kOctSpan;
  const Seg sg = seg_of(seg, first);
  const int b = static_cast<int>(first - sg.base);
  const int n = static_cast<int>(sg.count);  // n_brt
  uint2* far_items = a.far_item ? a.far_item + static_cast<size_t>(tile) * kOctFarCap : nullptr;
  // the tile's queue starts empty (k_oct_far skips unused items)
  if (far_items && threadIdx.x < kOctFarCap) far_items[threadIdx.x] = make_uint2(kNoFarItem, 0u);
  if (first < sg.base || b >= n) return;  // a segment's padding (uniform across the workgroup)
  const uint32_t tid = threadIdx.x;
  const OctSeg os = oct_seg(a, sg);
  const int64_t rows = os.rows;
  const uint32_t* codes = a.codes + sg.base;
  int4* rows8 = reinterpret_cast<int4*>(a.children + static_cast<size_t>(os.base) * 8);
  float4* corner = a.corner + os.base;
  float* cell = a.cell_size + os.base;
  int32_t* mask = a.child_node_mask + os.base;
  int32_t* leaf_mask = a.child_leaf_mask + os.base;
  const int nv = min(kSpan, n - b);  // valid nodes of the tile
  uint32_t flags = 0;  // RWH_STATUS_* bits of this thread's walks

  Window w;
  w.desc = s_desc;
  w.w0 = max(b - kOctHalo, 0);
  w.w1 = min(b + kSpan + kOctHalo, n);

  // 1. the tile's nodes (four per loader thread) and the window's
  // descriptors (the margins read by the other threads meanwhile)
  if (tid < 32) s_cell[tid] = __fdiv_rn(a.range, static_cast<float>(1u << tid));
  if (tid == 0) s_nfar = 0;
  constexpr int kHaloGroups = kOctHalo / 4;
  static_assert(kOctBlock - kOctLoaders >= 2 * kHaloGroups, "one thread per margin group");
  const int ht = static_cast<int>(tid) - kOctLoaders;
  if (ht >= 0 && ht < 2 * kHaloGroups) {
    const int h = (ht < kHaloGroups) ? b - kOctHalo + 4 * ht : b + kSpan + 4 * (ht - kHaloGroups);
    if (h >= w.w0 && h < w.w1) {
      const Nodes4 hg = load_nodes4(a, sg.base, h, n);
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (h + e < w.w1) s_desc[h + e - w.w0] = hg.desc(e);
    }
  }
  const int q0 = b + 4 * static_cast<int>(tid);
  // offsets of the tile's nodes for the slot owners' search; past the last
  // valid node they read as INT32_MAX, above every slot
  constexpr int32_t kPastEnd = 0x7fffffff;
  if (tid < kOctLoaders && q0 >= n)
    reinterpret_cast<int4*>(s_off)[tid] = make_int4(kPastEnd, kPastEnd, kPastEnd, kPastEnd);
  if (tid < kOctLoaders && q0 < n) {
    const Nodes4 g = load_nodes4(a, sg.base, q0, n);
    uint32_t lcf[4], lv = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (q0 + e < n) {
        s_desc[q0 + e - w.w0] = g.desc(e);
        const int lc = g.lc[e];
        // a child index outside the tree: the pull skips it, reported here
        if (kCheck && ((!g.ll(e) && static_cast<uint32_t>(lc) >= static_cast<uint32_t>(n)) ||
                       (!g.lr(e) && static_cast<uint32_t>(lc + 1) >= static_cast<uint32_t>(n))))
          flags |= RWH_STATUS_PULL;
      }
      lcf[e] = (static_cast<uint32_t>(g.lc[e]) & 0x3fffffffu) | (g.ll(e) ? 1u << 30 : 0u) |
               (g.lr(e) ? 1u << 31 : 0u);
      lv |= static_cast<uint32_t>(g.prefix(e) / 3) << (8 * e);
      // the last valid node closes the tile's slot range
      if (q0 + e == b + nv - 1) s_end = static_cast<int64_t>(g.off[e]) + max(g.cnt[e], 0);
    }
    reinterpret_cast<int4*>(s_off)[tid] = make_int4(g.off[0], q0 + 1 < n ? g.off[1] : kPastEnd,
                                                   q0 + 2 < n ? g.off[2] : kPastEnd, q0 + 3 < n ? g.off[3] : kPastEnd);
    reinterpret_cast<uint4*>(s_code)[tid] = make_uint4(g.code[0], g.code[1], g.code[2], g.code[3]);
    reinterpret_cast<uint4*>(s_lcf)[tid] = make_uint4(lcf[0], lcf[1], lcf[2], lcf[3]);
    s_lvl[tid] = lv;
  }
  const int root_level = a.prefix_n[sg.base] / 3;
  if (b == 0 && tid == 0) {
    const uint32_t cap = a.oct_cap ? a.oct_cap[sg.id] : a.single_cap;
    if (a.n_oct[sg.id] == 0 && cap > 0) {
      const uint32_t sh = level_shift(root_level);
      corner[0] = morton_decode((codes[0] >> sh) << sh, a.min_coord, a.range);
      cell[0] = a.range;
    }
  }
  __syncthreads();

  // 2. the slots, one per thread
  int32_t* row = s_row + tid;
  const FarQueue far{&s_nfar, far_items};
  // slots below rows, at most the octree capacity (< 2^31): 32-bit indices
  const int lo = max(s_off[0], 0);
  const int hi = static_cast<int>(min<int64_t>(s_end, rows));
  for (int slot = lo + static_cast<int>(tid); slot < hi; slot += kOctBlock) {
    int k = 0;  // last node with offset <= slot (the padding past nv is never taken)
#pragma unroll
    for (int step = kSpan / 2; step >= 1; step >>= 1) k = s_off[k + step] <= slot ? k + step : k;
    if (b + k == 0) continue;  // node 0 owns no chain (the reference starts at 1)
    const int j = static_cast<int>(slot - s_off[k]);
    const uint32_t ci = s_code[k];
    const int lvl = static_cast<int>((s_lvl[k >> 2] >> (8 * (k & 3))) & 0xffu);
    const int level = lvl - j;
    const uint32_t sh = level_shift(level);
    store_stream(corner + slot, morton_decode((ci >> sh) << sh, a.min_coord, a.range));
    store_stream(cell + slot, s_cell[static_cast<uint32_t>(level - root_level) & 31u]);
    int4 r0, r1;
    int32_t m;
    if (j > 0) {  // the child below, in the same chain
      const uint32_t which = (ci >> level_shift(lvl - (j - 1))) & 7u;
      const int32_t c = static_cast<int32_t>(slot - 1);
      r0 = make_int4(which == 0 ? c : 0, which == 1 ? c : 0, which == 2 ? c : 0, which == 3 ? c : 0);
      r1 = make_int4(which == 4 ? c : 0, which == 5 ? c : 0, which == 6 ? c : 0, which == 7 ? c : 0);
      m = static_cast<int32_t>(1u << which);
    } else {  // the bottom: pull the chains that link here
#pragma unroll
      for (int e = 0; e < 8; ++e) row[e * kOctBlock] = 0;
      m = pull_pair<kCheck>(a, sg.base, w, n, s_lcf[k], row, flags, far, static_cast<uint32_t>(slot));
      if (slot == 0) m |= pull_root<kCheck>(a, sg.base, n, row, flags);
      r0 = make_int4(row[0], row[kOctBlock], row[2 * kOctBlock], row[3 * kOctBlock]);
      r1 = make_int4(row[4 * kOctBlock], row[5 * kOctBlock], row[6 * kOctBlock], row[7 * kOctBlock]);
    }
    // final outputs, read by no later kernel of the build: nontemporal stores
    store_stream(rows8 + slot * 2, r0);
    store_stream(rows8 + slot * 2 + 1, r1);
    store_stream(mask + slot, m);
    store_stream(leaf_mask + slot, 0);
  }
  if (kCheck && flags && a.status) atomicOr(&a.status[sg.id], flags);
}

// The queued far walks: thread k of a tile's group of kOctFarCap walks the
// tile's item k (if queued) from global memory and merges its links into the
// bottom's row: atomicMax per child slot and atomicOr into the node mask,
// over the row k_oct_nodes wrote.  The walk goes level by level, the loads of
// a level's nodes issued together: below a node of count > 0 at most two
// nodes of count 0 follow each other on a path, so on a tree built from
// sorted unique codes three levels (one round trip each) reach every link.
// What a malformed tree leaves after them (more levels, a wider frontier)
// continues depth-first with pull_links' bounds and checks.
constexpr int kFarBlock = 256;
static_assert(kFarBlock % kOctFarCap == 0, "a workgroup covers whole tiles");
constexpr int kFarLevels = 3;
constexpr int kFarWidth = 8;  // frontier nodes per level

template <bool kCheck>
__global__ __launch_bounds__(kFarBlock) void k_oct_far(OctArgs a, SegArgs seg, uint32_t n_tiles) {
  const uint32_t gid = blockIdx.x * kFarBlock + threadIdx.x;
  const uint32_t tile = gid / kOctFarCap;
  if (tile >= n_tiles) return;
  const uint2 it = a.far_item[gid];
  if (it.x == kNoFarItem) return;
  const Seg sg = seg_of(seg, tile * kOctSpan);
  const OctSeg os = oct_seg(a, sg);
  const int n = static_cast<int>(sg.count);
  int32_t r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int32_t m = 0;
  uint32_t flags = 0;
  int stack[16];  // the depth-first rest
  int sp = 0;
  int fr[kFarWidth];
  int nf = 1;
  fr[0] = static_cast<int>(it.y);
  int steps = 0;
  for (int level = 0; level < kFarLevels && nf > 0; ++level) {
    uint2 d[kFarWidth];
#pragma unroll
    for (int i = 0; i < kFarWidth; ++i) {
      const bool in = i < nf && static_cast<uint32_t>(fr[i]) < static_cast<uint32_t>(n);
      if (kCheck && i < nf && !in) flags |= RWH_STATUS_PULL;
      d[i] = in ? desc_global(a, sg.base + static_cast<uint32_t>(fr[i])) : make_uint2(0u, 3u);
    }
    steps += nf;
    int nx[kFarWidth];
    int nn = 0;
#pragma unroll
    for (int i = 0; i < kFarWidth; ++i) {
      if (i >= nf) continue;
      if (d[i].y & kLinkBit) {
        const uint32_t which = d[i].y & 7u;
        const int32_t top = static_cast<int32_t>(d[i].x);
#pragma unroll
        for (uint32_t w = 0; w < 8; ++w) r[w] = (w == which) ? max(r[w], top) : r[w];
        m |= static_cast<int32_t>(1u << which);
        continue;
      }
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        if ((d[i].y >> side) & 1u) continue;  // a leaf (or a node skipped above)
        const int c = static_cast<int>(d[i].x) + side;
        if (nn < kFarWidth) nx[nn++] = c;
        else if (sp < 16) stack[sp++] = c;
        else flags |= RWH_STATUS_PULL;
      }
    }
#pragma unroll
    for (int i = 0; i < kFarWidth; ++i) fr[i] = nx[i];
    nf = nn;
  }
  for (int i = 0; i < nf && sp < 16; ++i) stack[sp++] = fr[i];
  if (sp > 0) pull_links<kCheck>(a, sg.base, n, stack, sp, r, m, flags, steps);
  const size_t slot = static_cast<size_t>(os.base) + it.x;
#pragma unroll
  for (int w = 0; w < 8; ++w)
    if ((m >> w) & 1) atomicMax(&a.children[slot * 8 + w], r[w]);
  if (m) atomicOr(&a.child_node_mask[slot], m);
  if (kCheck && flags && a.status) atomicOr(&a.status[sg.id], flags);
}

void launch_octree(const OctArgs& a, SegArgs seg, uint32_t n_index, bool check, hipStream_t stream) {
  const uint32_t n_tiles = oct_tiles(n_index);
  if (n_tiles == 0) return;
  seg.count_sub = 1;
  const uint32_t far_grid = (n_tiles * kOctFarCap + kFarBlock - 1) / kFarBlock;
  if (check) {
    hipLaunchKernelGGL(k_oct_nodes<true>, dim3(xcd_grid(n_tiles)), dim3(kOctBlock), 0, stream, a, seg, n_tiles);
    if (a.far_item)
      hipLaunchKernelGGL(k_oct_far<true>, dim3(far_grid), dim3(kFarBlock), 0, stream, a, seg, n_tiles);
  } else {
    hipLaunchKernelGGL(k_oct_nodes<false>, dim3(xcd_grid(n_tiles)), dim3(kOctBlock), 0, stream, a, seg, n_tiles);
    if (a.far_item)
      hipLaunchKernelGGL(k_oct_far<false>, dim3(far_grid), dim3(kFarBlock), 0, stream, a, seg, n_tiles);
  }
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
  const int n = static_cast<int>(sg.count);
  const int lvl = a.prefix_n[sg.base + i] / 3 + 1;
  const uint32_t sh = level_shift(lvl);
  uint32_t flags = 0;
  for (int side = 0; side < 2; ++side) {
    const bool is_leaf = side == 0 ? a.has_leaf_left[sg.base + i] : a.has_leaf_right[sg.base + i];
    if (!is_leaf) continue;
    const int leaf = a.left_child[sg.base + i] + side;
    if (leaf < 0 || leaf > n) {  // keys are [0, n]: a malformed tree
      flags |= RWH_STATUS_LEAF;
      continue;
    }
    const uint32_t which = (a.codes[sg.base + leaf] >> sh) & 7u;
    // the walk stops after 31 steps (func_octree.hpp:83-92, the left walk of
    // 07_octree.cu:135-146; the right one there is unbounded and would not
    // end at node 0, whose parent is itself)
    int rt = i;
    int steps = 0;
    bool inside = true;
    while (counts[rt] == 0) {
      rt = parent[rt];
      if (rt < 0 || rt >= n) {
        inside = false;
        break;
      }
      if (++steps > 30) break;
    }
    if (!inside) {
      flags |= RWH_STATUS_LEAF;
      continue;
    }
    const int64_t bottom = a.edge_offset[sg.base + rt];
    if (bottom < 0 || bottom >= static_cast<int64_t>(os.rows)) continue;
    atomicMax(&children[bottom * 8 + which], leaf);
    atomicAnd(&a.child_leaf_mask[os.base + bottom], ~static_cast<int32_t>(1u << which));
  }
  if (flags && a.status) atomicOr(&a.status[sg.id], flags);
}

void launch_link_leaf(const OctArgs& a, SegArgs seg, uint32_t n_index, hipStream_t stream) {
  const uint32_t grid = (n_index + kElemBlock - 1) / kElemBlock;
  if (grid == 0) return;
  seg.count_sub = 1;
  hipLaunchKernelGGL(k_link_leaf, dim3(grid), dim3(kElemBlock), 0, stream, a, seg);
}

}  // namespace rw
