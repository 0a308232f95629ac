// Stages 1, 3, 4, 5, 6 of the octree build: Morton encode, duplicate removal,
// Karras radix tree, edge count, exclusive prefix sum.  Every kernel works on
// segments (clouds) of one index space; a single build is one segment.
#include "rw_device.hpp"
#include "rw_internal.hpp"

namespace rw {

// ---------------------------------------------------------------------------
// S1  Morton encode (reference: apps/tree/cuda/01_morton.cu:8-19,
// apps/tree/host/func_morton.hpp:31-39).  A workgroup covers 4*256 points;
// every load is a whole float4 (16 bytes, nontemporal: each point is read
// once) and lanes read consecutive points.
// ---------------------------------------------------------------------------
constexpr uint32_t kMortonPerThread = 4;
constexpr uint32_t kMortonTile = kElemBlock * kMortonPerThread;  // 1024

__global__ __launch_bounds__(kElemBlock) void k_morton(const float4* __restrict__ pts,
                                                       uint32_t* __restrict__ codes, SegArgs seg,
                                                       Quant q, uint32_t* __restrict__ status) {
  const uint32_t first = blockIdx.x * kMortonTile;
  const Seg sg = seg_of(seg, first);
  const uint32_t lim = sg.base + sg.count;
  // a batched build's first kernel: the segment's first block clears its status
  if (status && threadIdx.x == 0 && first == sg.base) status[sg.id] = 0;
#pragma unroll
  for (uint32_t j = 0; j < kMortonPerThread; ++j) {
    const uint32_t i = first + j * kElemBlock + threadIdx.x;
    if (i < lim) {
      const float4 p = load_stream(pts + i);
      codes[i] = morton_encode(p.x, p.y, p.z, q);
    }
  }
}

void launch_morton(const float4* pts, uint32_t* codes, SegArgs seg, uint32_t n_index,
                   float min_coord, float range, uint32_t* status, hipStream_t stream) {
  const uint32_t grid = (n_index + kMortonTile - 1) / kMortonTile;
  if (grid == 0) return;
  hipLaunchKernelGGL(k_morton, dim3(grid), dim3(kElemBlock), 0, stream, pts, codes, seg,
                     make_quant(min_coord, range), status);
}

// dst[i] = value for i < count (counts of empty builds, state cleared by
// kernels rather than memset nodes).
__global__ __launch_bounds__(kElemBlock) void k_fill_u32(uint32_t* __restrict__ dst, uint32_t value,
                                                         uint32_t count) {
  const uint32_t i = blockIdx.x * kElemBlock + threadIdx.x;
  if (i < count) dst[i] = value;
}

void launch_fill_u32(uint32_t* dst, uint32_t value, uint32_t count, hipStream_t stream) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_fill_u32, dim3((count + kElemBlock - 1) / kElemBlock), dim3(kElemBlock), 0, stream,
                     dst, value, count);
}

// ---------------------------------------------------------------------------
// Tile-wide exclusive scan for BLOCK threads.
// ---------------------------------------------------------------------------
template <int BLOCK>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_w, uint32_t* total) {
  constexpr int W = BLOCK / kWave;
  const uint32_t t = threadIdx.x, w = t / kWave;
  const uint32_t incl = wave_incl_scan(v);
  if ((t % kWave) == kWave - 1) s_w[w] = incl;
  __syncthreads();
  uint32_t off = 0, sum = 0;
#pragma unroll
  for (int i = 0; i < W; ++i) {
    const uint32_t x = s_w[i];
    off += (static_cast<uint32_t>(i) < w) ? x : 0u;
    sum += x;
  }
  __syncthreads();
  *total = sum;
  return off + incl - v;
}

// Tile prefix over the segment's earlier tiles by decoupled look-back: thread 0
// publishes the tile's aggregate, wave 0 walks back 64 tiles per round trip,
// thread 0 publishes the inclusive prefix; the result is broadcast via LDS.
// All threads must call it.
__device__ __forceinline__ uint32_t tile_prefix(uint32_t tile_total, uint32_t tile,
                                                uint32_t tile_in_seg, uint32_t epoch,
                                                uint64_t* lookback, uint32_t* s_bcast) {
  if (threadIdx.x == 0)
    lb_store(&lookback[tile], lb_pack(epoch, tile_in_seg == 0 ? kIncl : kAgg, tile_total));
  if (threadIdx.x < kWave) {
    uint32_t excl = 0;
    if (tile_in_seg > 0) {
      excl = lb_exclusive_wave(lookback, tile, tile_in_seg, epoch);
      if (threadIdx.x == 0) lb_store(&lookback[tile], lb_pack(epoch, kIncl, excl + tile_total));
    }
    if (threadIdx.x == 0) *s_bcast = excl;
  }
  __syncthreads();
  return *s_bcast;
}

// ---------------------------------------------------------------------------
// S3  unique (reference: CPU std::unique_copy host_dispatchers.cpp:183-194;
// CUDA flag/scan/scatter cu_dispatcher.cu:597-627).  Single pass: head flags,
// tile scan, look-back, scatter; the segment's last tile writes n_unique.
// Layout: wave w of a tile owns kScanKpt*64 consecutive keys, read as uint4
// groups g (lane l holds keys 4*(g*64+l) .. +3 of the wave's range).
// ---------------------------------------------------------------------------
template <int BLOCK, int KPT>
__global__ __launch_bounds__(BLOCK) void k_unique(const uint32_t* __restrict__ in,
                                                       uint32_t* __restrict__ out, SegArgs seg,
                                                       uint32_t* __restrict__ n_unique,
                                                       uint64_t* __restrict__ lookback,
                                                       uint32_t* __restrict__ ctrl) {
  constexpr int G = KPT / 4;
  constexpr uint32_t TILE = BLOCK * KPT;
  __shared__ uint32_t s_w[BLOCK / kWave];
  __shared__ uint32_t s_bcast;
  __shared__ Grab s_grab;
  __shared__ uint32_t s_out[TILE];  // the tile's heads in output order
  const uint32_t tid = threadIdx.x;
  const uint32_t wave = tid / kWave, lane = tid % kWave;
  if (tid == 0) s_grab = grab_tiles(ctrl, gridDim.x);
  __syncthreads();
  const uint32_t epoch = s_grab.epoch;
  // One tile per workgroup, in grab order: a tile only ever waits on tiles
  // whose workgroups took their grab earlier and are therefore running, so
  // the look-back spin ends whatever the residency of the grid.
  {
    const uint32_t tile = s_grab.index;
    const uint32_t first = tile * TILE;
    const Seg sg = seg_of(seg, first);
    const uint32_t lim = sg.base + sg.count;
    if (first >= lim) {
      // an empty segment still reports n_unique = 0 from its first tile
      if (tid == 0 && first == sg.base) n_unique[sg.id] = 0;
      return;
    }
    const uint32_t tile_in_seg = (first - sg.base) / TILE;
    const uint32_t wbase = first + wave * (KPT * kWave);

    uint32_t k[G][4];
    uint32_t flags[G];  // 4 head bits per group
    uint32_t cnt[G];
    if (first + TILE <= lim) {  // full tile: plain 16-byte loads
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const uint4 q = *reinterpret_cast<const uint4*>(in + wbase + 4 * (g * kWave + lane));
        k[g][0] = q.x; k[g][1] = q.y; k[g][2] = q.z; k[g][3] = q.w;
      }
    } else {
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const uint32_t i0 = wbase + 4 * (g * kWave + lane);
#pragma unroll
        for (int c = 0; c < 4; ++c) k[g][c] = (i0 + c < lim) ? in[i0 + c] : 0u;
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const uint32_t i0 = wbase + 4 * (g * kWave + lane);
      // predecessor of the group's first key: previous lane's last key, or memory
      uint32_t prev = __shfl_up(k[g][3], 1, kWave);
      if (lane == 0 && i0 > sg.base && i0 < lim) prev = in[i0 - 1];
      uint32_t f = 0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const uint32_t i = i0 + c;
        const uint32_t pk = (c == 0) ? prev : k[g][c - 1];
        const bool head = (i < lim) && (i == sg.base || k[g][c] != pk);
        f |= (head ? 1u : 0u) << c;
      }
      flags[g] = f;
      cnt[g] = __popc(f);
    }
    // Order is (wave, group, lane, c): per group a wave scan over lanes.
    uint32_t wave_run = 0;
    uint32_t gexcl[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const uint32_t incl = wave_incl_scan(cnt[g]);
      gexcl[g] = wave_run + incl - cnt[g];
      wave_run += __shfl(incl, kWave - 1, kWave);
    }
    uint32_t tile_total;
    const uint32_t wexcl = block_excl_scan<BLOCK>(lane == 0 ? wave_run : 0u, s_w, &tile_total);
    const uint32_t wave_off = __shfl(wexcl, 0, kWave);
    // The heads staged in LDS in output order (the look-back's barrier
    // orders these stores before the reads), then written out by consecutive
    // threads: the scattered head stores of a wave would each touch a
    // quarter of a line.
#pragma unroll
    for (int g = 0; g < G; ++g) {
      uint32_t pos = wave_off + gexcl[g];
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if ((flags[g] >> c) & 1u) s_out[pos++] = k[g][c];
    }
    const uint32_t excl = tile_prefix(tile_total, tile, tile_in_seg, epoch, lookback, &s_bcast);
    uint32_t* __restrict__ dst = out + sg.base + excl;
    for (uint32_t i = tid; i < tile_total; i += BLOCK) dst[i] = s_out[i];
    // the tile holding the segment's last key reports the count
    if (tid == 0 && lim - first <= TILE) n_unique[sg.id] = excl + tile_total;
  }
}

void launch_unique(const uint32_t* sorted, uint32_t* unique_out, SegArgs seg, uint32_t n_index,
                   uint32_t* n_unique, uint64_t* lookback, uint32_t* ctrl, hipStream_t stream) {
  if (n_index == 0) return;
  if (seg.n_segs == 1) {  // one cloud: the larger tile (fewer grabs and look-back steps)
    const uint32_t n_tiles = (n_index + kScanBigTile - 1) / kScanBigTile;
    hipLaunchKernelGGL((k_unique<kScanBigBlock, kScanBigKpt>), dim3(n_tiles), dim3(kScanBigBlock), 0, stream,
                       sorted, unique_out, seg, n_unique, lookback, ctrl);
    return;
  }
  const uint32_t n_tiles = (n_index + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL((k_unique<kScanBlock, kScanKpt>), dim3(n_tiles), dim3(kScanBlock), 0, stream, sorted,
                     unique_out, seg, n_unique, lookback, ctrl);
}

// ---------------------------------------------------------------------------
// S4  Karras binary radix tree (reference: apps/tree/host/func_brt.hpp:45-138,
// apps/tree/cuda/04_radix_tree.cu:12-144) for nodes [0, n_brt).
// Quirks kept: delta = uint8(clz(a^b) - 1) (Q1), the root's range is
// [0, n_brt - 1] (Q2), clz(0) = 32 and log2_ceil(0) = 0 (D3).
// Parent slots: every internal node other than the root of the complete tree
// over [0, n_brt] has exactly one claimer among nodes i >= 1 (the Karras
// construction); the Q2 root's own claims lose to any such writer under the
// max-writer rule (D7) and otherwise write 0.  So nodes i >= 1 store plainly,
// node 0 stores nothing for its Q2 claims, and 0 goes to parent[0] and to the
// internal children of the complete-range root (the only slots no node i >= 1
// claims).  The result equals max-of-writers over zeroed slots without
// atomics or a clearing pass.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int delta8(uint32_t a, uint32_t b) {
  return static_cast<uint8_t>(__clz(a ^ b) - 1);  // __clz(0) == 32
}
__device__ __forceinline__ int log2_ceil_u32(uint32_t x) {
  if (x == 0) return 0;
  const int fl = 31 - __clz(x);
  return fl + (((1u << fl) < x) ? 1 : 0);
}

// ---- reference search (any 32-bit keys) ------------------------------------
// Range and split of node i exactly as func_brt.hpp:57-104 computes them,
// reading codes from global memory; used when a key has bit 31 set (uint8
// wrap of delta, so the searches are not plain boundary searches).
__device__ __forceinline__ int ref_split(const uint32_t* codes, int i, int n, int d, int l, int dnode,
                                         uint32_t ci) {
  int s = 0;
  const uint32_t max_div = 1u << log2_ceil_u32(static_cast<uint32_t>(l));
  const int scut = (d == -1) ? i - 1 : n - i - 1;
  uint32_t t = (static_cast<uint32_t>(l) + 1) / 2;
  for (uint32_t div = 2; div <= max_div; div <<= 1) {
    if (s + static_cast<int>(t) <= scut && delta8(ci, codes[i + (s + static_cast<int>(t)) * d]) > dnode)
      s += static_cast<int>(t);
    t = (static_cast<uint32_t>(l) + (div << 1) - 1) / (div << 1);
  }
  return i + s * d + min(d, 0);
}

struct RefNode {
  int dnode, gamma;
  bool ll, lr;
};

__device__ RefNode ref_node(const uint32_t* codes, int i, int n) {
  const uint32_t ci = codes[i];
  int d = 1, l = 0;
  if (i == 0) {
    l = n - 1;  // Q2
  } else {
    const int diff = delta8(ci, codes[i + 1]) - delta8(ci, codes[i - 1]);
    d = (diff > 0) - (diff < 0);
    const int dmin = delta8(ci, codes[i - d]);
    int lmax = 2;
    while (i + lmax * d >= 0 && i + lmax * d <= n && delta8(ci, codes[i + lmax * d]) > dmin) lmax *= 2;
    const int cutoff = (d == -1) ? i : n - i;
    for (int t = lmax / 2; t >= 1; t /= 2)
      if (l + t <= cutoff && delta8(ci, codes[i + (l + t) * d]) > dmin) l += t;
  }
  const int j = i + l * d;
  RefNode r;
  r.dnode = delta8(ci, codes[j]);
  r.gamma = ref_split(codes, i, n, d, l, r.dnode, ci);
  r.ll = min(i, j) == r.gamma;
  r.lr = max(i, j) == r.gamma + 1;
  return r;
}

// The reference's search over a whole segment.  Such a tree need not be a
// proper binary tree (a parent slot can have several claimers), so parent[]
// is the maximum claimer over zeroed slots (D7), built with atomics by the
// segment's first workgroup alone (ref_parents: zeroing, barrier, claims);
// the other outputs are written by each node's own thread (ref_nodes).
__device__ void ref_nodes(const TreeArgs& a, const uint32_t* codes, uint32_t base, int b_begin, int b_end,
                          int n) {
  for (int i = b_begin + static_cast<int>(threadIdx.x); i < b_end; i += static_cast<int>(blockDim.x)) {
    const RefNode r = ref_node(codes, i, n);
    a.prefix_n[base + i] = static_cast<uint8_t>(r.dnode);
    a.left_child[base + i] = r.gamma;
    a.has_leaf_left[base + i] = r.ll ? 1 : 0;
    a.has_leaf_right[base + i] = r.lr ? 1 : 0;
  }
}

__device__ void ref_parents(const TreeArgs& a, const uint32_t* codes, uint32_t base, int n) {
  const int tid = static_cast<int>(threadIdx.x);
  int32_t* parent = a.parent + base;
  for (int i = tid; i < n; i += static_cast<int>(blockDim.x)) parent[i] = 0;
  __syncthreads();
  // node 0's claims (its Q2 children, the complete root's children) are all 0
  for (int i = tid; i < n; i += static_cast<int>(blockDim.x)) {
    if (i == 0) continue;
    const RefNode r = ref_node(codes, i, n);
    if (!r.ll) atomicMax(&parent[r.gamma], i);
    if (!r.lr) atomicMax(&parent[r.gamma + 1], i);
  }
}

// ---- keys below 2^31: the tree grown from its leaves ------------------------
// With every key below 2^31, clz(a ^ b) >= 1, so delta8 = clz - 1 has no
// uint8 wrap and  delta8(ci, x) > t  <=>  (ci ^ x) < agree_below(t).
__device__ __forceinline__ uint32_t agree_below(int t) {
  return t >= 31 ? 0u : (1u << (30 - t));
}

// x * d for a direction d of +1 or -1 given as m = (d < 0 ? -1 : 0).
__device__ __forceinline__ int signed_step(int x, int m) { return (x ^ m) - m; }

// Over unique keys the reference's searches (func_brt.hpp:57-104) find the
// canonical binary radix tree: node i's range [A, B] is the largest run of
// keys around i that share a prefix longer than i's other neighbour does,
// its split gamma is where that run's keys first differ, and i is the end
// of the range next to its parent's split (B for a left child, A for a
// right child).  The same tree grows from the leaves (Apetrei, "Fast and
// Simple Agglomerative LBVH Construction", 2014): a node [a, b] is the left
// child of a parent split at b when delta(b, b+1) > delta(a-1, a), else the
// right child of a parent split at a-1 (the two never tie over unique keys;
// deltas outside [0, n] are -1), and the second child to arrive at a split
// forms the parent.  Each node is formed once, by one thread, with one LDS
// exchange, instead of three binary searches per node.
//
// k_tree_local grows every node whose range lies inside one tile of
// kTreeTile leaves, in LDS, and writes the tile's outputs coalesced; the
// nodes whose range crosses a tile edge (at most 62 per tile: those holding
// the pair of keys at one of its two edges, at most one per prefix length
// and edge, plus node 0) are listed per tile, and k_tree_cross computes
// them with the searches above, eight lanes per node.

// LDS slot of an interior split: 0 while empty, else the first child's far
// end (its end away from the split) - lo + 1, and that end's outer delta + 1
// in the upper half.
__global__ __launch_bounds__(kTreeBlock) void k_tree_local(TreeArgs a, SegArgs seg, uint32_t* __restrict__ list) {
  constexpr int T = kTreeTile;
  __shared__ uint32_t s_ks[T + 2];  // keys lo-1 .. hi at [k - lo + 1]; then the split slots
  __shared__ uint8_t s_e[T + 2];    // delta(k, k+1) (clz) + 1 for k in [lo-1, hi-1] at [k - lo + 1]; 0 outside
  // node lo + x: its split - lo | (0x80 formed here | leaf left << 6 | leaf right << 5 | prefix) << 16
  // (entries T + lane of s_rec and s_par: each lane's scratch sink for the
  // stores that do not apply, distinct so that they share no LDS bank word)
  __shared__ alignas(16) uint32_t s_rec[T + kWave];
  __shared__ alignas(16) uint16_t s_par[T + 2 * kWave];  // node lo + x: its parent - lo (0xffff: not formed here)
  __shared__ uint16_t s_pairs[kTreeBlock / kWave][kTreePairCap];  // per wave: left leaves of leaf pairs
  __shared__ uint32_t s_n_cross;
  __shared__ uint32_t s_count_sum;  // the tile's edge counts, summed (with a.tile_sum)
  __shared__ uint32_t s_flags;      // RWH_STATUS_* bits of the bounds this tile hit
  const int tid = static_cast<int>(threadIdx.x);
  const uint32_t first = blockIdx.x * static_cast<uint32_t>(T);
  const Seg sg = seg_of(seg, first);
  const int n = static_cast<int>(sg.count);  // internal nodes; leaves (keys) 0..n
  const int lo = static_cast<int>(first - sg.base);
  if (lo >= n) return;  // whole tile past the segment's nodes
  const uint32_t* __restrict__ codes = a.codes + sg.base;
  uint32_t* __restrict__ my_list = list + static_cast<size_t>(blockIdx.x) * kTreeListStride;
  if (codes[n] >= 0x80000000u) {  // keys with bit 31 set: the reference's own search
    ref_nodes(a, codes, sg.base, lo, min(lo + T, n), n);
    if (lo == 0) ref_parents(a, codes, sg.base, n);
    return;
  }
  // leaves [lo, hi): the segment's last tile also takes leaf n
  const int hi = (lo + T < n) ? lo + T : n + 1;
  const int k_end = min(hi, n);
  if (tid == 0) s_flags = 0;
  {
    // every load in flight before the first LDS store
    constexpr int kLoads = (T + 2 + kTreeBlock - 1) / kTreeBlock;
    const int k0 = max(lo - 1, 0) + tid;
    uint32_t v[kLoads];
#pragma unroll
    for (int j = 0; j < kLoads; ++j) v[j] = (k0 + j * kTreeBlock <= k_end) ? codes[k0 + j * kTreeBlock] : 0u;
#pragma unroll
    for (int j = 0; j < kLoads; ++j)
      if (k0 + j * kTreeBlock <= k_end) s_ks[k0 + j * kTreeBlock - lo + 1] = v[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < (T + 2 + kTreeBlock - 1) / kTreeBlock; ++j) {
    const int k = lo - 1 + tid + j * kTreeBlock;
    if (k < hi) {
      const bool pair = k >= 0 && k < n;
      s_e[k - lo + 1] = pair ? static_cast<uint8_t>(__clz(s_ks[k - lo + 1] ^ s_ks[k - lo + 2]) + 1) : uint8_t{0};
      // the growth below needs strictly increasing keys (no two adjacent
      // deltas then tie); other input is reported and the tile skipped
      if (pair && s_ks[k - lo + 1] >= s_ks[k - lo + 2]) s_flags = RWH_STATUS_KEY_ORDER;
    }
  }
  __syncthreads();
  if (s_flags & RWH_STATUS_KEY_ORDER) {
    if (tid == 0) {
      my_list[0] = 0;
      if (a.tile_sum) a.tile_sum[blockIdx.x] = 0;
      if (a.status) atomicOr(&a.status[sg.id], RWH_STATUS_KEY_ORDER);
    }
    return;
  }
  uint32_t* s_slot = s_ks;  // split g in [lo, hi-2] at [g - lo]
  for (int x = tid; x < T; x += kTreeBlock) {
    s_slot[x] = 0;
    s_par[x] = 0xffff;
    s_rec[x] = 0;
  }
  if (tid == 0) {
    s_n_cross = 0;
    s_count_sum = 0;
  }
  __syncthreads();

  // Each wave grows the nodes of its block of leaves: a lane whose climb has
  // ended takes the block's next leaf, so the wave's lanes stay busy; a lane
  // climbs while it is the second child at a split, and stops at a split on
  // the tile's edge (that parent is computed by k_tree_cross).  A node
  // [na, nb] (tile offsets) is held as two words, each end with its outer
  // delta biased by one:
  //   pa = kEnd | (delta(na-1, na) + 1) << 16 | na
  //   pb = kEnd | (delta(nb, nb+1) + 1) << 16 | (nb + 1)
  // The biased deltas sit above the offsets, so pb > pa exactly when the
  // node is its parent's left child (the deltas never tie below the root).
  // A child leaves its far end's word at the split; the sibling takes it as
  // its own far end, so a level costs one LDS exchange and no LDS read.
  constexpr uint32_t kEnd = 0x80000000u;  // never 0: an empty slot is 0
  constexpr int kWaves = kTreeBlock / kWave;
  const int wave = tid / kWave;
  const int n_leaf = hi - lo;  // at most T + 1
  const int wlo = min(wave * (T / kWaves), n_leaf);
  const int whi = wave == kWaves - 1 ? n_leaf : min(wlo + T / kWaves, n_leaf);
  const uint32_t n_split = static_cast<uint32_t>(n_leaf - 1);  // interior splits: offsets [0, n_leaf - 2]
  // First every leaf arrives at its parent's split, 64 at a time with no
  // divergence; two leaves that meet form their parent at once (both
  // children leaves) and it is queued for the climb.  A single leaf's sibling
  // is an internal node, formed later, so every leaf is done here.
  uint16_t* pairs = s_pairs[wave];
  int n_pairs = 0;
  const int sink = T + static_cast<int>(tid % kWave);  // this lane's scratch entry of s_rec (s_par: 2 sink - T)
  for (int k0 = wlo; k0 < whi; k0 += kWave) {
    const int k = k0 + static_cast<int>(tid % kWave);
    const int kk = min(k, whi - 1);  // lanes past the block read a valid leaf and do nothing
    const uint32_t e0 = s_e[kk], e1 = s_e[kk + 1];
    const bool right = e1 > e0;
    const int g = right ? k : k - 1;
    const bool go = k < whi && static_cast<uint32_t>(g) < n_split;
    uint32_t other = 0;
    if (go)
      other = atomicExch(&s_slot[g], right ? (kEnd | e0 << 16 | static_cast<uint32_t>(k))
                                           : (kEnd | e1 << 16 | static_cast<uint32_t>(k + 1)));
    // the sibling is a leaf too: the parent is [g, g + 1]; two leaves over
    // the complete range [0, 1] form the root, index 0, whose record is Q2's
    const uint32_t es = (other >> 16) & 0xffu;
    const uint32_t ea = right ? e0 : es, eb = right ? es : e1;
    const uint32_t ep = right ? e1 : e0;  // delta at the split + 1
    const bool met = other != 0;
    const bool root = met && (ea | eb) == 0;
    const bool queued = met && !root;
    const int left = g;  // the queued pair's left leaf
    s_rec[queued ? (eb > ea ? g + 1 : g) : sink] = static_cast<uint32_t>(g) | (0xe0u | (ep - 2)) << 16;
    s_par[root ? 0 : 2 * sink - T] = 0;
    const uint64_t qm = ballot(queued);
    const int at = n_pairs + static_cast<int>(popc_below(qm));
    if (queued) {
      if (at < kTreePairCap) pairs[at] = static_cast<uint16_t>(left);
      else atomicOr(&s_flags, RWH_STATUS_PAIR_CAP);
    }
    n_pairs += static_cast<int>(__popcll(qm));
  }
  __syncthreads();  // every leaf has arrived: from here on only internal nodes do

  // Then the climb: a lane takes the next queued pair and climbs while it is
  // the second child at a split.
  n_pairs = min(n_pairs, kTreePairCap);
  int next = 0;
  uint32_t pa = 0, pb = 0;
  bool have = false;
  int guard = 0;
  for (; guard < kTreeClimbGuard; ++guard) {
    // lanes without a node take the wave's next queued pairs
    const uint64_t need = ballot(!have);
    const int q = next + static_cast<int>(popc_below(need));
    if (!have && q < n_pairs) {
      const int k = pairs[q];  // the pair [k, k + 1]
      pa = kEnd | static_cast<uint32_t>(s_e[k]) << 16 | static_cast<uint32_t>(k);
      pb = kEnd | static_cast<uint32_t>(s_e[k + 2]) << 16 | static_cast<uint32_t>(k + 2);
      have = true;
    }
    next += static_cast<int>(__popcll(need));
    if (ballot(have) == 0) break;
    // a lane whose parent's split lies inside the tile leaves its far end
    // there; the second to arrive forms the parent and climbs on, every other
    // lane stops (at a split on the tile's edge, or first at its split)
    const bool right = pb > pa;
    const uint32_t near = right ? pb : pa;
    const int g = static_cast<int>(near & 0xffffu) - 1;  // the parent's split
    const bool go = have && static_cast<uint32_t>(g) < n_split;
    uint32_t other = 0;
    if (go) other = atomicExch(&s_slot[g], right ? pa : pb);
    have = go && other != 0;
    if (have) {
      pb = right ? other : pb;
      pa = right ? pa : other;
      const int na = static_cast<int>(pa & 0xffffu), nb = static_cast<int>(pb & 0xffffu) - 1;
      const bool ll = g == na, lr = g + 1 == nb;
      // the complete root [0, n] (lo == 0) is index 0, whose record is Q2's:
      // its internal children's parent entries and parent[0] get 0, no record
      const bool root = ((pa | pb) & 0x00ff0000u) == 0;
      const int i = root ? 0 : (pb > pa ? nb : na);
      const uint32_t ep = (near >> 16) & 0xffu;  // delta at the split + 1: the parent's prefix + 2
      // stores that do not apply go to the lane's scratch entries past the tile
      s_rec[root ? sink : i] = static_cast<uint32_t>(g) |
                               (0x80u | (ll ? 0x40u : 0u) | (lr ? 0x20u : 0u) | (ep - 2)) << 16;
      s_par[ll ? 2 * sink - T : g] = static_cast<uint16_t>(i);
      s_par[lr ? 2 * sink - T + 1 : g + 1] = static_cast<uint16_t>(i);
      s_par[root ? 0 : 2 * sink - T] = 0;
      have = !root;
    }
  }
  // the bound ended the loop with climbs or queued pairs left
  if (guard == kTreeClimbGuard && (ballot(have) != 0 || next < n_pairs) && tid % kWave == 0)
    atomicOr(&s_flags, RWH_STATUS_CLIMB);
  __syncthreads();

  // The tile's nodes, four per thread with 16- and 4-byte stores (the
  // buffers are 16-byte aligned and lo + sg.base is a multiple of 4); those
  // formed elsewhere get placeholders that k_tree_cross overwrites, and are
  // listed for it.
  const int m = k_end - lo;
  uint8_t* __restrict__ pfx = a.prefix_n + sg.base + lo;
  uint8_t* __restrict__ hll = a.has_leaf_left + sg.base + lo;
  uint8_t* __restrict__ hlr = a.has_leaf_right + sg.base + lo;
  int32_t* __restrict__ lc = a.left_child + sg.base + lo;
  int32_t* __restrict__ par = a.parent + sg.base + lo;
  const auto list_node = [&](int x) {
    const uint32_t at = atomicAdd(&s_n_cross, 1u);
    if (at < static_cast<uint32_t>(kTreeListCap)) my_list[1 + at] = static_cast<uint32_t>(lo + x);
    else atomicOr(&s_flags, RWH_STATUS_CROSS_CAP);
  };
  const auto parent_of = [&](uint32_t p) { return p == 0xffffu ? -1 : lo + static_cast<int>(p); };
  // S5's count prefix/3 - parent's prefix/3 when the parent was formed here
  // and is not node 0 (whose prefix is the Q2 record's); else a placeholder
  // that k_tree_cross overwrites (the parent crosses the tile's edge, or is
  // the root)
  int32_t* __restrict__ ec = a.edge_count ? a.edge_count + sg.base + lo : nullptr;
  const auto count_of = [&](uint32_t rec, uint32_t p) {
    if (p == 0xffffu || lo + static_cast<int>(p) == 0) return 0;
    return static_cast<int>((rec >> 16) & 31u) / 3 - static_cast<int>((s_rec[p] >> 16) & 31u) / 3;
  };
  uint32_t count_sum = 0;  // this thread's edge counts (placeholders add 0)
  for (int x = 4 * tid; x < m; x += 4 * kTreeBlock) {
    if (x + 4 <= m) {
      const uint4 r4 = *reinterpret_cast<const uint4*>(&s_rec[x]);
      const uint2 p2 = *reinterpret_cast<const uint2*>(&s_par[x]);
      // the four meta bytes, node x in the low byte
      const uint32_t meta = __builtin_amdgcn_perm(r4.y, r4.x, 0x0c0c0602u) |
                            __builtin_amdgcn_perm(r4.w, r4.z, 0x06020c0cu);
      *reinterpret_cast<uint32_t*>(pfx + x) = meta & 0x1f1f1f1fu;
      *reinterpret_cast<uint32_t*>(hll + x) = (meta >> 6) & 0x01010101u;
      *reinterpret_cast<uint32_t*>(hlr + x) = (meta >> 5) & 0x01010101u;
      *reinterpret_cast<int4*>(lc + x) =
          make_int4(lo + static_cast<int>(r4.x & 0xffffu), lo + static_cast<int>(r4.y & 0xffffu),
                    lo + static_cast<int>(r4.z & 0xffffu), lo + static_cast<int>(r4.w & 0xffffu));
      // no later kernel of the build reads parent[]: a nontemporal store
      store_stream(reinterpret_cast<int4*>(par + x), make_int4(parent_of(p2.x & 0xffffu), parent_of(p2.x >> 16),
                                                               parent_of(p2.y & 0xffffu), parent_of(p2.y >> 16)));
      if (ec) {
        const int4 c4 = make_int4(count_of(r4.x, p2.x & 0xffffu), count_of(r4.y, p2.x >> 16),
                                  count_of(r4.z, p2.y & 0xffffu), count_of(r4.w, p2.y >> 16));
        *reinterpret_cast<int4*>(ec + x) = c4;
        count_sum += static_cast<uint32_t>(c4.x + c4.y + c4.z + c4.w);
      }
      const uint32_t cross = ~meta & 0x80808080u;
      if (cross) {
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (cross & (0x80u << (8 * c))) list_node(x + c);
      }
    } else {
      for (int c = 0; x + c < m; ++c) {
        const uint32_t rec = s_rec[x + c];
        const uint32_t meta = (rec >> 16) & 0xffu;
        pfx[x + c] = static_cast<uint8_t>(meta & 31u);
        hll[x + c] = static_cast<uint8_t>((meta >> 6) & 1u);
        hlr[x + c] = static_cast<uint8_t>((meta >> 5) & 1u);
        lc[x + c] = lo + static_cast<int>(rec & 0xffffu);
        par[x + c] = parent_of(s_par[x + c]);
        if (ec) {
          const int cc = count_of(rec, s_par[x + c]);
          ec[x + c] = cc;
          count_sum += static_cast<uint32_t>(cc);
        }
        if (!(meta & 0x80u)) list_node(x + c);
      }
    }
  }
  if (a.tile_sum) {
    const uint32_t ws = wave_sum(count_sum);
    if (tid % kWave == 0 && ws != 0) atomicAdd(&s_count_sum, ws);
  }
  __syncthreads();
  if (tid == 0) {
    my_list[0] = min(s_n_cross, static_cast<uint32_t>(kTreeListCap));
    // k_tree_cross adds the counts it writes over this tile's placeholders
    if (a.tile_sum) a.tile_sum[blockIdx.x] = s_count_sum;
    if (s_flags && a.status) atomicOr(&a.status[sg.id], s_flags);
  }
}

// Largest x in [lo, hi] with p(x), for p true at lo and true-then-false over
// [lo, hi], by the kCrossGroup lanes of one group (lane r of the group, its
// lanes at bit gshift of the wave's ballot): each round probes kCrossGroup
// evenly spaced points and keeps the span after the last true one.  All
// lanes of the group call it with the same lo and hi.
template <class P>
__device__ __forceinline__ int group_last_true(const P& p, int lo, int hi, int r, int gshift) {
  while (lo < hi) {
    const int step = (hi - lo + kCrossGroup - 1) / kCrossGroup;
    const int x = lo + (r + 1) * step;
    const bool t = x <= hi && p(x);
    const uint32_t bits = static_cast<uint32_t>(ballot(t) >> gshift) & kCrossMask;
    lo += __popc(bits) * step;
    hi = min(lo + step - 1, hi);
  }
  return lo;
}

// The same search for the predicate (ci ^ key(x)) < thr with key(x) =
// codes[i + x*d] (m = d < 0 ? -1 : 0), which also keeps the probed keys at
// the ends: on entry klo = key(lo) and khi = key(hi + 1) (if known); on
// return klo = key(x) and khi = key(x + 1) when x < hi on entry, or the
// entry khi otherwise.
__device__ __forceinline__ int group_search(const uint32_t* __restrict__ codes, int i, int m, uint32_t ci,
                                            uint32_t thr, int lo, int hi, uint32_t& klo, uint32_t& khi, int r,
                                            int gshift) {
  while (lo < hi) {
    const int step = (hi - lo + kCrossGroup - 1) / kCrossGroup;
    const int x = lo + (r + 1) * step;
    uint32_t key = 0;
    bool t = false;
    if (x <= hi) {
      key = codes[i + signed_step(x, m)];
      t = (ci ^ key) < thr;
    }
    const int k = __popc(static_cast<uint32_t>(ballot(t) >> gshift) & kCrossMask);
    // the last true probe (lane k-1) and the first false one (lane k)
    const uint32_t key_t = __shfl(key, gshift + max(k - 1, 0), kWave);
    const uint32_t key_f = __shfl(key, gshift + min(k, kCrossGroup - 1), kWave);
    const int nlo = lo + k * step;
    const int nhi = min(nlo + step - 1, hi);
    if (k > 0) klo = key_t;
    if (nhi < hi) khi = key_f;  // nhi + 1 = nlo + step is lane k's probe
    lo = nlo;
    hi = nhi;
  }
  return lo;
}

// The edge count of node gi (index space) over k_tree_local's placeholder 0,
// added to its tile's sum when the tile sums are kept.
__device__ __forceinline__ void put_count(const TreeArgs& a, uint32_t gi, int count) {
  a.edge_count[gi] = count;
  if (a.tile_sum && count != 0) atomicAdd(&a.tile_sum[gi / kTreeTile], static_cast<uint32_t>(count));
}

// Node i (i >= 1) of a tree over keys below 2^31, by one group: its range
// end and split with the reference's predicates, then its record and its
// internal children's parent entries (lane 0 of the group).
__device__ void cross_node(const TreeArgs& a, uint32_t base, const uint32_t* __restrict__ codes, int n, int i,
                           int r, int gshift) {
  const uint32_t ci = codes[i];
  const int cp = __clz(ci ^ codes[i + 1]), cm = __clz(ci ^ codes[i - 1]);
  const int d = cp > cm ? 1 : -1;  // never equal over unique keys
  const int m = d < 0 ? -1 : 0;
  // delta(i, i+x*d) > delta(i, i-d) for x in [0, l]; x <= cutoff keeps the probe in [0, n]
  const uint32_t thr = agree_below((d > 0 ? cm : cp) - 1);
  const int cutoff = d < 0 ? i : n - i;
  // bracket l: lane r probes 2^(r * kCrossBracket) (x = 1 always holds)
  const int x4 = 1 << (kCrossBracket * r);
  uint32_t key4 = 0;
  bool t4 = false;
  if (x4 <= cutoff) {
    key4 = codes[i + signed_step(x4, m)];
    t4 = (ci ^ key4) < thr;
  }
  const int k = max(__popc(static_cast<uint32_t>(ballot(t4) >> gshift) & kCrossMask), 1);
  const int blo = 1 << (kCrossBracket * (k - 1));
  const int bhi = k < kCrossGroup ? min((1 << (kCrossBracket * k)) - 1, cutoff) : cutoff;
  uint32_t cj = __shfl(key4, gshift + k - 1, kWave), unused = 0;
  const int l = group_search(codes, i, m, ci, thr, blo, bhi, cj, unused, r, gshift);
  const int j = i + signed_step(l, m);
  const int dnode = delta8(ci, cj);
  // the split: the largest s in [0, l) with delta(i, i+s*d) > dnode; the keys
  // at s and s + 1 are those at gamma and gamma + 1
  uint32_t ks = ci, ks1 = cj;
  const int s = group_search(codes, i, m, ci, agree_below(dnode), 0, l - 1, ks, ks1, r, gshift);
  const int gamma = i + signed_step(s, m) + min(d, 0);
  if (r == 0) {
    const bool ll = min(i, j) == gamma, lr = max(i, j) == gamma + 1;
    a.prefix_n[base + i] = static_cast<uint8_t>(dnode);
    a.left_child[base + i] = gamma;
    a.has_leaf_left[base + i] = ll ? 1 : 0;
    a.has_leaf_right[base + i] = lr ? 1 : 0;
    if (!ll) a.parent[base + gamma] = i;
    if (!lr) a.parent[base + gamma + 1] = i;
    if (a.edge_count) {  // the internal children's counts: their prefixes from their end keys
      const uint32_t c_lo = d > 0 ? ci : cj, c_hi = d > 0 ? cj : ci;
      const uint32_t c_g = d > 0 ? ks : ks1, c_g1 = d > 0 ? ks1 : ks;  // keys at gamma, gamma + 1
      if (!ll) put_count(a, base + gamma, delta8(c_lo, c_g) / 3 - dnode / 3);
      if (!lr) put_count(a, base + gamma + 1, delta8(c_g1, c_hi) / 3 - dnode / 3);
    }
  }
}

// Node 0: the Q2 record over [0, n-1], parent[0] = 0, and 0 for the internal
// children of the complete root over [0, n] (the slots no node i >= 1 claims).
__device__ void cross_root(const TreeArgs& a, uint32_t base, const uint32_t* __restrict__ codes, int n, int r,
                           int gshift) {
  const uint32_t c0 = codes[0];
  const int l = n - 1;  // Q2
  const int dnode = delta8(c0, codes[l]);
  const uint32_t thr = agree_below(dnode);
  const int gamma = group_last_true([&](int x) { return (c0 ^ codes[x]) < thr; }, 0, max(l - 1, 0), r, gshift);
  const uint32_t thr_full = agree_below(delta8(c0, codes[n]));
  const int g_full = group_last_true([&](int x) { return (c0 ^ codes[x]) < thr_full; }, 0, n - 1, r, gshift);
  if (r == 0) {
    a.prefix_n[base] = static_cast<uint8_t>(dnode);
    a.left_child[base] = gamma;
    a.has_leaf_left[base] = gamma == 0 ? 1 : 0;
    a.has_leaf_right[base] = l == gamma + 1 ? 1 : 0;
    a.parent[base] = 0;
    if (g_full != 0) a.parent[base + g_full] = 0;
    if (g_full + 1 < n) a.parent[base + g_full + 1] = 0;
    if (a.edge_count) {  // count[0] = 0; the complete root's children count against node 0's Q2 prefix
      a.edge_count[base] = 0;
      if (g_full != 0) put_count(a, base + g_full, delta8(c0, codes[g_full]) / 3 - dnode / 3);
      if (g_full + 1 < n) put_count(a, base + g_full + 1, delta8(codes[g_full + 1], codes[n]) / 3 - dnode / 3);
    }
  }
}

__global__ __launch_bounds__(kCrossBlock) void k_tree_cross(TreeArgs a, SegArgs seg,
                                                           const uint32_t* __restrict__ list) {
  const uint32_t first = blockIdx.x * static_cast<uint32_t>(kTreeTile);
  const Seg sg = seg_of(seg, first);
  const int n = static_cast<int>(sg.count);
  const int lo = static_cast<int>(first - sg.base);
  if (lo >= n) return;
  const uint32_t* __restrict__ codes = a.codes + sg.base;
  if (codes[n] >= 0x80000000u) return;  // done by the reference's search
  const uint32_t* __restrict__ my_list = list + static_cast<size_t>(blockIdx.x) * kTreeListStride;
  const int cnt = static_cast<int>(my_list[0]);
  const int tid = static_cast<int>(threadIdx.x);
  const int grp = tid / kCrossGroup, r = tid % kCrossGroup;
  const int gshift = (tid % kWave) & ~(kCrossGroup - 1);
  for (int e = grp; e < cnt; e += kCrossBlock / kCrossGroup) {
    const int i = static_cast<int>(my_list[1 + e]);
    if (i == 0) {
      cross_root(a, sg.base, codes, n, r, gshift);
    } else {
      cross_node(a, sg.base, codes, n, i, r, gshift);
    }
  }
}

void launch_radix_tree(const TreeArgs& a, SegArgs seg, uint32_t n_index, uint32_t* list, hipStream_t stream) {
  const uint32_t grid = (n_index + kTreeTile - 1) / kTreeTile;
  if (grid == 0) return;
  seg.count_sub = 1;
  hipLaunchKernelGGL(k_tree_local, dim3(grid), dim3(kTreeBlock), 0, stream, a, seg, list);
  hipLaunchKernelGGL(k_tree_cross, dim3(grid), dim3(kCrossBlock), 0, stream, a, seg, list);
}

// ---------------------------------------------------------------------------
// S5  edge count (reference: apps/tree/host/func_edge.hpp:9-16,
// apps/tree/cuda/05_edge_count.cu:16-29): count[i] = prefix_n[i]/3 -
// prefix_n[parent[i]]/3, count[0] = 0.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kElemBlock) void k_edge_count(const uint8_t* __restrict__ prefix_n,
                                                           const int32_t* __restrict__ parent,
                                                           int32_t* __restrict__ edge_count,
                                                           SegArgs seg) {
  const uint32_t gi = blockIdx.x * kElemBlock + threadIdx.x;
  const Seg sg = seg_of(seg, blockIdx.x * kElemBlock);
  const uint32_t i = gi - sg.base;
  if (gi < sg.base || i >= sg.count) return;
  const int p = parent[gi];
  const int c = static_cast<int>(prefix_n[gi] / 3) - static_cast<int>(prefix_n[sg.base + p] / 3);
  edge_count[gi] = (i == 0) ? 0 : c;
}

void launch_edge_count(const uint8_t* prefix_n, const int32_t* parent, int32_t* edge_count,
                       SegArgs seg, uint32_t n_index, hipStream_t stream) {
  const uint32_t grid = (n_index + kElemBlock - 1) / kElemBlock;
  if (grid == 0) return;
  seg.count_sub = 1;
  hipLaunchKernelGGL(k_edge_count, dim3(grid), dim3(kElemBlock), 0, stream, prefix_n, parent,
                     edge_count, seg);
}

// ---------------------------------------------------------------------------
// S6  exclusive prefix sum (reference: apps/tree/cuda/06_prefix_sum.cu:26-37,
// a one-workgroup scan there; D4).  Single pass; the segment's last tile
// writes the total (n_octree_nodes).  A tile's prefix over the segment's
// earlier tiles comes from a decoupled look-back, or (kTiles) from the sums
// of the earlier tree tiles that the radix tree kernels left (tile_sum).
// ---------------------------------------------------------------------------
template <int BLOCK, int KPT, bool kTiles = false>
__global__ __launch_bounds__(BLOCK) void k_prefix_sum(const int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                      SegArgs seg, uint32_t* __restrict__ total,
                                                      uint64_t* __restrict__ lookback,
                                                      uint32_t* __restrict__ ctrl,
                                                      const uint32_t* __restrict__ tile_sum) {
  constexpr int G = KPT / 4;
  constexpr uint32_t TILE = BLOCK * KPT;
  static_assert(!kTiles || TILE % kTreeTile == 0, "a scan tile holds whole tree tiles");
  __shared__ uint32_t s_w[BLOCK / kWave];
  __shared__ uint32_t s_w2[BLOCK / kWave];
  __shared__ uint32_t s_bcast;
  __shared__ Grab s_grab;
  const uint32_t tid = threadIdx.x;
  const uint32_t wave = tid / kWave, lane = tid % kWave;
  if (!kTiles) {
    if (tid == 0) s_grab = grab_tiles(ctrl, gridDim.x);
    __syncthreads();
  }
  const uint32_t epoch = kTiles ? 0u : s_grab.epoch;
  // One tile per workgroup.  With the look-back, in grab order: a tile only
  // ever waits on tiles whose workgroups took their grab earlier and are
  // therefore running, so the look-back spin ends whatever the residency of
  // the grid.
  {
    const uint32_t tile = kTiles ? blockIdx.x : s_grab.index;
    const uint32_t first = tile * TILE;
    const Seg sg = seg_of(seg, first);
    const uint32_t lim = sg.base + sg.count;
    if (first >= lim) {
      if (tid == 0 && first == sg.base) total[sg.id] = 0;
      return;
    }
    const uint32_t tile_in_seg = (first - sg.base) / TILE;
    const uint32_t wbase = first + wave * (KPT * kWave);
    uint32_t v[G][4];
    uint32_t gsum[G];
    const bool full = first + TILE <= lim;
    if (full) {  // plain 16-byte loads
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int4 q = *reinterpret_cast<const int4*>(in + wbase + 4 * (g * kWave + lane));
        v[g][0] = q.x; v[g][1] = q.y; v[g][2] = q.z; v[g][3] = q.w;
      }
    } else {
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const uint32_t i0 = wbase + 4 * (g * kWave + lane);
#pragma unroll
        for (int c = 0; c < 4; ++c) v[g][c] = (i0 + c < lim) ? static_cast<uint32_t>(in[i0 + c]) : 0u;
      }
    }
    // kTiles: the segment's tree tiles before this tile, summed by the block
    uint32_t before = 0;
    if (kTiles) {
      const uint32_t t0 = sg.base / kTreeTile, nt = (first - sg.base) / kTreeTile;
      for (uint32_t j = tid; j < nt; j += BLOCK) before += tile_sum[t0 + j];
    }
#pragma unroll
    for (int g = 0; g < G; ++g) gsum[g] = v[g][0] + v[g][1] + v[g][2] + v[g][3];
    uint32_t wave_run = 0;
    uint32_t gexcl[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const uint32_t incl = wave_incl_scan(gsum[g]);
      gexcl[g] = wave_run + incl - gsum[g];
      wave_run += __shfl(incl, kWave - 1, kWave);
    }
    uint32_t tile_total;
    const uint32_t wexcl = block_excl_scan<BLOCK>(lane == 0 ? wave_run : 0u, s_w, &tile_total);
    const uint32_t wave_off = __shfl(wexcl, 0, kWave);
    uint32_t excl;
    if (kTiles) {
      block_excl_scan<BLOCK>(before, s_w2, &excl);
    } else {
      excl = tile_prefix(tile_total, tile, tile_in_seg, epoch, lookback, &s_bcast);
    }
    const uint32_t base = excl + wave_off;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const uint32_t i0 = wbase + 4 * (g * kWave + lane);
      uint32_t r = base + gexcl[g];
      uint32_t o[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        o[c] = r;
        r += v[g][c];
      }
      if (full) {
        *reinterpret_cast<int4*>(out + i0) = make_int4(static_cast<int>(o[0]), static_cast<int>(o[1]),
                                                       static_cast<int>(o[2]), static_cast<int>(o[3]));
      } else {
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (i0 + c < lim) out[i0 + c] = static_cast<int>(o[c]);
      }
    }
    if (tid == 0 && lim - first <= TILE) total[sg.id] = excl + tile_total;
  }
}

// Launches the scan of the live count of each segment on tiles of BLOCK x KPT
// (one cloud: the larger tile).
void launch_prefix_sum(const int32_t* in, int32_t* out, SegArgs seg, uint32_t n_index,
                       uint32_t* total, uint64_t* lookback, uint32_t* ctrl, hipStream_t stream) {
  if (n_index == 0) return;
  seg.count_sub = 1;
  if (seg.n_segs == 1) {
    const uint32_t n_tiles = (n_index + kScanBigTile - 1) / kScanBigTile;
    hipLaunchKernelGGL((k_prefix_sum<kScanBigBlock, kScanBigKpt>), dim3(n_tiles), dim3(kScanBigBlock), 0,
                       stream, in, out, seg, total, lookback, ctrl, nullptr);
    return;
  }
  const uint32_t n_tiles = (n_index + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL((k_prefix_sum<kScanBlock, kScanKpt>), dim3(n_tiles), dim3(kScanBlock), 0, stream, in,
                     out, seg, total, lookback, ctrl, nullptr);
}

// (No chain between tiles here, so the smaller tile of any cloud layout,
// which keeps more of them in flight, is also the faster one for one cloud.)
void launch_prefix_sum_tiles(const int32_t* in, int32_t* out, SegArgs seg, uint32_t n_index,
                             uint32_t* total, const uint32_t* tile_sum, hipStream_t stream) {
  if (n_index == 0) return;
  seg.count_sub = 1;
  const uint32_t n_tiles = (n_index + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL((k_prefix_sum<kScanBlock, kScanKpt, true>), dim3(n_tiles), dim3(kScanBlock), 0, stream,
                     in, out, seg, total, nullptr, nullptr, tile_sum);
}

}  // namespace rw
