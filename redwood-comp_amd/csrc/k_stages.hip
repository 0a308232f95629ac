// Stages 1, 3, 4, 5, 6 of the octree build: Morton encode, duplicate removal,
// Karras radix tree, edge count, exclusive prefix sum.  Every kernel works on
// segments (clouds) of one index space; a single build is one segment.
#include "rw_device.hpp"
#include "rw_internal.hpp"

namespace rw {

// ---------------------------------------------------------------------------
// S1  Morton encode (reference: apps/tree/cuda/01_morton.cu:8-19,
// apps/tree/host/func_morton.hpp:31-39).  A workgroup covers 4*256 points;
// every load is a whole float4 (16 bytes) and lanes read consecutive points.
// ---------------------------------------------------------------------------
constexpr uint32_t kMortonPerThread = 4;
constexpr uint32_t kMortonTile = kElemBlock * kMortonPerThread;  // 1024

__global__ __launch_bounds__(kElemBlock) void k_morton(const float4* __restrict__ pts,
                                                       uint32_t* __restrict__ codes, SegArgs seg,
                                                       float min_coord, float range) {
  const uint32_t first = blockIdx.x * kMortonTile;
  const Seg sg = seg_of(seg, first);
  const uint32_t lim = sg.base + sg.count;
#pragma unroll
  for (uint32_t j = 0; j < kMortonPerThread; ++j) {
    const uint32_t i = first + j * kElemBlock + threadIdx.x;
    if (i < lim) {
      const float4 p = pts[i];
      codes[i] = morton_encode(p.x, p.y, p.z, min_coord, range);
    }
  }
}

void launch_morton(const float4* pts, uint32_t* codes, SegArgs seg, uint32_t n_index,
                   float min_coord, float range, hipStream_t stream) {
  const uint32_t grid = (n_index + kMortonTile - 1) / kMortonTile;
  if (grid == 0) return;
  hipLaunchKernelGGL(k_morton, dim3(grid), dim3(kElemBlock), 0, stream, pts, codes, seg, min_coord,
                     range);
}

// ---------------------------------------------------------------------------
// Tile-wide exclusive scan for kScanBlock threads.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_w, uint32_t* total) {
  constexpr int W = kScanBlock / kWave;
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
__global__ __launch_bounds__(kScanBlock) void k_unique(const uint32_t* __restrict__ in,
                                                       uint32_t* __restrict__ out, SegArgs seg,
                                                       uint32_t* __restrict__ n_unique,
                                                       uint64_t* __restrict__ lookback,
                                                       uint32_t* __restrict__ ctrl) {
  constexpr int G = kScanKpt / 4;
  __shared__ uint32_t s_w[kScanBlock / kWave];
  __shared__ uint32_t s_bcast;
  __shared__ Grab s_grab;
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
    const uint32_t first = tile * kScanTile;
    const Seg sg = seg_of(seg, first);
    const uint32_t lim = sg.base + sg.count;
    if (first >= lim) {
      // an empty segment still reports n_unique = 0 from its first tile
      if (tid == 0 && first == sg.base) n_unique[sg.id] = 0;
      return;
    }
    const uint32_t tile_in_seg = (first - sg.base) / kScanTile;
    const uint32_t wbase = first + wave * (kScanKpt * kWave);

    uint32_t k[G][4];
    uint32_t flags[G];  // 4 head bits per group
    uint32_t cnt[G];
    if (first + kScanTile <= lim) {  // full tile: plain 16-byte loads
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
    const uint32_t wexcl = block_excl_scan(lane == 0 ? wave_run : 0u, s_w, &tile_total);
    const uint32_t wave_off = __shfl(wexcl, 0, kWave);
    const uint32_t excl = tile_prefix(tile_total, tile, tile_in_seg, epoch, lookback, &s_bcast);
    const uint32_t base = sg.base + excl + wave_off;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      uint32_t pos = base + gexcl[g];
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if ((flags[g] >> c) & 1u) out[pos++] = k[g][c];
    }
    // the tile holding the segment's last key reports the count
    if (tid == 0 && lim - first <= kScanTile) n_unique[sg.id] = excl + tile_total;
  }
}

void launch_unique(const uint32_t* sorted, uint32_t* unique_out, SegArgs seg, uint32_t n_index,
                   uint32_t* n_unique, uint64_t* lookback, uint32_t* ctrl, hipStream_t stream) {
  const uint32_t n_tiles = (n_index + kScanTile - 1) / kScanTile;
  if (n_tiles == 0) return;
  hipLaunchKernelGGL(k_unique, dim3(n_tiles), dim3(kScanBlock), 0, stream, sorted, unique_out, seg,
                     n_unique, lookback, ctrl);
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

// The whole segment with the reference's search.  Such a tree need not be a
// proper binary tree (a parent slot can have several claimers), so parent[]
// is the maximum claimer over zeroed slots (D7) built with atomics by the
// segment's first workgroup alone (zeroing, barrier, claims); the other
// outputs are written by each node's own thread.
__device__ void ref_tile(const TreeArgs& a, const uint32_t* codes, uint32_t base, int b0, int n) {
  const int tid = static_cast<int>(threadIdx.x);
  if (b0 + tid < n) {
    const int i = b0 + tid;
    const RefNode r = ref_node(codes, i, n);
    a.prefix_n[base + i] = static_cast<uint8_t>(r.dnode);
    a.left_child[base + i] = r.gamma;
    a.has_leaf_left[base + i] = r.ll ? 1 : 0;
    a.has_leaf_right[base + i] = r.lr ? 1 : 0;
  }
  if (b0 != 0) return;
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

// ---- nearest smaller adjacent deltas (keys below 2^31) ---------------------
// With a[k] = delta8(code[k], code[k+1]) and keys sorted, unique and below
// 2^31, delta8(code[x], code[y]) = min(a[x..y-1]), and the reference's searches
// reduce to nearest-smaller-value queries: the internal node split at k covers
// keys [L, R] = [psv(k) + 1, nsv(k)] (psv/nsv: nearest position on the left /
// right with a smaller a, -1 / n if none), its prefix is a[k], its children
// are k and k + 1, and its index is L if a[psv] > a[nsv] (it is its parent's
// right child) and R otherwise.  The split with no smaller value on either
// side is the root of the complete range [0, n]; node 0 reports the range
// [0, n-1] instead (Q2): the same split unless that is n - 1, else the split
// whose range is exactly [0, n-1].  Pinned against the reference search on
// the CPU by the tests.
//
// A workgroup owns kTreeBlock splits and keeps a[] for a window reaching
// kTreeHalo further on each side in LDS, in blocks of 64 (one wave):
//   * inside the split's own block: binary lifting over a per-wave sparse
//     table of minima (6 steps);
//   * across blocks: a bitmask per threshold t of the blocks whose minimum is
//     below t picks the block, and a binary search over that block's prefix
//     (or suffix) minima the position;
//   * beyond the window: a wave-cooperative 64-ary search in global memory.
constexpr int kTreeBlock = 1024;  // splits per workgroup, one per thread
constexpr int kTreeHalo = 1024;   // window margin on each side (multiple of 64)
constexpr int kTreeWin = kTreeBlock + 2 * kTreeHalo;
constexpr int kTreeWinBlocks = kTreeWin / kWave;
static_assert(kTreeWinBlocks <= 64, "one bit per window block");
static_assert(kChunk % kTreeBlock == 0, "tiles must not straddle segments");

// First t in [lo, hi] where pred holds, for pred false ... false true ... true
// on [lo, hi]; hi + 1 if it never holds.  Wave-uniform arguments; every lane
// calls it; 64 probes per round trip.
template <class Pred>
__device__ int wave_first_true(int lo, int hi, Pred pred) {
  const uint32_t lane = lane_id();
  for (;;) {
    if (lo > hi) return lo;
    const uint32_t span = static_cast<uint32_t>(hi - lo) + 1u;
    if (span <= static_cast<uint32_t>(kWave)) {
      const int t = lo + static_cast<int>(lane);
      const uint64_t b = ballot(t <= hi && pred(t));
      return b ? lo + static_cast<int>(__ffsll(static_cast<long long>(b)) - 1) : hi + 1;
    }
    // strictly increasing probes, the last one at hi
    const int t = lo + static_cast<int>((static_cast<uint64_t>(lane + 1) * span) / kWave) - 1;
    const uint64_t b = ballot(pred(t));
    if (b == 0) return hi + 1;
    const int c = __ffsll(static_cast<long long>(b)) - 1;
    const int tc = __shfl(t, c, kWave);
    const int tp = (c == 0) ? lo - 1 : __shfl(t, c - 1, kWave);
    lo = tp + 1;
    hi = tc;
  }
}

// Writes the tile's entries [0, nv) that this workgroup produced (own[e] =
// 1) to global memory: whole 16-byte groups with one store, the rest entry by
// entry (dst 16-byte aligned).
template <typename T>
__device__ __forceinline__ void flush_owned(T* __restrict__ dst, const T* src, const uint8_t* own,
                                            int nv, int first_group, int lanes) {
  constexpr int kPer = 16 / sizeof(T);
  for (int g = first_group; g * kPer < nv; g += lanes) {
    bool whole = (g + 1) * kPer <= nv;
    if (whole) {
      if constexpr (kPer == 16) {
        const uint4 o = reinterpret_cast<const uint4*>(own)[g];
        whole = (o.x & o.y & o.z & o.w) == 0x01010101u;
      } else {
        whole = reinterpret_cast<const uint32_t*>(own)[g] == 0x01010101u;
      }
    }
    if (whole) {
      *reinterpret_cast<uint4*>(dst + g * kPer) = reinterpret_cast<const uint4*>(src)[g];
    } else {
      for (int e = g * kPer; e < min(nv, (g + 1) * kPer); ++e)
        if (own[e]) dst[e] = src[e];
    }
  }
}

__global__ __launch_bounds__(kTreeBlock) void k_radix_tree(TreeArgs a, SegArgs seg, uint32_t n_tiles) {
  __shared__ uint8_t s_a[kTreeWin];                   // a[w0 + p], 0xff past the window
  __shared__ uint8_t s_pmin[kTreeWin], s_smin[kTreeWin];  // per-block prefix / suffix minima
  __shared__ uint64_t s_below[32];                    // bit b: block b's minimum < t
  __shared__ __align__(16) int32_t s_lc[kTreeBlock];
  __shared__ __align__(16) uint8_t s_pfx[kTreeBlock];
  __shared__ __align__(16) uint8_t s_ll[kTreeBlock];
  __shared__ __align__(16) uint8_t s_lr[kTreeBlock];
  __shared__ __align__(16) uint8_t s_own[kTreeBlock];

  // neighbouring tiles share window codes: XCD-grouped order
  const uint32_t tile = xcd_tile();
  if (tile >= n_tiles) return;
  const uint32_t first = tile * kTreeBlock;
  const Seg sg = seg_of(seg, first);
  const int n = static_cast<int>(sg.count);  // n_brt; codes hold n + 1 keys
  const int b0 = static_cast<int>(first - sg.base);
  if (first < sg.base || b0 >= n) return;  // whole workgroup past the segment's nodes
  const uint32_t* __restrict__ gcodes = a.codes + sg.base;
  const int tid = static_cast<int>(threadIdx.x);

  if (gcodes[n] >= 0x80000000u) {  // keys with bit 31 set: the reference's own search
    ref_tile(a, gcodes, sg.base, b0, n);
    return;
  }
  int32_t* __restrict__ parent = a.parent + sg.base;
  if (n == 1) {  // two keys: node 0 alone, over [0, 0] (Q2)
    if (tid == 0) {
      a.prefix_n[sg.base] = static_cast<uint8_t>(delta8(gcodes[0], gcodes[0]));
      a.left_child[sg.base] = 0;
      a.has_leaf_left[sg.base] = 1;
      a.has_leaf_right[sg.base] = 0;
      parent[0] = 0;
    }
    return;
  }

  const int lane = tid % kWave, wv = tid / kWave;
  const int w0 = max(b0 - kTreeHalo, 0);
  const int w1 = min(b0 + kTreeBlock + kTreeHalo, n);  // a[] window [w0, w1)
  const int nblk = (w1 - w0 + kWave - 1) / kWave;

  // 1. a[] over the window
  for (int p = tid; p < kTreeWin; p += kTreeBlock) {
    const int k = w0 + p;
    s_a[p] = (k < w1) ? static_cast<uint8_t>(delta8(gcodes[k], gcodes[k + 1])) : 0xffu;
  }
  s_own[tid] = 0;
  __syncthreads();

  // 2. per-block prefix / suffix minima
  for (int blk = wv; blk < nblk; blk += kTreeBlock / kWave) {
    const int p = blk * kWave + lane;
    uint32_t pm = s_a[p], sm = pm;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t u = __shfl_up(pm, o, kWave);
      const uint32_t v = __shfl_down(sm, o, kWave);
      if (lane >= o) pm = min(pm, u);
      if (lane + o < kWave) sm = min(sm, v);
    }
    s_pmin[p] = static_cast<uint8_t>(pm);
    s_smin[p] = static_cast<uint8_t>(sm);
  }
  __syncthreads();

  // 3. blocks below each threshold (one wave)
  if (wv == 0) {
    const uint32_t bm = (lane < nblk) ? s_pmin[lane * kWave + kWave - 1] : 0xffu;
    uint64_t mine = 0;
#pragma unroll
    for (int t = 0; t < 32; ++t) {
      const uint64_t m = ballot(bm < static_cast<uint32_t>(t));
      if (lane == t) mine = m;
    }
    if (lane < 32) s_below[lane] = mine;
  }
  __syncthreads();

  // 4. nearest smaller values of this thread's split k
  const int k = b0 + tid;
  const bool valid = k < n;
  const uint32_t v = s_a[k - w0];
  const int wbase = b0 + wv * kWave;          // first split of this wave
  const int wblk = (wbase - w0) / kWave;      // its window block
  // sparse table of minima over the wave: lvl[j] at lane i = min(a[i .. i + 2^j - 1])
  uint32_t lvl[6];
  lvl[0] = v;
#pragma unroll
  for (int j = 1; j < 6; ++j) {
    const uint32_t u = __shfl_down(lvl[j - 1], 1 << (j - 1), kWave);
    lvl[j] = (lane + (1 << (j - 1)) < kWave) ? min(lvl[j - 1], u) : lvl[j - 1];
  }
  // right: skip maximal runs >= v (binary lifting); 64 = none in the wave
  int nx = lane + 1;
#pragma unroll
  for (int j = 5; j >= 0; --j) {
    const uint32_t m = __shfl(lvl[j], min(nx, kWave - 1), kWave);
    if (nx + (1 << j) <= kWave && m >= v) nx += 1 << j;
  }
  // left: runs ending at px; -1 = none in the wave
  int px = lane - 1;
#pragma unroll
  for (int j = 5; j >= 0; --j) {
    const int st = px - (1 << j) + 1;
    const uint32_t m = __shfl(lvl[j], max(st, 0), kWave);
    if (st >= 0 && m >= v) px -= 1 << j;
  }
  int nsv = n, psv = -1;
  bool far_n = false, far_p = false;
  if (valid) {
    if (nx < kWave) {
      nsv = wbase + nx;
    } else {
      const uint64_t cand = s_below[v] & (wblk + 1 < 64 ? (~0ull << (wblk + 1)) : 0ull);
      if (cand) {
        const int b2 = __ffsll(static_cast<long long>(cand)) - 1;
        const uint8_t* pm = s_pmin + b2 * kWave;
        int q = 0;  // first position of block b2 below v
#pragma unroll
        for (int st = 32; st >= 1; st >>= 1)
          if (pm[q + st - 1] >= v) q += st;
        nsv = w0 + b2 * kWave + q;
      } else if (w1 < n) {
        far_n = true;
      }
    }
    if (px >= 0) {
      psv = wbase + px;
    } else {
      const uint64_t cand = s_below[v] & ((1ull << wblk) - 1);
      if (cand) {
        const int b2 = 63 - __clzll(static_cast<long long>(cand));
        const uint8_t* sm = s_smin + b2 * kWave;
        int q = kWave - 1;  // last position of block b2 below v
#pragma unroll
        for (int st = 32; st >= 1; st >>= 1)
          if (sm[q - st + 1] >= v) q -= st;
        psv = w0 + b2 * kWave + q;
      } else if (w0 > 0) {
        far_p = true;
      }
    }
  }
  // 5. the node split at k (splits whose range leaves the window wait: step 7)
  const auto a_at = [&](int x) -> int {
    int r = s_a[min(max(x - w0, 0), kTreeWin - 1)];
    asm volatile("" : "+v"(r));  // keep the LDS read; global only outside the window
    if (x < w0 || x >= w1) r = delta8(gcodes[x], gcodes[x + 1]);
    return r;
  };
  // staged: into the tile's LDS copy (when inside the tile), else straight out
  const auto put = [&](int idx, int pfx, int lc, bool ll, bool lr, bool staged) {
    const int e = idx - b0;
    if (staged && e >= 0 && e < kTreeBlock) {
      s_pfx[e] = static_cast<uint8_t>(pfx);
      s_lc[e] = lc;
      s_ll[e] = ll ? 1 : 0;
      s_lr[e] = lr ? 1 : 0;
      s_own[e] = 1;
    } else {
      a.prefix_n[sg.base + idx] = static_cast<uint8_t>(pfx);
      a.left_child[sg.base + idx] = lc;
      a.has_leaf_left[sg.base + idx] = ll ? 1 : 0;
      a.has_leaf_right[sg.base + idx] = lr ? 1 : 0;
    }
  };
  const auto emit = [&](bool staged) {
    const int L = psv + 1, R = nsv;
    int idx = 0;  // the complete-range root has index 0 (its outputs are Q2's)
    if (psv >= 0 || nsv < n) {
      const int ap = (psv >= 0) ? a_at(psv) : -1;
      const int an = (nsv < n) ? a_at(nsv) : -1;
      idx = (ap > an) ? L : R;
      put(idx, static_cast<int>(v), k, L == k, R == k + 1, staged);
    }
    // node 0 (Q2): the split of [0, n-1]
    if (psv == -1 && (nsv == n - 1 || (nsv == n && k < n - 1)))
      put(0, static_cast<int>(v), k, k == 0, k + 1 == n - 1, staged);
    if (L != k) parent[k] = idx;
    if (R != k + 1) parent[k + 1] = idx;
    if (k == 0) parent[0] = 0;
  };
  const bool far = far_n || far_p;
  if (valid && !far) emit(true);
  __syncthreads();

  // 6. the tile's staged outputs, 16 bytes per lane where whole
  const int nv = min(kTreeBlock, n - b0);
  const uint32_t at = sg.base + static_cast<uint32_t>(b0);
  if (tid < 64) flush_owned(a.prefix_n + at, s_pfx, s_own, nv, tid, 64);
  else if (tid < 128) flush_owned(a.has_leaf_left + at, s_ll, s_own, nv, tid - 64, 64);
  else if (tid < 192) flush_owned(a.has_leaf_right + at, s_lr, s_own, nv, tid - 128, 64);
  else if (tid >= 256 && tid < 512) flush_owned(a.left_child + at, s_lc, s_own, nv, tid - 256, 256);

  // 7. splits whose range leaves the window: wave-cooperative searches in
  // global memory, after the barrier so that the other waves are not held,
  // and outputs written straight out (the flush skipped their entries).
  if (!ballot(far)) return;
  // min(a[w1 .. t]) < v, i.e. delta8(code[w1], code[t + 1]) < v
  for (uint64_t m = ballot(far_n); m; m &= m - 1) {
    const int o = __ffsll(static_cast<long long>(m)) - 1;
    const int ov = __shfl(static_cast<int>(v), o, kWave);
    const uint32_t c1 = gcodes[w1];
    const int r = wave_first_true(w1, n - 1, [&](int t) { return delta8(c1, gcodes[t + 1]) < ov; });
    if (lane == o) nsv = r;
  }
  // min(a[t .. w0 - 1]) >= v, i.e. delta8(code[t], code[w0]) >= v; psv is one before
  for (uint64_t m = ballot(far_p); m; m &= m - 1) {
    const int o = __ffsll(static_cast<long long>(m)) - 1;
    const int ov = __shfl(static_cast<int>(v), o, kWave);
    const uint32_t c0 = gcodes[w0];
    const int r = wave_first_true(0, w0 - 1, [&](int t) { return delta8(gcodes[t], c0) >= ov; });
    if (lane == o) psv = r - 1;
  }
  if (far) emit(false);
}

void launch_radix_tree(const TreeArgs& a, SegArgs seg, uint32_t n_index, hipStream_t stream) {
  const uint32_t n_tiles = (n_index + kTreeBlock - 1) / kTreeBlock;
  if (n_tiles == 0) return;
  seg.count_sub = 1;
  hipLaunchKernelGGL(k_radix_tree, dim3(xcd_grid(n_tiles)), dim3(kTreeBlock), 0, stream, a, seg,
                     n_tiles);
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
// a one-workgroup scan there; D4).  Single pass with decoupled look-back;
// the segment's last tile writes the total (n_octree_nodes).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kScanBlock) void k_prefix_sum(const int32_t* __restrict__ in,
                                                           int32_t* __restrict__ out, SegArgs seg,
                                                           uint32_t* __restrict__ total,
                                                           uint64_t* __restrict__ lookback,
                                                           uint32_t* __restrict__ ctrl) {
  constexpr int G = kScanKpt / 4;
  __shared__ uint32_t s_w[kScanBlock / kWave];
  __shared__ uint32_t s_bcast;
  __shared__ Grab s_grab;
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
    const uint32_t first = tile * kScanTile;
    const Seg sg = seg_of(seg, first);
    const uint32_t lim = sg.base + sg.count;
    if (first >= lim) {
      if (tid == 0 && first == sg.base) total[sg.id] = 0;
      return;
    }
    const uint32_t tile_in_seg = (first - sg.base) / kScanTile;
    const uint32_t wbase = first + wave * (kScanKpt * kWave);
    uint32_t v[G][4];
    uint32_t gsum[G];
    const bool full = first + kScanTile <= lim;
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
    const uint32_t wexcl = block_excl_scan(lane == 0 ? wave_run : 0u, s_w, &tile_total);
    const uint32_t wave_off = __shfl(wexcl, 0, kWave);
    const uint32_t excl = tile_prefix(tile_total, tile, tile_in_seg, epoch, lookback, &s_bcast);
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
    if (tid == 0 && lim - first <= kScanTile) total[sg.id] = excl + tile_total;
  }
}

void launch_prefix_sum(const int32_t* in, int32_t* out, SegArgs seg, uint32_t n_index,
                       uint32_t* total, uint64_t* lookback, uint32_t* ctrl, hipStream_t stream) {
  const uint32_t n_tiles = (n_index + kScanTile - 1) / kScanTile;
  if (n_tiles == 0) return;
  seg.count_sub = 1;
  hipLaunchKernelGGL(k_prefix_sum, dim3(n_tiles), dim3(kScanBlock), 0, stream, in, out, seg,
                     total, lookback, ctrl);
}

}  // namespace rw
