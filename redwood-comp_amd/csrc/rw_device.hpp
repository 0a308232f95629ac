// Device-side building blocks shared by the octree-build kernels (gfx950,
// wave64).  Everything here is header-only and inlined into each kernel TU.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <stdint.h>

namespace rw {

constexpr int kWave = 64;

// Segments.  A batch lays clouds out in one "point index space"; every cloud
// (segment) starts at a multiple of kChunk, so a tile of at most kChunk items
// that starts at a multiple of its size never straddles two clouds, and the
// cloud of any index is chunk_seg[index >> kChunkShift].  A single build is the
// one-segment case with base 0 and no table.
constexpr uint32_t kChunkShift = 13;
constexpr uint32_t kChunk = 1u << kChunkShift;  // 8192

struct SegArgs {
  uint32_t n_segs;            // 1: single segment at base 0 (tables unused)
  uint32_t single_count;      // element count when count == nullptr (single segment)
  const uint32_t* base;       // [n_segs] first index of each segment (nullptr: 0)
  const uint32_t* count;      // [n_segs] live element count per segment (device)
  const uint32_t* chunk_seg;  // [index >> kChunkShift] -> segment
  uint32_t count_sub;         // live count = max(count - count_sub, 0)
};

struct Seg {
  uint32_t id, base, count;
};

__device__ __forceinline__ Seg seg_of(SegArgs a, uint32_t index) {
  Seg s;
  if (a.n_segs == 1) {
    s.id = 0;
    s.base = a.base ? a.base[0] : 0u;
    s.count = a.count ? a.count[0] : a.single_count;
  } else {
    s.id = a.chunk_seg[index >> kChunkShift];
    s.base = a.base[s.id];
    s.count = a.count[s.id];
  }
  s.count = s.count > a.count_sub ? s.count - a.count_sub : 0u;
  return s;
}

// A float4 that is read once (the input points): a nontemporal load, so the
// stream does not displace from the caches the data the next kernels reuse.
__device__ __forceinline__ float4 load_stream(const float4* p) {
  typedef float f4v __attribute__((vector_size(16)));
  const f4v t = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p));
  return make_float4(t[0], t[1], t[2], t[3]);
}

__device__ __forceinline__ uint4 load_stream(const uint4* p) {
  typedef uint32_t u4v __attribute__((vector_size(16)));
  const u4v t = __builtin_nontemporal_load(reinterpret_cast<const u4v*>(p));
  return make_uint4(t[0], t[1], t[2], t[3]);
}
template <class T>
__device__ __forceinline__ T load_stream(const T* p) {
  return __builtin_nontemporal_load(p);
}

// Stores of outputs that no later kernel of the build reads (the octree
// arrays): nontemporal, for the same reason.
__device__ __forceinline__ void store_stream(int4* p, int4 v) {
  typedef int i4v __attribute__((vector_size(16)));
  const i4v t = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(t, reinterpret_cast<i4v*>(p));
}
__device__ __forceinline__ void store_stream(float4* p, float4 v) {
  typedef float f4v __attribute__((vector_size(16)));
  const f4v t = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(t, reinterpret_cast<f4v*>(p));
}
template <class T>
__device__ __forceinline__ void store_stream(T* p, T v) {
  __builtin_nontemporal_store(v, p);
}

// ---------------------------------------------------------------------------
// Wave-level helpers.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Number of set bits of m in lanes below this lane.
__device__ __forceinline__ uint32_t popc_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Inclusive wave scan (sum) of 32-bit values, Kogge-Stone over shuffles.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t t = __shfl_up(v, o, kWave);
    if (lane >= static_cast<uint32_t>(o)) v += t;
  }
  return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// True when every active lane holds the same x (wave-uniform).
__device__ __forceinline__ bool wave_same(uint32_t x) {
  return ballot(x != __builtin_amdgcn_readfirstlane(x)) == 0;
}

// Adds one to h[d] for every active lane.  When all active lanes hit the
// same counter (sorted or clustered keys: consecutive keys share their high
// digits) one lane adds their number instead: 64 atomics on one LDS address
// would otherwise be serialised.  try_wave (wave-uniform) false skips the
// check: the caller saw no digit the wave shares (or only a few keys).
__device__ __forceinline__ void hist_add(uint32_t* h, uint32_t d, bool try_wave = true) {
  if (try_wave && wave_same(d)) {
    const uint64_t act = ballot(true);
    if (lane_id() == static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(act)) - 1))
      atomicAdd(&h[d], static_cast<uint32_t>(__popcll(act)));
  } else {
    atomicAdd(&h[d], 1u);
  }
}

// atomicAdd(&h[d], 1) for every active lane, returning each lane's old
// value (distinct ranks, in no particular order within the wave), with one
// atomic when all active lanes hold the same digit (checked only when
// try_wave, as for hist_add).
__device__ __forceinline__ uint32_t rank_add(uint32_t* h, uint32_t d, bool try_wave) {
  if (try_wave && wave_same(d)) {
    const uint64_t act = ballot(true);
    const uint32_t leader = static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(act)) - 1);
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(&h[d], static_cast<uint32_t>(__popcll(act)));
    return __shfl(base, leader, kWave) + popc_below(act);
  }
  return atomicAdd(&h[d], 1u);
}

// hist_add and rank_add for the lanes with `valid` set, called by every lane
// of the wave (the collectives then never run in divergent code); the rank
// of a lane without `valid` is meaningless.
__device__ __forceinline__ bool wave_same_valid(uint32_t x, bool valid, uint64_t* act, uint32_t* leader) {
  *act = ballot(valid);
  if (*act == 0) return false;
  *leader = static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(*act)) - 1);
  const uint32_t xl = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x), static_cast<int>(*leader)));
  return ballot(valid && x != xl) == 0;
}
__device__ __forceinline__ void hist_add_if(uint32_t* h, uint32_t d, bool valid, bool try_wave) {
  uint64_t act;
  uint32_t leader;
  if (try_wave && wave_same_valid(d, valid, &act, &leader)) {
    if (lane_id() == leader) atomicAdd(&h[d], static_cast<uint32_t>(__popcll(act)));
    return;
  }
  if (valid) atomicAdd(&h[d], 1u);
}
__device__ __forceinline__ uint32_t rank_add_if(uint32_t* h, uint32_t d, bool valid, bool try_wave) {
  uint64_t act;
  uint32_t leader;
  if (try_wave && wave_same_valid(d, valid, &act, &leader)) {
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(&h[d], static_cast<uint32_t>(__popcll(act)));
    return __shfl(base, static_cast<int>(leader), kWave) + popc_below(act);
  }
  return valid ? atomicAdd(&h[d], 1u) : 0u;
}

// Stable multisplit within a wave: the lanes holding the same BITS-bit digit
// d as this lane (this lane included), found with BITS ballots (a lane
// differs from this one where some ballot differs from this lane's own bit).
template <int BITS>
__device__ __forceinline__ uint64_t match_peers(uint32_t d) {
  uint32_t own = static_cast<uint32_t>(__builtin_amdgcn_sbfe(d, 0, 1));
  uint64_t bal = __builtin_amdgcn_uicmp(own, 0u, 33 /* ne */);
  uint32_t xl = own ^ static_cast<uint32_t>(bal);
  uint32_t xh = own ^ static_cast<uint32_t>(bal >> 32);
#pragma unroll
  for (int b = 1; b < BITS; ++b) {
    own = static_cast<uint32_t>(__builtin_amdgcn_sbfe(d, b, 1));
    bal = __builtin_amdgcn_uicmp(own, 0u, 33 /* ne */);
    // truth tables indexed by (x << 2 | own << 1 | bal): 0xF6 is
    // x | (own ^ bal); the last bit takes its complement 0x09, so that
    // x ends as the peer mask
    if (b + 1 < BITS) {
      xl = __builtin_amdgcn_bitop3_b32(xl, own, static_cast<uint32_t>(bal), 0xF6);
      xh = __builtin_amdgcn_bitop3_b32(xh, own, static_cast<uint32_t>(bal >> 32), 0xF6);
    } else {
      xl = __builtin_amdgcn_bitop3_b32(xl, own, static_cast<uint32_t>(bal), 0x09);
      xh = __builtin_amdgcn_bitop3_b32(xh, own, static_cast<uint32_t>(bal >> 32), 0x09);
    }
  }
  if (BITS == 1) {
    xl = ~xl;
    xh = ~xh;
  }
  return (static_cast<uint64_t>(xh) << 32) | xl;
}


// ---------------------------------------------------------------------------
// Decoupled look-back.  A tile publishes a 64-bit word:
//   [63:34] epoch  [33:32] status (1 aggregate, 2 inclusive prefix)  [31:0] value
// The epoch is the launch serial of the kernel using the table, so a table is
// never cleared: words of older launches simply do not match.  The words are
// read and written with device-scope (agent) relaxed atomics, which keeps them
// coherent across the XCDs' separate L2s; the word carries all of its data.
// ---------------------------------------------------------------------------
constexpr uint64_t kAgg = 1, kIncl = 2;

__device__ __forceinline__ uint64_t lb_pack(uint32_t epoch, uint64_t status, uint32_t value) {
  return (static_cast<uint64_t>(epoch) << 34) | (status << 32) | value;
}

__device__ __forceinline__ void lb_store(uint64_t* p, uint64_t w) {
  __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t lb_load(const uint64_t* p) {
  return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool lb_ready(uint64_t w, uint32_t epoch) {
  return static_cast<uint32_t>(w >> 34) == epoch && ((w >> 32) & 3u) != 0;
}

// Exclusive prefix of one lane's column of look-back words: walks back from
// tile-1 over the n_prev predecessor tiles of the same segment, W words at a
// time (independent loads in flight), summing aggregates until an inclusive
// prefix is found.  Every predecessor was handed out by the tile counter
// before this tile, so it is resident and will publish: the spin ends.  The
// segment's first tile always publishes an inclusive prefix.
template <int W>
__device__ __forceinline__ uint32_t lb_exclusive(const uint64_t* table, uint32_t stride,
                                                 uint32_t tile, uint32_t n_prev, uint32_t epoch) {
  uint32_t sum = 0;
  uint32_t k = 0;  // predecessors consumed
  while (k < n_prev) {
    uint64_t w[W];
#pragma unroll
    for (int j = 0; j < W; ++j)
      w[j] = (k + j < n_prev) ? lb_load(table + static_cast<size_t>(tile - 1 - k - j) * stride) : 0;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      if (k + j >= n_prev) return sum;
      const uint64_t* p = table + static_cast<size_t>(tile - 1 - k - j) * stride;
      while (!lb_ready(w[j], epoch)) w[j] = lb_load(p);
      sum += static_cast<uint32_t>(w[j]);
      if (((w[j] >> 32) & 3u) == kIncl) return sum;
    }
    k += W;
  }
  return sum;
}

// The same walk for a single column done by a whole wave: lane l inspects
// predecessor tile-1-(k+l), 64 words per round trip.  All lanes must call it;
// every lane returns the exclusive prefix.
__device__ __forceinline__ uint32_t lb_exclusive_wave(const uint64_t* table, uint32_t tile,
                                                      uint32_t n_prev, uint32_t epoch) {
  const uint32_t lane = lane_id();
  uint32_t sum = 0;
  for (uint32_t k = 0; k < n_prev; k += kWave) {
    const bool live = k + lane < n_prev;
    const uint64_t* p = table + (tile - 1 - k - (live ? lane : 0));
    uint64_t w = live ? lb_load(p) : 0;
    for (;;) {
      const bool ready = !live || lb_ready(w, epoch);
      const uint64_t incl = ballot(live && ready && ((w >> 32) & 3u) == kIncl);
      // lanes that count: up to and including the nearest inclusive prefix
      const uint64_t upto = incl ? (((incl & (~incl + 1)) << 1) - 1) : ~0ull;
      const uint64_t waiting = ballot(!ready) & upto;
      if (waiting == 0) {
        const uint32_t v = (live && ((upto >> lane) & 1u)) ? static_cast<uint32_t>(w) : 0u;
        sum += wave_sum(v);
        if (incl) return sum;
        break;
      }
      if (!ready) w = lb_load(p);
    }
  }
  return sum;
}

// Launch bookkeeping of one look-back user: one 64-bit word, the launch
// epoch in the high half and a grab counter in the low half.  A launch has
// exactly n_grabs workgroups; each takes the next grab (its tile) with one
// atomic add, so tiles are handed out in order to running workgroups and a
// tile's predecessors all belong to workgroups that are already running (the
// look-back spin always ends, however few workgroups fit on the chip).  The epoch comes back with the
// same atomic; the workgroup that takes the last grab stores (epoch + 1, 0),
// which is safe because no other workgroup of the launch touches the word
// again.  Consecutive launches and graph replays need no memset, and no
// second atomic or fence per workgroup is spent on bookkeeping.
constexpr int kCtrlWords = 4;  // 16 bytes per user (8 used)

struct Grab {
  uint32_t index;  // which run of tiles
  uint32_t epoch;  // look-back epoch of this launch, in [1, 2^30)
};

// Thread 0 only.
__device__ __forceinline__ Grab grab_tiles(uint32_t* ctrl, uint32_t n_grabs) {
  unsigned long long* w = reinterpret_cast<unsigned long long*>(ctrl);
  const unsigned long long v = atomicAdd(w, 1ull);
  const uint32_t e = static_cast<uint32_t>(v >> 32);
  Grab g;
  g.index = static_cast<uint32_t>(v);
  g.epoch = (e % 0x3fffffffu) + 1u;
  if (g.index == n_grabs - 1)
    __hip_atomic_store(w, static_cast<unsigned long long>(e + 1u) << 32, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  return g;
}

// ---------------------------------------------------------------------------
// XCD-grouped tile order.  MI355X deals the workgroups of a launch to its 8
// XCDs round robin, and each XCD has its own L2.  Kernels whose neighbouring
// tiles read overlapping input launch a grid padded to a multiple of 8
// (xcd_grid) and map workgroup b to tile (b % 8) * (G / 8) + b / 8, so each
// XCD walks a contiguous run of tiles and the overlap hits its L2.
// ---------------------------------------------------------------------------
constexpr uint32_t kXcds = 8;

__device__ __forceinline__ uint32_t xcd_tile() {
  return (blockIdx.x % kXcds) * (gridDim.x / kXcds) + blockIdx.x / kXcds;
}

// Morton code: 10 bits per axis, x in bit 0 (apps/tree/host/func_morton.hpp:15-39).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__device__ __forceinline__ uint32_t compact3(uint32_t v) {
  v &= 0x09249249u;
  v = (v ^ (v >> 2)) & 0x030c30c3u;
  v = (v ^ (v >> 4)) & 0x0300f00fu;
  v = (v ^ (v >> 8)) & 0x030000ffu;
  v = (v ^ (v >> 16)) & 0x000003ffu;
  return v;
}

// The box of one encode: (v - min) / range.  When range is a power of two
// whose reciprocal is a normal float, inv_range holds that reciprocal and the
// quotient is taken as the product with it: the same real number, so the same
// rounded float as the IEEE division (a result in the subnormal range
// quantises to 0 either way).  Otherwise inv_range is 0 and the encode divides.
struct Quant {
  float min_coord, range, inv_range;
};

inline Quant make_quant(float min_coord, float range) {
  Quant q{min_coord, range, 0.0f};
  int e = 0;
  if (range > 0.0f && range <= 0x1p126f && range >= 0x1p-126f && std::frexp(range, &e) == 0.5f)
    q.inv_range = std::ldexp(1.0f, 1 - e);  // range = 2^(e-1)
  return q;
}

// u32((v - min) / range * 1024): IEEE sub, div, mul (no contraction), then a
// saturating truncation (NaN and negatives -> 0, >= 2^32 -> 2^32-1).
__device__ __forceinline__ uint32_t quantize(float v, const Quant& q) {
  const float d = __fsub_rn(v, q.min_coord);
  const float t = q.inv_range != 0.0f ? __fmul_rn(d, q.inv_range) : __fdiv_rn(d, q.range);
  const float s = __fmul_rn(t, 1024.0f);
  if (!(s > 0.0f)) return 0u;
  if (s >= 4294967296.0f) return 0xffffffffu;
  return static_cast<uint32_t>(s);
}

__device__ __forceinline__ uint32_t morton_encode(float x, float y, float z, const Quant& q) {
  return spread3(quantize(x, q)) | (spread3(quantize(y, q)) << 1) | (spread3(quantize(z, q)) << 2);
}

// Corner of a cell: (float(d) / 1024) * range + min per axis, w = 1.
// The reference divides by 1024.0f; the grid coordinate is an integer in
// [0, 1023], so the quotient is exact and equals the product with 2^-10.
__device__ __forceinline__ float4 morton_decode(uint32_t code, float min_coord, float range) {
  constexpr float kInv1024 = 0x1p-10f;
  float4 r;
  r.x = __fadd_rn(__fmul_rn(static_cast<float>(compact3(code)) * kInv1024, range), min_coord);
  r.y = __fadd_rn(__fmul_rn(static_cast<float>(compact3(code >> 1)) * kInv1024, range), min_coord);
  r.z = __fadd_rn(__fmul_rn(static_cast<float>(compact3(code >> 2)) * kInv1024, range), min_coord);
  r.w = 1.0f;
  return r;
}

}  // namespace rw
