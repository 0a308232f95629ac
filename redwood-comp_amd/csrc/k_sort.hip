// Stage 2: LSD radix sort of 32-bit keys as an 8-bit-digit onesweep.
//
// What it replaces: the reference's CPU binning passes
// (apps/tree/host/func_sort.cpp:117-211) and its CUDA onesweep
// (apps/tree/cuda/cu_dispatcher.cu:82-468).  The design is the published
// Onesweep scheme (Adinets & Merrill 2022) with decoupled look-back
// (Merrill & Garland 2016), laid out for gfx950:
//   * one sweep over the keys builds the digit histograms of every pass
//     (16-byte loads, LDS histograms, spread global atomics); for a
//     single-cloud sort it also counts pass 0's digits per tile;
//   * a small kernel turns them into per-segment exclusive digit starts;
//   * one launch per digit pass: a workgroup takes the next tile from an atomic
//     counter, counts its digits and publishes them for its successors, looks
//     back over its predecessors, ranks its keys (64-lane ballots: a stable
//     wave-level multisplit), stages the tile in LDS in digit order, then
//     writes each digit run to its final place (coalesced along the run);
//   * pass 0 needs no stability (the keys carry no order yet), so it ranks
//     with LDS atomics, and for a single cloud it takes its tile prefixes from
//     a scan of the sweep's tile counts instead of the look-back.
// Segments (clouds) are sorted independently in one launch: tiles never cross
// a segment boundary and the look-back stops at the segment's first tile.
#include <type_traits>

#include "rw_device.hpp"
#include "rw_internal.hpp"

namespace rw {

// ---------------------------------------------------------------------------
// Histogram sweep.  Workgroup u reads histogram unit u (unit_keys keys, within
// one segment) with uint4 loads and counts the digits of all passes in LDS;
// the unit's counts are added to copy (u % kHistCopies) of its segment's
// histogram, which spreads the global atomics over many cache lines.
// ---------------------------------------------------------------------------
// Pass 0's digit counts of the sort tiles a histogram unit covers (the unit
// starts at a multiple of the tile size), from LDS to the digit-major table
// tile_counts[digit][tile] of a sort of n keys.
__device__ __forceinline__ void store_tile_counts(const uint32_t* s_tc, uint32_t first, uint32_t end,
                                                  uint32_t n, uint32_t tile_shift, uint32_t* tile_counts) {
  const uint32_t n_tiles = ((n - 1) >> tile_shift) + 1;
  const uint32_t t0 = first >> tile_shift;
  const uint32_t nt = ((end - first - 1) >> tile_shift) + 1;
  for (uint32_t i = threadIdx.x; i < kTilesPerHistUnit * 256; i += kHistBlock) {
    const uint32_t d = i / kTilesPerHistUnit, j = i % kTilesPerHistUnit;
    if (j < nt) tile_counts[static_cast<size_t>(d) * n_tiles + t0 + j] = s_tc[j * 256 + d];
  }
}

// Pass 0's unit histogram as the sum of its tiles' counts (tiles past the
// unit's keys hold zeros).  Ends with a barrier.
__device__ __forceinline__ void finish_pass0(const uint32_t* s_tc, uint32_t* s_h) {
  if (threadIdx.x < 256) {
    uint32_t c = 0;
#pragma unroll
    for (uint32_t t = 0; t < kTilesPerHistUnit; ++t) c += s_tc[t * 256 + threadIdx.x];
    s_h[threadIdx.x] = c;
  }
  __syncthreads();
}

__global__ __launch_bounds__(kHistBlock) void k_digit_hist(const uint32_t* __restrict__ keys,
                                                           SegArgs seg, uint32_t unit_keys,
                                                           uint32_t n_units, DigitPlan plan,
                                                           uint32_t* __restrict__ hist,
                                                           uint32_t tile_shift,
                                                           uint32_t* __restrict__ tile_counts,
                                                           const uint32_t* __restrict__ skip) {
  if (skip && *skip) return;  // the bucket sort ran instead
  // the digit of pass p is taken exactly as the binning pass takes it, so
  // the counts match the ranks whatever the key values
  const uint32_t n_passes = plan.n_passes;
  uint32_t mask[kMaxPasses], shift[kMaxPasses];
#pragma unroll
  for (uint32_t p = 0; p < kMaxPasses; ++p) {
    mask[p] = p < n_passes ? (1u << plan.bits[p]) - 1u : 0u;
    shift[p] = p < n_passes ? plan.shift[p] : 0u;
  }
  __shared__ uint32_t s_h[kMaxPasses * 256];
  __shared__ uint32_t s_tc[kTilesPerHistUnit * 256];  // pass 0 digit counts per sort tile
  // With tile counts requested, pass 0 is counted per tile only and its unit
  // histogram is summed from the tiles afterwards.
  const bool per_tile = tile_counts != nullptr;
  for (uint32_t i = threadIdx.x; i < kMaxPasses * 256; i += kHistBlock) s_h[i] = 0;
  if (per_tile)
    for (uint32_t i = threadIdx.x; i < kTilesPerHistUnit * 256; i += kHistBlock) s_tc[i] = 0;
  __syncthreads();
  const uint32_t u = blockIdx.x;
  const uint32_t first = u * unit_keys;
  const Seg sg = seg_of(seg, first);
  const uint32_t lim = sg.base + sg.count;
  if (first < lim) {
    const uint32_t end = min(first + unit_keys, lim);
    // aligned uint4 body (first is a multiple of 4), scalar tail
    const uint32_t body = first + ((end - first) & ~3u);
    const uint4* v4 = reinterpret_cast<const uint4*>(keys + first);
    const uint32_t n4 = (body - first) >> 2;
    // four 16-byte loads in flight per thread before their counts
    constexpr uint32_t kDepth = 4;
    for (uint32_t i0 = threadIdx.x; i0 < n4; i0 += kDepth * kHistBlock) {
      uint4 q[kDepth];
#pragma unroll
      for (uint32_t r = 0; r < kDepth; ++r)
        if (i0 + r * kHistBlock < n4) q[r] = v4[i0 + r * kHistBlock];
#pragma unroll
      for (uint32_t r = 0; r < kDepth; ++r) {
        if (i0 + r * kHistBlock >= n4) break;
        const uint32_t k[4] = {q[r].x, q[r].y, q[r].z, q[r].w};
        // the four keys share a tile (tiles hold a multiple of four keys)
        const uint32_t tc_row = ((4 * (i0 + r * kHistBlock)) >> tile_shift) * 256;
        // keys that are ordered, or share their low bits, show as a top or
        // a low digit common to the whole wave
        const bool shared = wave_same(k[0] >> shift[n_passes - 1]) || wave_same(k[0] & mask[0]);
        // one branch per group of keys; the plain counts stay a tight loop
        const auto count4 = [&](auto check) {
#pragma unroll
          for (int c = 0; c < 4; ++c) {
#pragma unroll
            for (uint32_t p = 0; p < kMaxPasses; ++p)
              if (p < n_passes && (p > 0 || !per_tile))
                hist_add(s_h, p * 256 + ((k[c] >> shift[p]) & mask[p]), decltype(check)::value);
            if (per_tile) hist_add(s_tc, tc_row + (k[c] & mask[0]), decltype(check)::value);
          }
        };
        if (shared) count4(std::true_type{});
        else count4(std::false_type{});
      }
    }
    for (uint32_t i = body + threadIdx.x; i < end; i += kHistBlock) {
      const uint32_t k = keys[i];
      for (uint32_t p = per_tile ? 1 : 0; p < n_passes; ++p)
        hist_add(s_h, p * 256 + ((k >> shift[p]) & mask[p]), false);
      if (per_tile) hist_add(s_tc, ((i - first) >> tile_shift) * 256 + (k & mask[0]), false);
    }
  }
  __syncthreads();
  if (per_tile) {
    if (first < lim) store_tile_counts(s_tc, first, min(first + unit_keys, lim), lim, tile_shift, tile_counts);
    finish_pass0(s_tc, s_h);
  }
  uint32_t* dst = hist + (static_cast<size_t>(sg.id) * kHistCopies + (u % kHistCopies)) *
                             (kMaxPasses * 256);
  for (uint32_t i = threadIdx.x; i < n_passes * 256; i += kHistBlock) {
    const uint32_t c = s_h[i];
    if (c) atomicAdd(&dst[i], c);
  }
}

// ---------------------------------------------------------------------------
// Histogram sweep fused with stage 1, for a whole build (one cloud, or every
// cloud of a batch): each workgroup encodes a unit of points of one cloud
// (16-byte nontemporal float4 loads, four in flight per thread), writes the
// codes and counts their digits in LDS into the cloud's histogram copies as
// k_digit_hist does, so the sort does not read the codes an extra time.  The
// first unit of each cloud clears its status word (the build's first kernel).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kHistBlock) void k_morton_hist(const float4* __restrict__ pts,
                                                            uint32_t* __restrict__ codes, SegArgs seg,
                                                            Quant quant,
                                                            uint32_t unit_keys, DigitPlan plan,
                                                            uint32_t* __restrict__ hist,
                                                            uint32_t tile_shift,
                                                            uint32_t* __restrict__ tile_counts,
                                                            uint32_t* __restrict__ status) {
  const uint32_t n_passes = plan.n_passes;
  uint32_t mask[kMaxPasses], shift[kMaxPasses];
#pragma unroll
  for (uint32_t p = 0; p < kMaxPasses; ++p) {
    mask[p] = p < n_passes ? (1u << plan.bits[p]) - 1u : 0u;
    shift[p] = p < n_passes ? plan.shift[p] : 0u;
  }
  __shared__ uint32_t s_h[kMaxPasses * 256];
  __shared__ uint32_t s_tc[kTilesPerHistUnit * 256];  // pass 0 digit counts per sort tile
  // With tile counts requested, pass 0 is counted per tile only and its unit
  // histogram is summed from the tiles afterwards.
  const bool per_tile = tile_counts != nullptr;
  for (uint32_t i = threadIdx.x; i < kMaxPasses * 256; i += kHistBlock) s_h[i] = 0;
  if (per_tile)
    for (uint32_t i = threadIdx.x; i < kTilesPerHistUnit * 256; i += kHistBlock) s_tc[i] = 0;
  const uint32_t u = blockIdx.x;
  const uint32_t first = u * unit_keys;
  const Seg sg = seg_of(seg, first);
  const uint32_t lim = sg.base + sg.count;
  if (threadIdx.x == 0 && status && first == sg.base) status[sg.id] = 0;  // the build's first kernel
  __syncthreads();
  const uint32_t end = min(first + unit_keys, lim);
  constexpr uint32_t kDepth = 4;
  for (uint32_t i0 = first + threadIdx.x; i0 < end; i0 += kDepth * kHistBlock) {
    float4 q[kDepth];
#pragma unroll
    for (uint32_t r = 0; r < kDepth; ++r)
      if (i0 + r * kHistBlock < end) q[r] = load_stream(pts + i0 + r * kHistBlock);
#pragma unroll
    for (uint32_t r = 0; r < kDepth; ++r) {
      const uint32_t i = i0 + r * kHistBlock;
      if (i >= end) break;
      const uint32_t k = morton_encode(q[r].x, q[r].y, q[r].z, quant);
      codes[i] = k;
      // points that are ordered, or share their low bits, show as a top or a
      // low digit common to the whole wave
      const bool shared = wave_same(k >> shift[n_passes - 1]) || wave_same(k & mask[0]);
#pragma unroll
      for (uint32_t p = 0; p < kMaxPasses; ++p)
        if (p < n_passes && (p > 0 || !per_tile))
          hist_add(s_h, p * 256 + ((k >> shift[p]) & mask[p]), shared);
      if (per_tile) hist_add(s_tc, ((i - first) >> tile_shift) * 256 + (k & mask[0]), shared);
    }
  }
  __syncthreads();
  if (per_tile) {
    if (first < end) store_tile_counts(s_tc, first, end, lim, tile_shift, tile_counts);
    finish_pass0(s_tc, s_h);
  }
  uint32_t* dst = hist + (static_cast<size_t>(sg.id) * kHistCopies + (u % kHistCopies)) *
                             (kMaxPasses * 256);
  for (uint32_t i = threadIdx.x; i < n_passes * 256; i += kHistBlock) {
    const uint32_t c = s_h[i];
    if (c) atomicAdd(&dst[i], c);
  }
}

// Pass 0 runs from the sweep's tile counts (no look-back) for a single cloud
// when its digit has a table kernel (6 to 8 bits) and another pass follows.
constexpr bool table_pass0(const DigitPlan& plan) {
  return plan.n_passes >= 2 && plan.bits[0] >= 6 && plan.bits[0] <= 8;
}

void launch_morton_hist(const float4* pts, uint32_t* codes, SegArgs seg, uint32_t n_index, uint32_t unit_keys,
                        float min_coord, float range, const DigitPlan& plan, uint32_t* hist, int sort_variant,
                        uint32_t* tile_counts, uint32_t* status, hipStream_t stream) {
  if (n_index == 0) return;
  const uint32_t units = (n_index + unit_keys - 1) / unit_keys;
  if (!table_pass0(plan) || seg.n_segs != 1) tile_counts = nullptr;
  hipLaunchKernelGGL(k_morton_hist, dim3(units), dim3(kHistBlock), 0, stream, pts, codes, seg,
                     make_quant(min_coord, range), unit_keys, plan, hist, sort_tile_shift(sort_variant),
                     tile_counts, status);
}

// Pass 0's exclusive digit prefix of every tile of a single-cloud sort, in
// place of the tile counts: one workgroup of 1024 threads scans one digit's
// row of the digit-major table, 1024 tiles per step (coalesced), carrying
// the total.  s_w holds 16 words.
__device__ __forceinline__ void tile_scan_row(uint32_t* tile_counts, uint32_t n_tiles, uint32_t digit,
                                              uint32_t* s_w) {
  const uint32_t t = threadIdx.x;
  uint32_t* row = tile_counts + static_cast<size_t>(digit) * n_tiles;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < n_tiles; base += 1024) {
    const uint32_t k = base + t;
    const uint32_t c = k < n_tiles ? row[k] : 0u;
    const uint32_t incl = wave_incl_scan(c);
    if (t % kWave == kWave - 1) s_w[t / kWave] = incl;
    __syncthreads();
    uint32_t off = 0, total = 0;
#pragma unroll
    for (uint32_t w = 0; w < 1024 / kWave; ++w) {
      const uint32_t x = s_w[w];
      off += (w < t / kWave) ? x : 0u;
      total += x;
    }
    if (k < n_tiles) row[k] = carry + off + incl - c;
    carry += total;
    __syncthreads();  // s_w is rewritten by the next step
  }
}

// Block-wide exclusive scan of one value per thread for the first 256 threads
// (4 waves); s_tmp holds 4 words.  Returns the exclusive prefix; *total gets
// the sum.  All threads of the block must call it.
__device__ __forceinline__ uint32_t scan256(uint32_t v, uint32_t* s_tmp, uint32_t* total) {
  const uint32_t t = threadIdx.x;
  const uint32_t w = t >> 6;
  const uint32_t incl = wave_incl_scan(v);
  if (t < 256 && (t & 63u) == 63u) s_tmp[w] = incl;
  __syncthreads();
  uint32_t off = 0, sum = 0;
#pragma unroll
  for (uint32_t i = 0; i < 4; ++i) {
    const uint32_t x = s_tmp[i];
    off += (i < w) ? x : 0u;
    sum += x;
  }
  __syncthreads();
  *total = sum;
  return off + incl - v;
}

// Digit starts of segment s, pass p: sums the histogram copies, clears them
// for the next sort, and writes the exclusive scan over digits (thread d <
// 256 owns digit d; every thread of the block calls it).
__device__ __forceinline__ void digit_starts(uint32_t* hist, uint32_t* digit_start, uint32_t n_passes,
                                             uint32_t s, uint32_t p, uint32_t* s_tmp) {
  const uint32_t d = threadIdx.x;
  uint32_t c = 0;
  if (d < 256) {
#pragma unroll
    for (uint32_t k = 0; k < kHistCopies; ++k) {
      uint32_t* h = hist + (static_cast<size_t>(s) * kHistCopies + k) * (kMaxPasses * 256) + p * 256 + d;
      c += *h;
      *h = 0;
    }
  }
  uint32_t total;
  const uint32_t ex = scan256(c, s_tmp, &total);
  if (d < 256) digit_start[(static_cast<size_t>(s) * n_passes + p) * 256 + d] = ex;
}

// Clears the histogram copies of segment s, pass p (thread d < 256: digit d).
__device__ __forceinline__ void clear_counts(uint32_t* hist, uint32_t s, uint32_t p) {
  const uint32_t d = threadIdx.x;
  if (d < 256) {
#pragma unroll
    for (uint32_t k = 0; k < kHistCopies; ++k)
      hist[(static_cast<size_t>(s) * kHistCopies + k) * (kMaxPasses * 256) + p * 256 + d] = 0;
  }
}

// Digit starts: one workgroup per (segment, pass).
__global__ __launch_bounds__(256) void k_digit_scan(uint32_t* __restrict__ hist,
                                                    uint32_t* __restrict__ digit_start,
                                                    uint32_t n_passes, const uint32_t* __restrict__ skip) {
  if (skip && *skip) {  // the bucket sort ran: the sweep's counts are only cleared
    clear_counts(hist, blockIdx.x, blockIdx.y);
    return;
  }
  __shared__ uint32_t s_tmp[4];
  digit_starts(hist, digit_start, n_passes, blockIdx.x, blockIdx.y, s_tmp);
}

// A single-cloud sort's two small scans in one launch: workgroups 0..255
// each scan one digit's row of the pass-0 tile table, the next n_passes
// workgroups compute the digit starts of one pass each.
__global__ __launch_bounds__(1024) void k_start_scans(uint32_t* __restrict__ hist,
                                                      uint32_t* __restrict__ digit_start,
                                                      uint32_t n_passes,
                                                      uint32_t* __restrict__ tile_counts,
                                                      uint32_t n_tiles, const uint32_t* __restrict__ skip) {
  if (skip && *skip) {  // the bucket sort ran: the sweep's counts are only cleared
    if (blockIdx.x >= 256) clear_counts(hist, 0, blockIdx.x - 256);
    return;
  }
  __shared__ uint32_t s_tmp[1024 / kWave];
  if (blockIdx.x < 256) {
    tile_scan_row(tile_counts, n_tiles, blockIdx.x, s_tmp);
  } else {
    digit_starts(hist, digit_start, n_passes, 0, blockIdx.x - 256, s_tmp);
  }
}

// ---------------------------------------------------------------------------
// One onesweep digit pass.
// Keys are loaded wave-striped (key k of lane l of wave w sits at
// tile + w*KPT*64 + k*64 + l), which is the order the ranking walks, so every
// pass after the first is stable.
// ---------------------------------------------------------------------------
// kFirst (pass 0): the keys carry no order from earlier passes, and equal
// digits of a keys-only sort need none, so ranks within a wave come from LDS
// atomics instead of the stable multisplit.
// kTable (pass 0 of a single-cloud sort): the tile is blockIdx.x and its
// digit prefixes come from tile_excl, scanned from the histogram sweep's tile
// counts, so the pass has no look-back chain.
template <int BLOCK, int KPT, int BITS, bool kTable = false, bool kFirst = kTable>
__global__ __launch_bounds__(BLOCK) void k_onesweep(
    const uint32_t* __restrict__ in, uint32_t* __restrict__ out, SegArgs seg,
    const uint32_t* __restrict__ digit_start, uint32_t pass,
    uint32_t n_passes, uint32_t shift, uint64_t* __restrict__ lookback, uint32_t* __restrict__ ctrl,
    const uint32_t* __restrict__ tile_excl, uint32_t n_tiles, const uint32_t* __restrict__ skip) {
  if (skip && *skip) return;  // before the grab: a skipped pass leaves the counter as it was
  constexpr int WAVES = BLOCK / kWave;
  constexpr uint32_t TILE = BLOCK * KPT;
  static_assert(BLOCK >= 256 && BLOCK % kWave == 0, "one thread per digit needed");
  // a tile never straddles two clouds: it divides a chunk, or (a multiple of
  // the chunk) it is only launched for a single cloud
  static_assert(kChunk % TILE == 0 || TILE % kChunk == 0, "tile and chunk must nest");
  static_assert(!kTable || kFirst, "the tile table serves pass 0");
  constexpr uint32_t dmask = (1u << BITS) - 1u;
  // digits of this pass: the per-digit work (counts, starts, look-back,
  // scans) runs on threads tid < RADIX only (64 for a 6-bit pass)
  constexpr uint32_t RADIX = 1u << BITS;

  __shared__ uint32_t s_stage[TILE];
  __shared__ uint32_t s_whist[WAVES][256];
  __shared__ uint32_t s_gbase[256];
  __shared__ uint32_t s_count[256];  // the tile's digit counts (valid keys)
  __shared__ uint32_t s_pre[256];    // the tile's digit starts (+ prefixes with kTable)
  __shared__ uint32_t s_tmp[4];
  __shared__ Grab s_grab;  // the tile taken next

  const uint32_t tid = threadIdx.x;
  const uint32_t wave = tid / kWave;
  const uint32_t lane = tid % kWave;

  // A workgroup sorts tiles one after the other.  With kTable it walks
  // blockIdx.x, + gridDim.x, ... (no tile waits on another).  Otherwise it
  // takes each tile from the counter, the next one while the current one is
  // staged, and its last grab finds none: a tile only ever waits on tiles
  // taken earlier, by running workgroups that finish their earlier tiles
  // first, so the look-back spin ends whatever the residency of the grid.
  // The next tile's keys load while the current one is scattered.
  const uint32_t n_grabs = n_tiles + gridDim.x;
  if (!kTable && tid == 0) s_grab = grab_tiles(ctrl, n_grabs);
  for (uint32_t i = tid; i < WAVES * 256; i += BLOCK) (&s_whist[0][0])[i] = 0;
  if (tid < RADIX) s_count[tid] = 0;
  __syncthreads();
  const uint32_t epoch = kTable ? 0u : s_grab.epoch;
  uint32_t tile = kTable ? blockIdx.x : s_grab.index;

  uint32_t keys[KPT];
  uint32_t rank[KPT];
  uint32_t* wh = s_whist[wave];  // this wave's digit counters
  // The segment of the tile whose keys are loaded, kept for its iteration:
  // looked up there, after the scatter, its read would wait for every store
  // of the scatter as well (one in-order counter for loads and stores).
  // So are the digit starts of its segment and, with kTable, the tile's
  // prefixes (thread d < 256: digit d), loaded before the keys and moved to
  // s_pre where the keys are first used.  A kTable sort has one segment,
  // found once.
  Seg g = {};
  uint32_t tdig = 0, texcl = 0;
  const Seg g_table = kTable ? seg_of(seg, 0) : Seg{};
  const auto load_keys = [&](uint32_t t) {
    const uint32_t f = t * TILE;
    if (kTable && tid < RADIX) texcl = tile_excl[static_cast<size_t>(tid) * n_tiles + t];
    g = kTable ? g_table : seg_of(seg, f);
    if (tid < RADIX) tdig = digit_start[(static_cast<size_t>(g.id) * n_passes + pass) * 256 + tid];
    const uint32_t l = g.base + g.count;
    if (f >= l) return;  // a segment's padding
    const uint32_t wb = f + wave * (KPT * kWave) + lane;
    if (f + TILE <= l) {
      const uint32_t* src = in + wb;  // one base, immediate offsets per key
#pragma unroll
      for (int k = 0; k < KPT; ++k) keys[k] = src[k * kWave];
    } else {
#pragma unroll
      for (int k = 0; k < KPT; ++k) {
        const uint32_t idx = wb + k * kWave;
        keys[k] = (idx < l) ? in[idx] : 0xffffffffu;  // padding ranks last in the top digit
      }
    }
  };
  // The digit counts of the loaded tile (LDS atomics into s_count), taken
  // before its iteration so that the aggregate is published first thing.
  // For every tile after the first this runs behind the previous tile's
  // scatter, in straight-line code: the wait for the keys then leaves the
  // scatter's stores in flight.
  const auto count_keys = [&](uint32_t t) {
    if (tid < RADIX) s_pre[tid] = tdig;
    const uint32_t f = t * TILE;
    const uint32_t l = g.base + g.count;
    if (f >= l) {  // a segment's padding: no keys were loaded
      __builtin_amdgcn_s_waitcnt(0);
      return;
    }
    const uint32_t wb = f + wave * (KPT * kWave) + lane;
    // a wave whose first keys share their digit may hold ordered keys
    const bool shared = wave_same((keys[0] >> shift) & dmask);
    if (f + TILE <= l) {
#pragma unroll
      for (int k = 0; k < KPT; ++k) hist_add(s_count, (keys[k] >> shift) & dmask, shared);
    } else {
      // the last tile of a segment.  Every key is read here (a select that
      // changes nothing): on no path does a key load stay in flight into the
      // iteration, where waiting for it would also wait for the stores
#pragma unroll
      for (int k = 0; k < KPT; ++k) {
        const bool in_seg = wb + k * kWave < l;
        keys[k] = in_seg ? keys[k] : 0xffffffffu;
        if (in_seg) hist_add(s_count, (keys[k] >> shift) & dmask, shared);
      }
    }
  };
  // Pass 0 ranks, per digit counter of the wave (LDS atomics; padding keys
  // take no rank: they are not staged).  With kTable they are taken before
  // the iteration, like the counts above.  One branch per tile: the plain
  // ranks stay a tight loop.
  const auto rank_first = [&](uint32_t t) {
    if (kTable && tid < RADIX) s_pre[tid] = tdig + texcl;
    const uint32_t f = t * TILE;
    const uint32_t l = g.base + g.count;
    if (f >= l) return;  // a segment's padding
    const uint32_t wb = f + wave * (KPT * kWave) + lane;
    const auto ranks = [&](auto check) {
      if (f + TILE <= l) {
#pragma unroll
        for (int k = 0; k < KPT; ++k) rank[k] = rank_add(wh, (keys[k] >> shift) & dmask, decltype(check)::value);
      } else {
        // every key read, as in count_keys
#pragma unroll
        for (int k = 0; k < KPT; ++k) {
          const bool in_seg = wb + k * kWave < l;
          keys[k] = in_seg ? keys[k] : 0xffffffffu;
          rank[k] = 0;
          if (in_seg) rank[k] = rank_add(wh, (keys[k] >> shift) & dmask, decltype(check)::value);
        }
      }
    };
    // a wave whose first keys share their digit may hold ordered keys
    if (wave_same((keys[0] >> shift) & dmask)) ranks(std::true_type{});
    else ranks(std::false_type{});
  };
  if (tile < n_tiles) {
    load_keys(tile);
    if (!kTable) {
      count_keys(tile);
      __syncthreads();
    } else {
      rank_first(tile);
    }
  }

  while (tile < n_tiles) {
    const uint32_t first = tile * TILE;
    const Seg sg = g;
    const uint32_t lim = sg.base + sg.count;
    const bool live = first < lim;  // else a segment's padding (uniform across the workgroup)
    const uint32_t n_valid = live ? min(TILE, lim - first) : 0u;
    const uint32_t wbase = first + wave * (KPT * kWave) + lane;
    if (live) {
      const uint32_t tile_in_seg = (first - sg.base) / TILE;
      // The tile's digit counts (taken before the iteration) are published,
      // and its own look-back done, before the ranking: successors waiting
      // on this tile are released early.
      uint32_t excl = 0;  // this digit's exclusive prefix over the segment's earlier tiles (look-back)
      if (!kTable && tid < RADIX) {
        const uint32_t d = tid;
        const uint32_t cnt_valid = s_count[d];
        s_count[d] = 0;  // for the next tile's counts (after two barriers)
        uint64_t* slot = &lookback[static_cast<size_t>(tile) * 256 + d];
        if (tile_in_seg == 0) {
          lb_store(slot, lb_pack(epoch, kIncl, cnt_valid));
        } else {
          lb_store(slot, lb_pack(epoch, kAgg, cnt_valid));
          excl = lb_exclusive<kLookbackWindow>(lookback + d, 256, tile, tile_in_seg, epoch);
          lb_store(slot, lb_pack(epoch, kIncl, excl + cnt_valid));
        }
      }

      // Ranks within the wave, per digit counter of the wave.
      if (kFirst && !kTable) rank_first(tile);
#pragma unroll
      for (int k = 0; k < KPT; ++k) {
        const uint32_t d = (keys[k] >> shift) & dmask;
        if (kFirst) continue;
        // Stable multisplit: the lowest lane of each group of lanes holding
        // the same digit advances the wave's counter for that digit.
        const uint64_t m = match_peers<BITS>(d);  // peers, this lane included
        const uint32_t below = popc_below(m);
        const uint32_t old = wh[d];
        if (below == 0) wh[d] = old + static_cast<uint32_t>(__popcll(m));
        rank[k] = old + below;
      }
      __syncthreads();

      // Per digit: the tile's digit starts, then each wave's start of the
      // digit in the staged tile (stable ranks count padding keys, staged
      // last and never written out; pass 0 does not count them).
      uint32_t cnt = 0;
      if (tid < RADIX) {
#pragma unroll
        for (int w = 0; w < WAVES; ++w) cnt += s_whist[w][tid];
      }
      uint32_t total;
      const uint32_t tstart = scan256(cnt, s_tmp, &total);
      if (tid < RADIX) {
        uint32_t run = tstart;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
          const uint32_t t = s_whist[w][tid];
          s_whist[w][tid] = run;
          run += t;
        }
        s_gbase[tid] = sg.base + s_pre[tid] + excl - tstart;
      }
      __syncthreads();

      // Stage the tile in digit order.
#pragma unroll
      for (int k = 0; k < KPT; ++k) {
        const uint32_t d = (keys[k] >> shift) & dmask;
        if (!kFirst || n_valid == TILE || wbase + k * kWave < lim) s_stage[wh[d] + rank[k]] = keys[k];
      }
      // from here on only this wave uses its counters: cleared for the next tile
      for (uint32_t i = lane; i < RADIX; i += kWave) wh[i] = 0;
    }
    // The next tile: taken now, its keys loading during the scatter.
    if (!kTable && tid == 0) s_grab = grab_tiles(ctrl, n_grabs);
    __syncthreads();
    const uint32_t next = kTable ? tile + gridDim.x : s_grab.index;
    if (next < n_tiles) load_keys(next);

    // Scatter: consecutive staged keys of one digit go to consecutive
    // addresses.  A whole tile's loop is unrolled, so the next tile's first
    // key waits only for its own load, not for these stores.
    const auto put = [&](uint32_t p) {
      const uint32_t key = s_stage[p];
      out[s_gbase[(key >> shift) & dmask] + p] = key;
    };
    // (the next tile's counts, or pass 0 ranks, follow in each branch: after
    // a merge the wait for its keys would have to assume that no store is in
    // flight)
    if (n_valid == TILE) {
#pragma unroll
      for (int k = 0; k < KPT; ++k) put(tid + k * BLOCK);
      if (!kTable && next < n_tiles) count_keys(next);
      if (kTable && next < n_tiles) rank_first(next);
    } else {
      for (uint32_t p = tid; p < n_valid; p += BLOCK) put(p);
      if (!kTable && next < n_tiles) count_keys(next);
      if (kTable && next < n_tiles) rank_first(next);
    }
    __syncthreads();  // the staged tile and the counters are reused
    tile = next;
  }
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
// Calls f with the variant index as a compile-time constant.
template <class F>
void with_variant(int v, F&& f) {
  static_assert(kNumSortVariants == 6, "one case per variant");
  switch (v) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 5: f(std::integral_constant<int, 5>{}); break;
    default: f(std::integral_constant<int, 0>{}); break;
  }
}

// Calls f with the digit width b (1..8) as a compile-time constant.
template <class F>
void with_bits(uint32_t b, F&& f) {
  switch (b) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 5: f(std::integral_constant<int, 5>{}); break;
    case 6: f(std::integral_constant<int, 6>{}); break;
    case 7: f(std::integral_constant<int, 7>{}); break;
    default: f(std::integral_constant<int, 8>{}); break;
  }
}
// The same for pass 0's own kernels, compiled for 6..8-bit digits.
template <class F>
void with_first_bits(uint32_t b, F&& f) {
  switch (b) {
    case 6: f(std::integral_constant<int, 6>{}); break;
    case 7: f(std::integral_constant<int, 7>{}); break;
    default: f(std::integral_constant<int, 8>{}); break;
  }
}

// Pass 0 of a single-cloud sort from the tile prefixes (no look-back): the
// scan of the histogram sweep's tile counts, then the pass.
template <int BLOCK, int KPT>
void launch_table_pass(const SortLaunch& L, uint32_t* dst, hipStream_t stream) {
  constexpr uint32_t TILE = BLOCK * KPT;
  const uint32_t n_tiles = (L.n_index + TILE - 1) / TILE;
  with_first_bits(L.plan.bits[0], [&](auto B) {
    constexpr auto kern = k_onesweep<BLOCK, KPT, B, true>;
    hipLaunchKernelGGL(kern, dim3(resident_grid<kern>(BLOCK, n_tiles)), dim3(BLOCK), 0, stream, L.keys_in,
                       dst, L.seg, L.digit_start, 0u, L.plan.n_passes, 0u, L.lookback, L.ctrl, L.tile_counts,
                       n_tiles, L.skip);
  });
}

// Digit passes p_begin .. n_passes-1 on tiles of BLOCK x KPT.  The buffers
// alternate in -> tmp -> out -> tmp -> out: an even number of passes ends in
// out, an odd count starts with in -> out.
template <int BLOCK, int KPT>
void launch_passes(const SortLaunch& L, uint32_t p_begin, hipStream_t stream) {
  constexpr uint32_t TILE = BLOCK * KPT;
  static_assert(TILE >= kSortMinTile, "look-back table sized for kSortMinTile");
  const uint32_t n_passes = L.plan.n_passes;
  const uint32_t n_tiles = (L.n_index + TILE - 1) / TILE;
  const auto dst_of = [&](uint32_t p) { return ((n_passes - p) % 2) == 1 ? L.keys_out : L.keys_tmp; };
  const uint32_t* src = p_begin == 0 ? L.keys_in : dst_of(p_begin - 1);
  for (uint32_t p = p_begin; p < n_passes; ++p) {
    uint32_t* dst = dst_of(p);
    const uint32_t bits = L.plan.bits[p], shift = L.plan.shift[p];
    if (p == 0 && bits >= 6) {  // no order to keep yet: atomic ranks
      with_first_bits(bits, [&](auto B) {
        constexpr auto kern = k_onesweep<BLOCK, KPT, B, false, true>;
        hipLaunchKernelGGL(kern, dim3(resident_grid<kern>(BLOCK, n_tiles)), dim3(BLOCK), 0, stream, src, dst,
                           L.seg, L.digit_start, 0u, n_passes, shift, L.lookback, L.ctrl, nullptr, n_tiles,
                           L.skip);
      });
    } else {
      with_bits(bits, [&](auto B) {
        constexpr auto kern = k_onesweep<BLOCK, KPT, B>;
        hipLaunchKernelGGL(kern, dim3(resident_grid<kern>(BLOCK, n_tiles)), dim3(BLOCK), 0, stream, src, dst,
                           L.seg, L.digit_start, p, n_passes, shift, L.lookback, L.ctrl, nullptr, n_tiles,
                           L.skip);
      });
    }
    src = dst;
  }
}

void launch_sort(const SortLaunch& L, hipStream_t stream) {
  const uint32_t n_passes = L.plan.n_passes;
  const int variant = (L.n_segs > 1 && !sort_tile_fits_chunk(L.variant)) ? kSortChunkVariant : L.variant;
  // pass 0 from tile prefixes: one cloud, whole histogram units, a pass 0
  // with a table kernel
  const bool table = L.tile_counts && L.n_segs == 1 && L.hist_unit_keys == kPlainHistUnit && table_pass0(L.plan);
  if (!L.hist_ready)
    hipLaunchKernelGGL(k_digit_hist, dim3(L.n_hist_units), dim3(kHistBlock), 0, stream, L.keys_in,
                       L.seg, L.hist_unit_keys, L.n_hist_units, L.plan, L.hist,
                       sort_tile_shift(variant), table ? L.tile_counts : nullptr, L.skip);
  if (table) {
    const uint32_t n_tiles = (L.n_index + (1u << sort_tile_shift(variant)) - 1) >> sort_tile_shift(variant);
    hipLaunchKernelGGL(k_start_scans, dim3(256 + n_passes), dim3(1024), 0, stream, L.hist, L.digit_start,
                       n_passes, L.tile_counts, n_tiles, L.skip);
  } else {
    hipLaunchKernelGGL(k_digit_scan, dim3(L.n_segs, n_passes), dim3(256), 0, stream, L.hist,
                       L.digit_start, n_passes, L.skip);
  }
  with_variant(variant, [&](auto V) {
    constexpr int B = kSortVariants[V].block, K = kSortVariants[V].kpt;
    if (table) launch_table_pass<B, K>(L, ((n_passes % 2) == 1) ? L.keys_out : L.keys_tmp, stream);
    launch_passes<B, K>(L, table ? 1u : 0u, stream);
  });
}

}  // namespace rw
