// Stage 2 of a single-cloud build of (2^20, 2^24] points: a bucket sort of
// the 30-bit Morton codes in two passes over HBM.
//
// What it replaces: the reference's binning passes
// (apps/tree/host/func_sort.cpp:117-211) and its CUDA onesweep
// (apps/tree/cuda/cu_dispatcher.cu:82-468), four 8-bit LSD passes, each of
// which reads and writes every key.  A keys-only sort needs stability only
// between LSD passes, so this sort is a hybrid (most significant digit first,
// then least significant first inside each bucket, in the manner of Stehle &
// Jacobsen 2017):
//   * stage 1 (k_morton_bucket) counts each 16384-key tile's codes per bucket
//     of their top D bits, into a tile-major table (one LDS atomic per key:
//     the digit-pass sort's five are left to its own sweep, which runs only
//     when the bucket sort cannot);
//   * k_bucket_scan turns the table into each tile's start inside each
//     bucket (32 buckets per workgroup, a column scan over the tiles) and
//     the buckets' totals;
//   * k_bucket_scatter scans the totals into the buckets' starts and raises
//     the flag that lets the bucket kernels run when every bucket fits
//     kBucketCap, then moves each tile's keys to their buckets: ranks from
//     LDS atomics (the order inside a bucket is free), the tile staged in LDS
//     by bucket, each bucket's run written to consecutive addresses;
//   * k_bucket_local sorts one bucket per workgroup in LDS by its low R bits:
//     an atomic-ranked pass on the low kBucketLowBits bits (the keys carry
//     no order yet), then one stable multisplit pass on the remaining bits,
//     and writes the sorted bucket.
// Equal keys are indistinguishable, so the result is the digit-pass sort's.
// When some bucket exceeds kBucketCap (clustered points) the flag stays 0,
// these kernels return at once and the digit-pass sort, gated on the same
// word, sorts the codes instead.
#include "rw_device.hpp"
#include "rw_internal.hpp"

namespace rw {

// Exclusive scan of one value per thread over a block of BLOCK threads;
// s_w holds BLOCK / 64 words; *total gets the sum.  Every thread of the block
// calls it; it ends with a barrier, so s_w can be reused at once.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_w, uint32_t* total) {
  constexpr uint32_t W = BLOCK / kWave;
  const uint32_t t = threadIdx.x;
  const uint32_t incl = wave_incl_scan(v);
  if (t % kWave == kWave - 1) s_w[t / kWave] = incl;
  __syncthreads();
  uint32_t off = 0, sum = 0;
#pragma unroll
  for (uint32_t w = 0; w < W; ++w) {
    const uint32_t x = s_w[w];
    off += (w < t / kWave) ? x : 0u;
    sum += x;
  }
  __syncthreads();
  *total = sum;
  return off + incl - v;
}

// ---------------------------------------------------------------------------
// Stage 1 with the bucket counts.  Workgroup t encodes the points of scatter
// tile t (16-byte nontemporal float4 loads, four in flight per thread),
// writes the codes and stores the tile's bucket counts as row t of the
// table.  The first workgroup clears the status word (the build's first
// kernel).  Every lane of a wave runs the same iterations, so the counts'
// collectives run in uniform code; the lanes past the end take no part.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kHistBlock) void k_morton_bucket(const float4* __restrict__ pts,
                                                              uint32_t* __restrict__ codes, uint32_t n,
                                                              Quant quant, uint32_t shift, uint32_t n_bins,
                                                              uint32_t* __restrict__ table,
                                                              uint32_t* __restrict__ status) {
  __shared__ uint32_t s_h[kBucketMaxBins];
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < n_bins; i += kHistBlock) s_h[i] = 0;
  const uint32_t first = blockIdx.x * kBucketTile;
  const uint32_t end = min(first + kBucketTile, n);
  if (tid == 0 && status && blockIdx.x == 0) status[0] = 0;
  __syncthreads();
  const uint32_t top = n_bins - 1;  // codes are below 2^30: the clamp only guards the LDS array
  const uint32_t wave0 = tid & ~(kWave - 1u);
  constexpr uint32_t kDepth = 4;
  for (uint32_t j = first; j < end; j += kDepth * kHistBlock) {
    float4 q[kDepth];
#pragma unroll
    for (uint32_t r = 0; r < kDepth; ++r) {
      const uint32_t i = j + r * kHistBlock + tid;
      q[r] = i < end ? load_stream(pts + i) : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (uint32_t r = 0; r < kDepth; ++r) {
      if (j + r * kHistBlock + wave0 >= end) break;  // the wave's row lies past the end (uniform)
      const uint32_t i = j + r * kHistBlock + tid;
      const bool valid = i < end;
      const uint32_t k = morton_encode(q[r].x, q[r].y, q[r].z, quant);
      if (valid) codes[i] = k;
      // ordered points put whole waves in one bucket: one atomic then
      hist_add_if(s_h, min(k >> shift, top), valid, true);
    }
  }
  __syncthreads();
  uint32_t* row = table + static_cast<size_t>(blockIdx.x) * n_bins;
  for (uint32_t i = tid; i < n_bins; i += kHistBlock) row[i] = s_h[i];
}

// ---------------------------------------------------------------------------
// Tile scan.  Workgroup g owns buckets 32g .. 32g+31: thread (c, r) = (tid %
// 32, tid / 32) adds up column 32g + c over its run of tiles (one 128-byte
// segment per row and pair of waves, the run held in registers so that all
// its loads are in flight at once), the runs are scanned in LDS, and the
// thread rewrites its run as exclusive prefixes.  The bucket's total goes to
// count[].  No workgroup waits for another and none fences: the scatter,
// the next kernel, turns the totals into the buckets' starts.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_bucket_scan(uint32_t* __restrict__ table, uint32_t n_tiles,
                                                      uint32_t n_bins, uint32_t* __restrict__ count) {
  __shared__ uint32_t s_part[32][33];
  const uint32_t tid = threadIdx.x, c = tid % 32, r = tid / 32;
  const uint32_t col = blockIdx.x * 32 + c;
  const uint32_t per = (n_tiles + 31) / 32;  // at most kBucketScanRun
  const uint32_t t0 = min(r * per, n_tiles), t1 = min(t0 + per, n_tiles);
  uint32_t v[kBucketScanRun];
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t j = 0; j < kBucketScanRun; ++j)
    v[j] = t0 + j < t1 ? table[static_cast<size_t>(t0 + j) * n_bins + col] : 0u;
#pragma unroll
  for (uint32_t j = 0; j < kBucketScanRun; ++j) sum += v[j];
  s_part[r][c] = sum;
  __syncthreads();
  uint32_t run = 0;
  for (uint32_t j = 0; j < r; ++j) run += s_part[j][c];
#pragma unroll
  for (uint32_t j = 0; j < kBucketScanRun; ++j) {
    if (t0 + j < t1) table[static_cast<size_t>(t0 + j) * n_bins + col] = run;
    run += v[j];
  }
  if (r == 31) count[col] = run;  // the last run's end: the bucket's total
}

// ---------------------------------------------------------------------------
// Scatter.  Each workgroup first turns the buckets' totals into their starts
// (an exclusive scan in LDS, 8 KiB of input) and checks that every bucket
// fits an in-bucket sort and that the totals cover the n keys; workgroup 0
// publishes the starts and that verdict as the flag for the kernels that
// follow.  When the flag is 0 every workgroup returns; otherwise each moves
// one tile's keys to their buckets.  A tile's run in a bucket (about 8
// keys) ends where the next tile's begins, so the tiles are dealt to the
// XCDs in contiguous ranges (xcd_tile): the consecutive tiles one XCD runs
// at once complete each other's cache lines in its L2.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBucketScatterBlock) void k_bucket_scatter(
    const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t n, uint32_t shift, uint32_t n_bins,
    const uint32_t* __restrict__ table, const uint32_t* __restrict__ count, uint32_t* __restrict__ start,
    uint32_t* __restrict__ ctl) {
  constexpr int BLOCK = kBucketScatterBlock;
  constexpr int KPT = kBucketTile / BLOCK;
  static_assert(kBucketMaxBins <= 2 * BLOCK, "two buckets per thread");
  __shared__ uint32_t s_stage[kBucketTile];
  __shared__ uint32_t s_cnt[kBucketMaxBins];  // counts, then tile starts, then write bases
  __shared__ uint32_t s_w[BLOCK / kWave];
  __shared__ uint32_t s_big;
  const uint32_t tile = xcd_tile();
  if (tile * kBucketTile >= n) return;  // the grid's padding (never workgroup 0)
  const uint32_t tid = threadIdx.x;
  const uint32_t first = tile * kBucketTile;
  const uint32_t m = min(kBucketTile, n - first);
  const uint32_t top = n_bins - 1;
  const bool full = m == kBucketTile;
  // thread t owns buckets 2t and 2t + 1
  const uint32_t b0 = 2 * tid;
  uint32_t c0 = 0, c1 = 0;
  if (b0 < n_bins) {
    c0 = count[b0];
    c1 = count[b0 + 1];
  }
  if (tid == 0) s_big = 0;
  __syncthreads();
  if (c0 > kBucketCap || c1 > kBucketCap) atomicOr(&s_big, 1u);
  uint32_t total;
  const uint32_t bstart = block_excl_scan<BLOCK>(c0 + c1, s_w, &total);  // (its barriers publish s_big)
  const bool ok = s_big == 0 && total == n;
  if (blockIdx.x == 0) {
    if (b0 < n_bins) {
      start[b0] = bstart;
      start[b0 + 1] = bstart + c0;
    }
    if (tid == 0) {
      start[n_bins] = total;
      ctl[0] = ok ? 1u : 0u;
    }
  }
  if (!ok) return;  // a bucket too large for LDS: the digit-pass sort runs
  for (uint32_t i = tid; i < n_bins; i += BLOCK) s_cnt[i] = 0;
  uint32_t keys[KPT], rank[KPT];
  if (full) {
#pragma unroll
    for (int k = 0; k < KPT; ++k) keys[k] = in[first + k * BLOCK + tid];
  } else {
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const uint32_t i = k * BLOCK + tid;
      keys[k] = i < m ? in[first + i] : 0u;
    }
  }
  // the tile's starts inside its buckets load while the keys are ranked
  uint32_t g0 = 0, g1 = 0;
  if (b0 < n_bins) {
    const uint32_t* trow = table + static_cast<size_t>(tile) * n_bins;
    g0 = bstart + trow[b0];
    g1 = bstart + c0 + trow[b0 + 1];
  }
  __syncthreads();
  const auto bucket = [&](uint32_t key) { return min(key >> shift, top); };
  // ranks inside the tile's buckets (any order); ordered keys put a whole
  // wave in one bucket, which then takes one atomic
  const uint32_t wave0 = tid & ~(kWave - 1u);
  const auto ranks = [&](auto check) {
    if (full) {
#pragma unroll
      for (int k = 0; k < KPT; ++k) rank[k] = rank_add(s_cnt, bucket(keys[k]), decltype(check)::value);
    } else {
      // the last tile: rows past its end are skipped per wave, the rest
      // ranked with every lane taking part
#pragma unroll
      for (int k = 0; k < KPT; ++k) {
        if (k * BLOCK + wave0 >= m) break;
        rank[k] = rank_add_if(s_cnt, bucket(keys[k]), k * BLOCK + tid < m, decltype(check)::value);
      }
    }
  };
  if (wave_same(bucket(keys[0]))) ranks(std::true_type{});
  else ranks(std::false_type{});
  __syncthreads();
  uint32_t k0 = 0, k1 = 0;
  if (b0 < n_bins) {
    k0 = s_cnt[b0];
    k1 = s_cnt[b0 + 1];
  }
  const uint32_t ex = block_excl_scan<BLOCK>(k0 + k1, s_w, &total);
  if (b0 < n_bins) {
    s_cnt[b0] = ex;
    s_cnt[b0 + 1] = ex + k0;
  }
  __syncthreads();
  // the tile in bucket order
#pragma unroll
  for (int k = 0; k < KPT; ++k)
    if (full || k * BLOCK + tid < m) s_stage[s_cnt[bucket(keys[k])] + rank[k]] = keys[k];
  __syncthreads();
  // staged position p of bucket b goes to g_b + (p - tile start of b)
  // (modulo 2^32: the base alone may wrap)
  if (b0 < n_bins) {
    s_cnt[b0] = g0 - ex;
    s_cnt[b0 + 1] = g1 - (ex + k0);
  }
  __syncthreads();
  if (full) {
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const uint32_t p = k * BLOCK + tid;
      const uint32_t key = s_stage[p];
      out[s_cnt[bucket(key)] + p] = key;
    }
  } else {
    for (uint32_t p = tid; p < m; p += BLOCK) {
      const uint32_t key = s_stage[p];
      out[s_cnt[bucket(key)] + p] = key;
    }
  }
}

// ---------------------------------------------------------------------------
// In-bucket sort.  Workgroup b sorts bucket b (at most kBucketCap keys) by
// its low kBucketLowBits + HI bits.  Keys are held wave-striped (key k of
// lane l of wave w is the bucket's key w * KPT * 64 + k * 64 + l), the order
// in which the stable pass walks them.
// ---------------------------------------------------------------------------
template <int HI>
__global__ __launch_bounds__(kBucketLocalBlock) void k_bucket_local(const uint32_t* __restrict__ in,
                                                                    uint32_t* __restrict__ out,
                                                                    const uint32_t* __restrict__ start,
                                                                    const uint32_t* __restrict__ ctl) {
  constexpr int BLOCK = kBucketLocalBlock, KPT = kBucketLocalKpt, WAVES = BLOCK / kWave;
  constexpr uint32_t LO = kBucketLowBits, NLO = 1u << LO, NHI = 1u << HI;
  static_assert(NLO % (4 * BLOCK) == 0, "the low pass scans whole uint4 groups per thread");
  static_assert(WAVES * NHI * sizeof(uint16_t) <= NLO * sizeof(uint32_t),
                "the stable pass's 16-bit wave counters fit in the low pass's counters");
  static_assert(kBucketCap < 65536, "16-bit wave counters");
  __shared__ uint32_t s_stage[kBucketCap];
  __shared__ alignas(16) uint32_t s_cnt[NLO];
  __shared__ uint32_t s_w[WAVES];
  if (ctl[0] == 0) return;  // a bucket too large for LDS: the digit-pass sort runs
  const uint32_t base = start[blockIdx.x];
  // at most kBucketCap (the flag says so); the clamp keeps every LDS index in
  // bounds whatever the starts hold
  const uint32_t m = min(start[blockIdx.x + 1] - base, kBucketCap);
  if (m == 0) return;
  const uint32_t tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
  const uint32_t wb = wave * (KPT * kWave) + lane;
  for (uint32_t i = tid; i < NLO; i += BLOCK) s_cnt[i] = 0;
  uint32_t keys[KPT], rank[KPT];
#pragma unroll
  for (int k = 0; k < KPT; ++k) keys[k] = wb + k * kWave < m ? in[base + wb + k * kWave] : 0u;
  __syncthreads();

  // Pass 1: the low bits, ranks from LDS atomics.
  // (rows past the bucket are skipped per wave; every lane of the others
  // takes part in the collectives)
  const auto lo_ranks = [&](auto check) {
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      if (wave * (KPT * kWave) + k * kWave >= m) break;
      rank[k] = rank_add_if(s_cnt, keys[k] & (NLO - 1), wb + k * kWave < m, decltype(check)::value);
    }
  };
  if (wave_same(keys[0] & (NLO - 1))) lo_ranks(std::true_type{});
  else lo_ranks(std::false_type{});
  __syncthreads();
  {
    // the digits' starts: thread t owns digits [PER t, PER t + PER)
    constexpr uint32_t PER = NLO / BLOCK;
    uint4* c4 = reinterpret_cast<uint4*>(s_cnt) + tid * (PER / 4);
    uint32_t c[PER];
#pragma unroll
    for (uint32_t j = 0; j < PER / 4; ++j) {
      const uint4 v = c4[j];
      c[4 * j] = v.x;
      c[4 * j + 1] = v.y;
      c[4 * j + 2] = v.z;
      c[4 * j + 3] = v.w;
    }
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < PER; ++j) sum += c[j];
    uint32_t total;
    uint32_t run = block_excl_scan<BLOCK>(sum, s_w, &total);
#pragma unroll
    for (uint32_t j = 0; j < PER; ++j) {
      const uint32_t x = c[j];
      c[j] = run;
      run += x;
    }
#pragma unroll
    for (uint32_t j = 0; j < PER / 4; ++j) c4[j] = make_uint4(c[4 * j], c[4 * j + 1], c[4 * j + 2], c[4 * j + 3]);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KPT; ++k)
    if (wb + k * kWave < m) s_stage[s_cnt[keys[k] & (NLO - 1)] + rank[k]] = keys[k];
  __syncthreads();

  // Pass 2: the high bits, stable.  Keys past the bucket (in its last rows)
  // read as all ones: they rank after every key of the top digit, which ends
  // the bucket, and are not staged.
#pragma unroll
  for (int k = 0; k < KPT; ++k) keys[k] = wb + k * kWave < m ? s_stage[wb + k * kWave] : 0xffffffffu;
  uint16_t* const cnt16 = reinterpret_cast<uint16_t*>(s_cnt);
  for (uint32_t i = tid; i < WAVES * NHI / 2; i += BLOCK) s_cnt[i] = 0;
  __syncthreads();  // every key read before the stage is rewritten; the counters clear
  uint16_t* const wh = cnt16 + wave * NHI;
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    if (wave * (KPT * kWave) + k * kWave >= m) break;  // rows past the bucket (wave-uniform)
    const uint32_t d = (keys[k] >> LO) & (NHI - 1);
    const uint64_t peers = match_peers<HI>(d);
    const uint32_t below = popc_below(peers);
    const uint32_t old = wh[d];
    // every lane has read its counter before a group's first lane writes it
    // (a wave issues the read for all its lanes first)
    __builtin_amdgcn_wave_barrier();
    if (below == 0) wh[d] = static_cast<uint16_t>(old + static_cast<uint32_t>(__popcll(peers)));
    rank[k] = old + below;
  }
  __syncthreads();
  {
    // per digit: its start in the bucket, then each wave's start of it
    constexpr uint32_t DP = NHI > BLOCK ? NHI / BLOCK : 1;
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < DP; ++j) {
      const uint32_t d = tid * DP + j;
      if (d < NHI)
#pragma unroll
        for (int w = 0; w < WAVES; ++w) sum += cnt16[w * NHI + d];
    }
    uint32_t total;
    uint32_t run = block_excl_scan<BLOCK>(sum, s_w, &total);
#pragma unroll
    for (uint32_t j = 0; j < DP; ++j) {
      const uint32_t d = tid * DP + j;
      if (d < NHI)
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
          const uint32_t x = cnt16[w * NHI + d];
          cnt16[w * NHI + d] = static_cast<uint16_t>(run);
          run += x;
        }
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KPT; ++k)
    if (wb + k * kWave < m) s_stage[wh[(keys[k] >> LO) & (NHI - 1)] + rank[k]] = keys[k];
  __syncthreads();
  for (uint32_t p = tid; p < m; p += BLOCK) out[base + p] = s_stage[p];
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
void launch_morton_bucket(const float4* pts, uint32_t* codes, uint32_t n, float min_coord, float range,
                          const BucketPlan& plan, uint32_t* table, uint32_t* status, hipStream_t stream) {
  if (n == 0 || !plan.ok) return;
  hipLaunchKernelGGL(k_morton_bucket, dim3(plan.n_tiles), dim3(kHistBlock), 0, stream, pts, codes, n,
                     make_quant(min_coord, range), plan.shift, plan.n_bins, table, status);
}

void launch_bucket_sort(const BucketLaunch& B, hipStream_t stream) {
  const BucketPlan& p = B.plan;
  if (B.n == 0 || !p.ok) return;
  hipLaunchKernelGGL(k_bucket_scan, dim3(p.n_bins / 32), dim3(1024), 0, stream, B.table, p.n_tiles, p.n_bins,
                     B.count);
  hipLaunchKernelGGL(k_bucket_scatter, dim3(xcd_grid(p.n_tiles)), dim3(kBucketScatterBlock), 0, stream, B.keys_in,
                     B.keys_tmp, B.n, p.shift, p.n_bins, B.table, B.count, B.start, B.ctl);
  const auto local = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(p.n_bins), dim3(kBucketLocalBlock), 0, stream, B.keys_tmp, B.keys_out, B.start,
                       B.ctl);
  };
  switch (p.hi_bits) {
    case 7: local(k_bucket_local<7>); break;
    case 8: local(k_bucket_local<8>); break;
    case 9: local(k_bucket_local<9>); break;
    default: local(k_bucket_local<10>); break;
  }
}

}  // namespace rw
