// Host-side internals of the HIP library: launch descriptors shared between
// the kernel translation units and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>

#include "rw_device.hpp"
#include "rwhip.h"

namespace rw {

// Sort shapes.  A digit-pass tile is BLOCK threads x KPT keys; the variants
// below are compiled, one is chosen per call (the autotuner of the host layer
// picks per size bucket).  A tile larger than kChunk could straddle two
// clouds, so such shapes serve single-cloud sorts only; a batched sort with
// one of them selected runs kSortChunkVariant instead.
struct SortVariant {
  int block, kpt;
};
constexpr SortVariant kSortVariants[] = {{256, 16}, {512, 16}, {256, 32}, {1024, 8}, {512, 8}, {1024, 16}};
constexpr int kNumSortVariants = sizeof(kSortVariants) / sizeof(kSortVariants[0]);
constexpr int kSortChunkVariant = 3;  // 1024 x 8: the chunk-sized tile
constexpr bool sort_tile_fits_chunk(int v) {
  return static_cast<uint32_t>(kSortVariants[v].block * kSortVariants[v].kpt) <= kChunk;
}
static_assert(sort_tile_fits_chunk(kSortChunkVariant), "fallback must fit a chunk");
constexpr uint32_t kSortMinTile = 4096;  // smallest tile: sizes the look-back table
constexpr int kHistBlock = 1024;
constexpr uint32_t kHistCopies = 16;
constexpr uint32_t kMaxPasses = 4;
constexpr uint32_t kPlainHistUnit = 16384;  // keys per histogram workgroup, one segment (the largest sort tile)
constexpr uint32_t kTilesPerHistUnit = kPlainHistUnit / kSortMinTile;
// log2 of a variant's tile size; a single-cloud histogram unit holds whole tiles
constexpr uint32_t sort_tile_shift(int v) {
  uint32_t s = 0;
  while ((1u << s) < static_cast<uint32_t>(kSortVariants[v].block * kSortVariants[v].kpt)) ++s;
  return s;
}
constexpr bool sort_tiles_nest() {
  for (int v = 0; v < kNumSortVariants; ++v) {
    const uint32_t t = kSortVariants[v].block * kSortVariants[v].kpt;
    if ((1u << sort_tile_shift(v)) != t || kPlainHistUnit % t != 0 || t < kSortMinTile) return false;
  }
  return true;
}
static_assert(sort_tiles_nest(), "tiles are powers of two dividing the histogram unit");
constexpr int kLookbackWindow = 4;           // look-back words loaded per round trip

// The digit passes of a sort: pass p (least significant first) orders the
// keys by bits [shift[p], shift[p] + bits[p]), with 1 <= bits[p] <= 8.
struct DigitPlan {
  uint32_t n_passes;
  uint32_t shift[kMaxPasses];
  uint32_t bits[kMaxPasses];
};
// Digit widths of the plans for 30-bit keys (Morton codes), least
// significant pass first; a workspace runs one of them (the host layer's
// autotuner picks per size bucket).  Sorts of keys of any other width run
// 8-bit passes with the remaining bits in the last one.
constexpr uint8_t kPlanWidths[][kMaxPasses] = {{8, 8, 8, 6}, {8, 7, 7, 8}, {8, 8, 7, 7},
                                                 {8, 6, 8, 8}, {7, 7, 8, 8}, {6, 8, 8, 8}};
constexpr int kNumSortPlans = sizeof(kPlanWidths) / sizeof(kPlanWidths[0]);
constexpr bool plans_cover_30_bits() {
  for (int p = 0; p < kNumSortPlans; ++p) {
    uint32_t sum = 0;
    for (uint32_t k = 0; k < kMaxPasses; ++k) {
      if (kPlanWidths[p][k] < 1 || kPlanWidths[p][k] > 8) return false;
      sum += kPlanWidths[p][k];
    }
    if (sum != 30 || kPlanWidths[p][0] < 6) return false;  // pass 0 has a 6..8-bit kernel
  }
  return true;
}
static_assert(plans_cover_30_bits(), "each plan sorts 30 bits in four passes of 1..8 bits");

constexpr DigitPlan make_plan(uint32_t key_bits, int plan) {
  DigitPlan d{};
  if (key_bits == 30 && plan >= 0 && plan < kNumSortPlans) {
    d.n_passes = kMaxPasses;
    uint32_t sh = 0;
    for (uint32_t p = 0; p < kMaxPasses; ++p) {
      d.shift[p] = sh;
      d.bits[p] = kPlanWidths[plan][p];
      sh += d.bits[p];
    }
    return d;
  }
  d.n_passes = (key_bits + 7) / 8;
  for (uint32_t p = 0; p < d.n_passes; ++p) {
    d.shift[p] = 8 * p;
    d.bits[p] = (p + 1 == d.n_passes) ? key_bits - 8 * p : 8u;
  }
  return d;
}

// Compaction / scan tiles.
constexpr int kScanBlock = 512;
constexpr int kScanKpt = 16;
constexpr uint32_t kScanTile = kScanBlock * kScanKpt;  // 8192: a chunk, any cloud layout
// A single cloud's compaction and scans run on a larger tile: fewer tiles mean
// fewer serialised grabs and shorter look-back chains.
constexpr int kScanBigBlock = 1024;
constexpr int kScanBigKpt = 16;
constexpr uint32_t kScanBigTile = kScanBigBlock * kScanBigKpt;  // 16384
static_assert(kScanTile == kChunk, "segmented scans tile by chunk");

// Elementwise tiles.
constexpr int kElemBlock = 256;

// Radix tree (S4): tiles of leaves grown in LDS, then the nodes that cross a
// tile edge (listed per tile) by groups of lanes.
constexpr int kTreeTile = 4096;   // leaves per tile (the segment's last tile also takes leaf n)
constexpr int kTreeBlock = 512;
constexpr int kTreeListCap = 64;  // edge-crossing nodes per tile: at most 31 per edge, plus node 0
constexpr int kTreeListStride = kTreeListCap + 1;  // count, then the nodes
// loop bound of a wave's climbs: every iteration takes a leaf, forms a node
// or ends a climb, and a tile has at most kTreeTile + 1 leaves and nodes
constexpr int kTreeClimbGuard = 3 * kTreeTile + 8;
// leaf pairs a wave queues: its block's kTreeTile / waves (+1) leaves pair up
// at most half each, plus one pair across each end of the block
constexpr int kTreePairCap = kTreeTile / (kTreeBlock / kWave) / 2 + 4;
constexpr int kCrossGroup = 8;    // lanes per edge-crossing node
constexpr uint32_t kCrossMask = (1u << kCrossGroup) - 1u;
constexpr int kCrossBracket = 32 / kCrossGroup;  // the range bracket: lane r probes 2^(r * this)
constexpr int kCrossBlock = 256;  // four waves: 32 nodes at a time
static_assert(kChunk % kTreeTile == 0, "a tree tile never straddles two clouds");
static_assert(kTreeTile < 0xffff, "LDS node offsets are 16-bit with 0xffff reserved");
static_assert(kWave % kCrossGroup == 0 && kCrossBracket * (kCrossGroup - 1) <= 30, "bracket probes below 2^31");
constexpr uint32_t tree_list_words(uint32_t n_index) {
  return ((n_index + kTreeTile - 1) / kTreeTile) * static_cast<uint32_t>(kTreeListStride);
}

// Workgroups of a kernel that fit on the device at once (at most n_tiles):
// the look-back kernels loop over their tiles, so a larger grid would only
// queue.
// Cached per kernel and device (a benign race computes the same value
// twice); any grid is correct, this one is the fastest.
template <auto Kernel>
uint32_t resident_grid(int block, uint32_t n_tiles) {
  constexpr int kMaxDevices = 64;
  static std::atomic<uint32_t> cache[kMaxDevices];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) dev = 0;
  uint32_t g = cache[dev].load(std::memory_order_relaxed);
  if (g == 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, Kernel, block, 0) != hipSuccess || per_cu < 1)
      per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
      cus = 1;
    g = static_cast<uint32_t>(per_cu) * static_cast<uint32_t>(cus);
    cache[dev].store(g, std::memory_order_relaxed);
  }
  return std::max<uint32_t>(1, std::min(g, n_tiles));
}

// Grid of n_tiles workgroups padded for xcd_tile().
inline uint32_t xcd_grid(uint32_t n_tiles) { return (n_tiles + kXcds - 1) / kXcds * kXcds; }

struct SortLaunch {
  const uint32_t* keys_in;
  uint32_t* keys_out;
  uint32_t* keys_tmp;
  SegArgs seg;
  uint32_t n_segs;
  int variant;              // index into kSortVariants
  uint32_t n_index;        // index-space size (tiles are derived per variant)
  uint32_t hist_unit_keys;  // keys per histogram workgroup
  uint32_t n_hist_units;
  DigitPlan plan;          // the digit passes (make_plan of the significant key bits)
  uint32_t* hist;          // [segs][kHistCopies][kMaxPasses*256], zero between sorts
  uint32_t* digit_start;   // [segs][passes][256]
  uint64_t* lookback;      // [n_tiles][256]
  uint32_t* ctrl;          // kCtrlWords
  bool hist_ready;         // hist already counted (launch_morton_hist)
  // [256][n_tiles]: pass 0 tile counts, for a single-cloud sort of 9+ bits
  // (nullptr, or any other sort: pass 0 runs the look-back chain)
  uint32_t* tile_counts;
  // nullptr, or a device word: when it is non-zero (the bucket sort below ran
  // instead) every kernel of the sort returns at once (the scans after
  // clearing the histogram copies, which then hold no counts)
  const uint32_t* skip;
};
void launch_sort(const SortLaunch& L, hipStream_t stream);
// Stage 1 fused with the sort's histogram sweep, over units of unit_keys
// points each inside one segment (the sort's plain unit for one cloud, a
// chunk for a batch): the digit histograms of the plan's passes go to each
// segment's hist copies, and, for one cloud only, pass 0's counts per tile of
// sort_variant to tile_counts unless it is nullptr.  status (per segment, or
// nullptr) is cleared by the launch: the build's first kernel.
void launch_morton_hist(const float4* pts, uint32_t* codes, SegArgs seg, uint32_t n_index, uint32_t unit_keys,
                        float min_coord, float range, const DigitPlan& plan, uint32_t* hist, int sort_variant,
                        uint32_t* tile_counts, uint32_t* status, hipStream_t stream);

// Bucket sort of a single cloud's 30-bit keys (Morton codes), two passes over
// HBM instead of four: the keys are scattered by their top D bits into 2^D
// buckets (about 8192 keys each; no order needed within a bucket, so ranks
// come from LDS atomics), then one workgroup per bucket sorts it in LDS by
// its low R = 30 - D bits: an atomic-ranked pass on the low kBucketLowBits
// bits and one stable multisplit pass on the rest.  A bucket larger than
// kBucketCap keys (clustered input) cannot be held in LDS: then the bucket
// kernels return at once and the digit-pass sort runs instead, gated by the
// same device flag.  Equal keys are indistinguishable, so the output is the
// same as the digit-pass sort's.
constexpr uint32_t kBucketTile = 16384;      // keys per scatter tile (and per S1 count unit)
constexpr int kBucketScatterBlock = 1024;    // 16 keys per thread
constexpr uint32_t kBucketMaxBins = 2048;    // D <= 11
constexpr int kBucketLocalBlock = 512;
constexpr int kBucketLocalKpt = 18;
constexpr uint32_t kBucketCap = kBucketLocalBlock * kBucketLocalKpt;  // 9216 keys
constexpr uint32_t kBucketLowBits = 12;      // the atomic-ranked in-bucket pass
constexpr uint32_t kBucketMinLog = 21, kBucketMaxLog = 24;  // sizes in (2^20, 2^24]
// tiles per thread of the start scan (32 runs of them cover every tile)
constexpr uint32_t kBucketScanRun = ((1u << kBucketMaxLog) / kBucketTile + 31) / 32;
static_assert(kBucketTile % (kHistBlock * 4) == 0, "stage 1 reads whole rows of four loads per thread");
static_assert(kBucketMaxBins % 32 == 0, "the start scan takes 32 buckets per workgroup");
struct BucketPlan {
  bool ok;            // n in the bucket sort's range
  uint32_t shift;     // R: the bucket of key k is k >> R
  uint32_t n_bins;    // 2^D
  uint32_t n_tiles;   // scatter tiles
  uint32_t hi_bits;   // R - kBucketLowBits: the stable in-bucket pass
};
// D = ceil(log2 n) - 13, so that a bucket averages (4096, 8192] keys of
// uniform input: kBucketCap is more than ten standard deviations above.
constexpr BucketPlan bucket_plan(uint32_t n) {
  BucketPlan p{};
  uint32_t lg = 0;
  while (lg < 32 && (uint64_t{1} << lg) < n) ++lg;
  if (lg < kBucketMinLog || lg > kBucketMaxLog) return p;
  const uint32_t d = lg - 13;
  p.ok = true;
  p.shift = 30 - d;
  p.n_bins = 1u << d;
  p.n_tiles = (n + kBucketTile - 1) / kBucketTile;
  p.hi_bits = p.shift - kBucketLowBits;
  return p;
}
// Words of the per-tile bucket count table for any size up to max_points.
constexpr size_t bucket_table_words(uint32_t max_points) {
  size_t w = 0;
  for (uint32_t lg = kBucketMinLog; lg <= kBucketMaxLog; ++lg) {
    const uint64_t n = std::min<uint64_t>(max_points, uint64_t{1} << lg);
    const BucketPlan p = bucket_plan(static_cast<uint32_t>(n));
    if (p.ok) w = std::max<size_t>(w, size_t{p.n_bins} * p.n_tiles);
  }
  return w;
}
constexpr uint32_t kBucketCtlWords = 4;  // [0]: the flag
struct BucketLaunch {
  const uint32_t* keys_in;  // n keys below 2^30
  uint32_t* keys_out;       // sorted (untouched when the flag refuses the bucket sort)
  uint32_t* keys_tmp;       // the scattered buckets
  uint32_t n;
  BucketPlan plan;
  uint32_t* table;   // [n_tiles][n_bins] bucket counts of each tile (S1), scanned in place
  uint32_t* count;   // [n_bins] bucket totals
  uint32_t* start;   // [n_bins + 1] bucket starts
  uint32_t* ctl;     // kBucketCtlWords: [0] the flag (1: the bucket sort runs), written by every launch
};
// Stage 1 for the bucket sort: codes, and the bucket counts of each scatter
// tile into table; status (or nullptr) is cleared by the launch.
void launch_morton_bucket(const float4* pts, uint32_t* codes, uint32_t n, float min_coord, float range,
                          const BucketPlan& plan, uint32_t* table, uint32_t* status, hipStream_t stream);
// The tile scan, the scatter (which sets the flag) and the in-bucket sorts.
void launch_bucket_sort(const BucketLaunch& B, hipStream_t stream);

// status (per segment, or nullptr) is cleared by the launch: the first kernel
// of a batched build.
void launch_morton(const float4* pts, uint32_t* codes, SegArgs seg, uint32_t n_index,
                   float min_coord, float range, uint32_t* status, hipStream_t stream);

// Compaction of adjacent-distinct keys; n_unique[s] per segment.
void launch_unique(const uint32_t* sorted, uint32_t* unique_out, SegArgs seg, uint32_t n_index,
                   uint32_t* n_unique, uint64_t* lookback, uint32_t* ctrl, hipStream_t stream);

struct TreeArgs {
  const uint32_t* codes;  // unique keys (index space)
  uint8_t* prefix_n;
  uint8_t* has_leaf_left;
  uint8_t* has_leaf_right;
  int32_t* left_child;
  int32_t* parent;
  // optional: S5's edge counts, written beside the tree (keys below 2^31)
  int32_t* edge_count = nullptr;
  // optional with edge_count: per tree tile (kTreeTile nodes of index space)
  // the sum of its nodes' edge counts, for a scan without look-back
  uint32_t* tile_sum = nullptr;
  // optional: per segment, RWH_STATUS_* bits OR-ed in where a bound stops the work
  uint32_t* status = nullptr;
};
// seg.count = n_unique per segment with count_sub = 1, i.e. the live count is n_brt
void launch_radix_tree(const TreeArgs& a, SegArgs seg, uint32_t n_index, uint32_t* list, hipStream_t stream);
void launch_edge_count(const uint8_t* prefix_n, const int32_t* parent, int32_t* edge_count,
                       SegArgs seg, uint32_t n_index, hipStream_t stream);
// Exclusive scan of the live count (n_brt) per segment; total per segment.
void launch_prefix_sum(const int32_t* in, int32_t* out, SegArgs seg, uint32_t n_index,
                       uint32_t* total, uint64_t* lookback, uint32_t* ctrl, hipStream_t stream);
// The same scan of edge counts whose tree tile sums (TreeArgs::tile_sum) are
// known: each tile adds up the sums of the segment's earlier tree tiles.
void launch_prefix_sum_tiles(const int32_t* in, int32_t* out, SegArgs seg, uint32_t n_index,
                             uint32_t* total, const uint32_t* tile_sum, hipStream_t stream);
constexpr uint32_t tree_tiles(uint32_t n_index) { return (n_index + kTreeTile - 1) / kTreeTile; }

// Octree nodes (S7): tiles of radix nodes.  A tile's pull walks that leave
// its LDS window are queued per tile (at most kOctFarCap, the rest walked in
// place) and walked from global memory by a second, short kernel.
constexpr uint32_t kOctSpan = 1024;   // radix nodes per tile
constexpr uint32_t kOctFarCap = 16;   // queued walks per tile
static_assert(kChunk % kOctSpan == 0, "octree tiles must not straddle segments");
constexpr uint32_t oct_tiles(uint32_t n_index) { return (n_index + kOctSpan - 1) / kOctSpan; }

struct OctArgs {
  const int32_t* edge_offset;
  const int32_t* edge_count;
  const uint32_t* codes;
  const uint8_t* prefix_n;
  const int32_t* parent;
  const uint8_t* has_leaf_left;
  const uint8_t* has_leaf_right;
  const int32_t* left_child;
  const uint32_t* n_oct;     // per segment
  const uint32_t* oct_base;  // per segment (nullptr: 0)
  const uint32_t* oct_cap;   // per segment (nullptr: single_cap)
  uint32_t single_cap;
  float min_coord, range;
  int32_t* children;
  float4* corner;
  float* cell_size;
  int32_t* child_node_mask;
  int32_t* child_leaf_mask;
  uint32_t* status;  // per segment, or nullptr: RWH_STATUS_* bits OR-ed in
  // queued far walks: kOctFarCap items per tile (bottom slot relative to
  // the segment's octree base, node relative to the segment); nullptr:
  // every walk runs in place
  uint2* far_item;
};
// check: a tree handed in through the C ABI (child indices checked, walk
// bounds reported in a.status); the whole builds pass the tree their own
// radix-tree kernels built, unchecked.
void launch_octree(const OctArgs& a, SegArgs seg, uint32_t n_index, bool check, hipStream_t stream);
// dst[0, count) = value with a kernel (no memset node, so a captured graph
// replays it as it ran)
void launch_fill_u32(uint32_t* dst, uint32_t value, uint32_t count, hipStream_t stream);
void launch_link_leaf(const OctArgs& a, SegArgs seg, uint32_t n_index, hipStream_t stream);

// ---- packed result (format version 1 of include/rwhip.h) ----
static_assert(sizeof(rwh_result_header) == RWH_RESULT_HEADER_BYTES, "header is 256 bytes");
__host__ __device__ constexpr uint32_t result_entry_bytes(int s) {
  return s <= RWH_SEC_UNIQUE_KEYS ? 4u
         : s <= RWH_SEC_HAS_LEAF_RIGHT ? 1u
         : s <= RWH_SEC_EDGE_OFFSET ? 4u
         : s == RWH_SEC_CHILDREN ? 32u
         : s == RWH_SEC_CORNER ? 16u
                               : 4u;
}
struct ResultLayout {
  uint64_t offset[RWH_SEC_COUNT];
  uint32_t count[RWH_SEC_COUNT];
  uint64_t total;
};
// Sections in order after the header, each at a multiple of RWH_RESULT_ALIGN;
// n_unique <= n_input and rows <= the octree capacity are the caller's.
__host__ __device__ constexpr ResultLayout result_layout(uint32_t n_input, uint32_t n_unique,
                                                         uint32_t rows, uint32_t sections) {
  ResultLayout L{};
  const uint32_t n_brt = n_unique > 0 ? n_unique - 1 : 0;
  uint64_t at = RWH_RESULT_HEADER_BYTES;
  for (int s = 0; s < RWH_SEC_COUNT; ++s) {
    L.offset[s] = 0;
    L.count[s] = 0;
    if (!((sections >> s) & 1u)) continue;
    const uint32_t c = s == RWH_SEC_SORTED_KEYS   ? n_input
                       : s == RWH_SEC_UNIQUE_KEYS ? n_unique
                       : s <= RWH_SEC_EDGE_OFFSET ? n_brt
                                                  : rows;
    L.offset[s] = at;
    L.count[s] = c;
    at = (at + static_cast<uint64_t>(c) * result_entry_bytes(s) + RWH_RESULT_ALIGN - 1) &
         ~static_cast<uint64_t>(RWH_RESULT_ALIGN - 1);
  }
  L.total = at;
  return L;
}
struct PackArgs {
  const unsigned char* src[RWH_SEC_COUNT];  // the cloud's arrays (absent sections: nullptr)
  const uint32_t* counts;                   // n_unique, n_oct, status at 0, stride, 2 stride
  uint32_t stride;
  uint32_t n_input, oct_capacity, sections;
  unsigned char* out;
};
// One kernel: header + sections, 16 bytes per work-item step; the grid covers
// max_bytes (the largest result of n_input points).
void launch_pack_result(const PackArgs& a, uint64_t max_bytes, hipStream_t stream);

}  // namespace rw
