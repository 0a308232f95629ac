// The C ABI of the octree-build backend (include/rwhip.h): argument checks,
// workspace management, HIP plumbing, and the stage / build entry points that
// enqueue the kernels of k_sort.hip, k_stages.hip and k_octree.hip.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "rw_internal.hpp"
#include "rwhip.h"

using namespace rw;

namespace {

thread_local std::string g_detail;
thread_local int g_fail_stage = -1;

int fail(int status, const std::string& msg) {
  g_detail = msg;
  return status;
}

#define RWH_HIP(call)                                                                   \
  do {                                                                                  \
    const hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                               \
      return fail(e_ == hipErrorNoDevice ? RWH_ERR_NO_DEVICE : RWH_ERR_HIP,             \
                  std::string(#call) + ": " + hipGetErrorString(e_));                   \
  } while (0)

#define RWH_CHECK(cond, msg) \
  do {                       \
    if (!(cond)) return fail(RWH_ERR_INVALID_ARG, msg); \
  } while (0)

#define RWH_TRY(expr)             \
  do {                            \
    const int s_ = (expr);        \
    if (s_ != RWH_OK) return s_;  \
  } while (0)

// The names of the bounds set in a status word.
std::string status_text(uint32_t bits) {
  static const struct {
    uint32_t bit;
    const char* what;
  } kNames[] = {{RWH_STATUS_PAIR_CAP, "radix tree: leaf-pair queue full"},
                {RWH_STATUS_CLIMB, "radix tree: climb bound reached"},
                {RWH_STATUS_CROSS_CAP, "radix tree: edge-crossing list full"},
                {RWH_STATUS_KEY_ORDER, "radix tree: keys not sorted and unique"},
                {RWH_STATUS_PULL, "octree: link walk bound reached or index outside the tree"},
                {RWH_STATUS_LEAF, "leaf linking: index outside the tree"}};
  std::string out;
  for (const auto& k : kNames)
    if (bits & k.bit) out += (out.empty() ? "" : "; ") + std::string(k.what);
  if (bits & ~0x3fu) out += (out.empty() ? "" : "; ") + std::string("unknown bits");
  return out;
}

int injected(int stage) {
  if (g_fail_stage < 0) return 0;
  if (g_fail_stage == 0 || g_fail_stage == stage) {
    g_fail_stage = -1;
    g_detail = "failure injected at stage " + std::to_string(stage);
    return 1;
  }
  return 0;
}

int launched() {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(RWH_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return RWH_OK;
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

constexpr uint32_t kMaxPoints = 1u << 30;

SegArgs single_seg(uint32_t n) {
  SegArgs s{};
  s.n_segs = 1;
  s.single_count = n;
  return s;
}

SegArgs single_seg_dev(const uint32_t* d_count) {
  SegArgs s{};
  s.n_segs = 1;
  s.count = d_count;
  return s;
}

}  // namespace

struct rwh_workspace {
  uint32_t magic = 0;       // ws_magic(): the layout this library built it with
  uint32_t max_points = 0;  // index-space capacity
  uint32_t max_segs = 0;
  uint32_t max_chunks = 0;
  uint32_t max_sort_tiles = 0;
  uint32_t max_scan_tiles = 0;
  size_t bytes = 0;
  void* block = nullptr;  // one device allocation for everything below (none for a view of the scratch)
  uint32_t* keys_tmp = nullptr;
  uint32_t* hist = nullptr;
  uint32_t* digit_start = nullptr;
  uint64_t* lb_sort = nullptr;
  uint32_t* tile_counts = nullptr;  // [256][tiles]: pass 0 tile prefixes (max_sort_tiles*256 words)
  uint64_t* lb_unique = nullptr;
  uint64_t* lb_scan = nullptr;
  uint32_t* ctrl = nullptr;    // 3 users x kCtrlWords
  uint32_t* counts = nullptr;  // n_unique[max_segs], n_oct[max_segs], status[max_segs]
  uint32_t* tables = nullptr;  // seg_base, seg_n, oct_base, oct_cap [max_segs each], chunk_seg
  uint32_t* tree_list = nullptr;  // radix tree: edge-crossing nodes per tile
  uint32_t* tree_sums = nullptr;  // radix tree: edge counts summed per tile (build path)
  uint2* far_item = nullptr;      // octree: far walks queued per tile (kOctFarCap each)
  uint32_t* bucket_table = nullptr;  // bucket sort: [tiles][bins] counts, then tile starts
  uint32_t* bucket_start = nullptr;  // bucket sort: [kBucketMaxBins + 1] starts, then [kBucketMaxBins] totals
  uint32_t* bucket_ctl = nullptr;    // bucket sort: the flag (kBucketCtlWords)
  uint32_t* h_stage = nullptr;  // pinned staging of tables
  uint32_t* h_counts = nullptr;  // pinned counts read-back
  hipEvent_t upload_done = nullptr;
  bool upload_pending = false;
  uint32_t batch_clouds = 0;  // shape of the last staged batch
  uint32_t batch_index = 0;
  int sort_variant = 0;       // index into kSortVariants
  int sort_plan = 0;          // index into kPlanWidths (sorts of 30-bit keys)
  bool bucket_sort = true;    // single-cloud builds in the bucket sort's range use it
  bool bucket_tried = false;  // the last whole build launched the bucket sort

  uint32_t* seg_base() { return tables; }
  uint32_t* seg_n() { return tables + max_segs; }
  uint32_t* oct_base() { return tables + 2 * max_segs; }
  uint32_t* oct_cap() { return tables + 3 * max_segs; }
  uint32_t* chunk_seg() { return tables + 4 * max_segs; }
  size_t table_words() const { return 4ull * max_segs + max_chunks; }
  uint32_t* n_unique() { return counts; }
  uint32_t* n_oct() { return counts + max_segs; }
  uint32_t* status() { return counts + 2 * max_segs; }
};

namespace {

// A fingerprint of rwh_workspace's layout.  Two builds of this library with
// different layouts can be loaded in one process; a workspace one of them
// made is refused by the other instead of being read at the wrong offsets.
constexpr uint32_t ws_magic() {
  return 0x52575300u ^ static_cast<uint32_t>(sizeof(rwh_workspace)) ^
         (static_cast<uint32_t>(offsetof(rwh_workspace, tree_sums)) << 8) ^
         (static_cast<uint32_t>(offsetof(rwh_workspace, far_item)) << 16) ^
         (static_cast<uint32_t>(offsetof(rwh_workspace, bucket_ctl)) << 20);
}

constexpr size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Byte offsets of a workspace's areas inside its one device block, for
// max_points index-space points and max_segs clouds.
struct WsLayout {
  uint32_t max_chunks, max_sort_tiles, max_scan_tiles;
  size_t tmp, hist, ds, lbs, tc, lbu, lbc, ctrl, cnt, tab, tl, ts, fi, bkt, bks, bkc;
  size_t bytes;
};

constexpr WsLayout ws_layout(uint32_t max_points, uint32_t max_segs) {
  WsLayout L{};
  const uint32_t mp = max_points > 0 ? max_points : 1u;
  L.max_chunks = (mp + kChunk - 1) / kChunk;
  L.max_sort_tiles = (mp + kSortMinTile - 1) / kSortMinTile;
  L.max_scan_tiles = (mp + kScanTile - 1) / kScanTile;
  size_t off = 0;
  const auto take = [&off](size_t bytes) {
    const size_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  L.tmp = take(sizeof(uint32_t) * static_cast<size_t>(mp));
  L.hist = take(sizeof(uint32_t) * static_cast<size_t>(max_segs) * kHistCopies * kMaxPasses * 256);
  L.ds = take(sizeof(uint32_t) * static_cast<size_t>(max_segs) * kMaxPasses * 256);
  L.lbs = take(sizeof(uint64_t) * static_cast<size_t>(L.max_sort_tiles) * 256);
  L.tc = take(sizeof(uint32_t) * static_cast<size_t>(L.max_sort_tiles) * 256);
  L.lbu = take(sizeof(uint64_t) * static_cast<size_t>(L.max_scan_tiles));
  L.lbc = take(sizeof(uint64_t) * static_cast<size_t>(L.max_scan_tiles));
  L.ctrl = take(sizeof(uint32_t) * 3 * kCtrlWords);
  L.cnt = take(sizeof(uint32_t) * 3ull * max_segs);
  L.tab = take(sizeof(uint32_t) * (4ull * max_segs + L.max_chunks));
  L.tl = take(sizeof(uint32_t) * static_cast<size_t>(tree_list_words(mp)));
  L.ts = take(sizeof(uint32_t) * static_cast<size_t>(tree_tiles(mp)));
  L.fi = take(sizeof(uint2) * static_cast<size_t>(oct_tiles(mp)) * kOctFarCap);
  L.bkt = take(sizeof(uint32_t) * std::max<size_t>(1, bucket_table_words(mp)));
  L.bks = take(sizeof(uint32_t) * (2 * kBucketMaxBins + 1));
  L.bkc = take(sizeof(uint32_t) * kBucketCtlWords);
  L.bytes = off;
  return L;
}

// Points the workspace's areas into base (a block laid out by L).
void bind_workspace(rwh_workspace* ws, char* base, const WsLayout& L, uint32_t max_points, uint32_t max_segs) {
  ws->magic = ws_magic();
  ws->max_points = max_points;
  ws->max_segs = max_segs;
  ws->max_chunks = L.max_chunks;
  ws->max_sort_tiles = L.max_sort_tiles;
  ws->max_scan_tiles = L.max_scan_tiles;
  ws->bytes = L.bytes;
  ws->keys_tmp = reinterpret_cast<uint32_t*>(base + L.tmp);
  ws->hist = reinterpret_cast<uint32_t*>(base + L.hist);
  ws->digit_start = reinterpret_cast<uint32_t*>(base + L.ds);
  ws->lb_sort = reinterpret_cast<uint64_t*>(base + L.lbs);
  ws->tile_counts = reinterpret_cast<uint32_t*>(base + L.tc);
  ws->lb_unique = reinterpret_cast<uint64_t*>(base + L.lbu);
  ws->lb_scan = reinterpret_cast<uint64_t*>(base + L.lbc);
  ws->ctrl = reinterpret_cast<uint32_t*>(base + L.ctrl);
  ws->counts = reinterpret_cast<uint32_t*>(base + L.cnt);
  ws->tables = reinterpret_cast<uint32_t*>(base + L.tab);
  ws->tree_list = reinterpret_cast<uint32_t*>(base + L.tl);
  ws->tree_sums = reinterpret_cast<uint32_t*>(base + L.ts);
  ws->far_item = reinterpret_cast<uint2*>(base + L.fi);
  ws->bucket_table = reinterpret_cast<uint32_t*>(base + L.bkt);
  ws->bucket_start = reinterpret_cast<uint32_t*>(base + L.bks);
  ws->bucket_ctl = reinterpret_cast<uint32_t*>(base + L.bkc);
}

// The allocation-free build's workspace: a static device array of the
// library, laid out as a one-cloud workspace of RWH_NOALLOC_MAX_POINTS
// points (each GPU holds its own copy until the process ends).  Host code
// never names it directly: its device address comes from
// hipGetSymbolAddress, once per device.  Zero at load, as a fresh workspace
// after its memset; every launch leaves it consistent for the next one
// (self-resetting look-back counters, histogram copies cleared by the digit
// scan).
constexpr WsLayout kNoallocLayout = ws_layout(RWH_NOALLOC_MAX_POINTS, 1);
// the sort tile shape of the allocation-free build (1024 x 16)
constexpr int kNoallocSortVariant = 5;
static_assert(kSortVariants[kNoallocSortVariant].block == 1024 && kSortVariants[kNoallocSortVariant].kpt == 16,
              "the allocation-free build runs the 1024 x 16 sort tile");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoallocLayout.bytes];

namespace {

int enqueue_sort(rwh_workspace* ws, const uint32_t* in, uint32_t* out, SegArgs seg, uint32_t n_segs,
                 uint32_t n_index, uint32_t key_bits, bool plain, hipStream_t st,
                 bool hist_ready = false, const uint32_t* skip = nullptr) {
  SortLaunch L{};
  L.keys_in = in;
  L.keys_out = out;
  L.keys_tmp = ws->keys_tmp;
  L.seg = seg;
  L.n_segs = n_segs;
  L.n_index = n_index;
  L.variant = ws->sort_variant;
  if (plain) {
    L.hist_unit_keys = kPlainHistUnit;
  } else {
    L.hist_unit_keys = kChunk;
  }
  L.n_hist_units = (n_index + L.hist_unit_keys - 1) / L.hist_unit_keys;
  L.plan = make_plan(key_bits, ws->sort_plan);
  L.hist = ws->hist;
  L.digit_start = ws->digit_start;
  L.lookback = ws->lb_sort;
  L.ctrl = ws->ctrl + 0 * kCtrlWords;
  L.hist_ready = hist_ready;
  L.tile_counts = ws->tile_counts;  // used when the sort is one cloud
  L.skip = skip;
  if (n_index == 0) return RWH_OK;
  launch_sort(L, st);
  return launched();
}

int check_ws(const rwh_workspace* ws, uint32_t n) {
  if (!ws) return fail(RWH_ERR_INVALID_ARG, "null workspace");
  if (ws->magic != ws_magic())
    return fail(RWH_ERR_INVALID_ARG, "not a workspace of this build of the library");
  if (n > ws->max_points)
    return fail(RWH_ERR_CAPACITY, "n = " + std::to_string(n) + " exceeds the workspace capacity " +
                                      std::to_string(ws->max_points));
  return RWH_OK;
}

OctArgs oct_args(const rwh_tree_buffers* b, float min_coord, float range) {
  OctArgs a{};
  a.edge_offset = b->edge_offset;
  a.edge_count = b->edge_count;
  a.codes = b->morton_keys;
  a.prefix_n = b->prefix_n;
  a.parent = b->parents;
  a.has_leaf_left = b->has_leaf_left;
  a.has_leaf_right = b->has_leaf_right;
  a.left_child = b->left_child;
  a.single_cap = b->oct_capacity;
  a.min_coord = min_coord;
  a.range = range;
  a.children = b->children;
  a.corner = reinterpret_cast<float4*>(b->corner);
  a.cell_size = b->cell_size;
  a.child_node_mask = b->child_node_mask;
  a.child_leaf_mask = b->child_leaf_mask;
  return a;
}

int check_bufs(const rwh_tree_buffers* b) {
  RWH_CHECK(b, "null buffers");
  RWH_CHECK(b->points && b->morton_keys && b->morton_keys_alt && b->prefix_n && b->has_leaf_left &&
                b->has_leaf_right && b->left_child && b->parents && b->edge_count &&
                b->edge_offset && b->children && b->corner && b->cell_size &&
                b->child_node_mask && b->child_leaf_mask,
            "a tree buffer is null");
  RWH_CHECK(aligned16(b->points) && aligned16(b->morton_keys) && aligned16(b->morton_keys_alt) &&
                aligned16(b->prefix_n) && aligned16(b->has_leaf_left) &&
                aligned16(b->has_leaf_right) && aligned16(b->left_child) && aligned16(b->parents) &&
                aligned16(b->edge_count) && aligned16(b->edge_offset) && aligned16(b->children) &&
                aligned16(b->corner),
            "tree buffers must be 16-byte aligned");
  return RWH_OK;
}

// The whole build of one cloud on a workspace (or a view of the scratch
// array); n_unique, n_octree_nodes and the status word go to d_nu, d_no and
// d_status (device).  marks (or nullptr): RWH_BUILD_MARKS events recorded
// at the stage boundaries.
int build_single(rwh_workspace* ws, const rwh_tree_buffers* bufs, uint32_t n, float min_coord, float range,
                 uint32_t* d_nu, uint32_t* d_no, uint32_t* d_status, hipStream_t st,
                 const rwh_event_t* marks = nullptr);

// Per-device state of the allocation-free build.
constexpr int kNoallocDevices = 64;
struct NoallocDevice {
  char* base = nullptr;  // the scratch array's address on this device
  bool dirty = false;       // the last call stopped between two launches
};
std::mutex g_noalloc_mu;
NoallocDevice g_noalloc[kNoallocDevices];

}  // namespace

extern "C" {

const char* rwh_error_string(int status) {
  switch (status) {
    case RWH_OK: return "ok";
    case RWH_ERR_INVALID_ARG: return "invalid argument";
    case RWH_ERR_HIP: return "HIP runtime error";
    case RWH_ERR_CAPACITY: return "capacity exceeded";
    case RWH_ERR_NO_DEVICE: return "no GPU device";
    case RWH_ERR_INJECTED: return "injected failure";
    case RWH_ERR_OVERFLOW: return "a kernel bound was exceeded (malformed input)";
    default: return "unknown status";
  }
}

const char* rwh_last_error(void) { return g_detail.c_str(); }
int rwh_version(void) { return 10000; }
void rwh_debug_fail_next(int stage) { g_fail_stage = stage; }

// ---- plumbing ---------------------------------------------------------------
int rwh_device_count(int* count) {
  RWH_CHECK(count, "null count");
  *count = 0;
  const hipError_t e = hipGetDeviceCount(count);
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) {
    *count = 0;
    return RWH_OK;
  }
  RWH_HIP(e);
  return RWH_OK;
}

int rwh_device_name(char* buf, size_t len) {
  RWH_CHECK(buf && len > 0, "bad buffer");
  int dev = 0;
  RWH_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  RWH_HIP(hipGetDeviceProperties(&prop, dev));
  std::snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  return RWH_OK;
}

int rwh_malloc(void** ptr, size_t bytes) {
  RWH_CHECK(ptr, "null out pointer");
  *ptr = nullptr;
  if (bytes == 0) return RWH_OK;
  RWH_HIP(hipMalloc(ptr, bytes));
  return RWH_OK;
}
int rwh_free(void* ptr) {
  if (ptr) RWH_HIP(hipFree(ptr));
  return RWH_OK;
}
int rwh_host_alloc(void** ptr, size_t bytes) {
  RWH_CHECK(ptr, "null out pointer");
  *ptr = nullptr;
  if (bytes == 0) return RWH_OK;
  RWH_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
  return RWH_OK;
}
int rwh_host_free(void* ptr) {
  if (ptr) RWH_HIP(hipHostFree(ptr));
  return RWH_OK;
}
int rwh_memcpy_h2d(void* dst, const void* src, size_t bytes, rwh_stream_t stream) {
  if (bytes == 0) return RWH_OK;
  RWH_CHECK(dst && src, "null pointer");
  RWH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)));
  return RWH_OK;
}
int rwh_memcpy_d2h(void* dst, const void* src, size_t bytes, rwh_stream_t stream) {
  if (bytes == 0) return RWH_OK;
  RWH_CHECK(dst && src, "null pointer");
  RWH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, static_cast<hipStream_t>(stream)));
  return RWH_OK;
}
int rwh_memcpy_d2d(void* dst, const void* src, size_t bytes, rwh_stream_t stream) {
  if (bytes == 0) return RWH_OK;
  RWH_CHECK(dst && src, "null pointer");
  RWH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)));
  return RWH_OK;
}
int rwh_memset(void* dst, int value, size_t bytes, rwh_stream_t stream) {
  if (bytes == 0) return RWH_OK;
  RWH_CHECK(dst, "null pointer");
  RWH_HIP(hipMemsetAsync(dst, value, bytes, static_cast<hipStream_t>(stream)));
  return RWH_OK;
}
int rwh_stream_create(rwh_stream_t* stream) {
  RWH_CHECK(stream, "null out pointer");
  hipStream_t s = nullptr;
  RWH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return RWH_OK;
}
int rwh_stream_destroy(rwh_stream_t stream) {
  if (stream) RWH_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream)));
  return RWH_OK;
}
int rwh_stream_sync(rwh_stream_t stream) {
  RWH_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return RWH_OK;
}
int rwh_device_sync(void) {
  RWH_HIP(hipDeviceSynchronize());
  return RWH_OK;
}
int rwh_event_create(rwh_event_t* event) {
  RWH_CHECK(event, "null out pointer");
  hipEvent_t e = nullptr;
  RWH_HIP(hipEventCreate(&e));
  *event = e;
  return RWH_OK;
}
int rwh_event_destroy(rwh_event_t event) {
  if (event) RWH_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
  return RWH_OK;
}
int rwh_event_record(rwh_event_t event, rwh_stream_t stream) {
  RWH_CHECK(event, "null event");
  RWH_HIP(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
  return RWH_OK;
}
int rwh_event_sync(rwh_event_t event) {
  RWH_CHECK(event, "null event");
  RWH_HIP(hipEventSynchronize(static_cast<hipEvent_t>(event)));
  return RWH_OK;
}
int rwh_event_query(rwh_event_t event) {
  RWH_CHECK(event, "null event");
  const hipError_t e = hipEventQuery(static_cast<hipEvent_t>(event));
  if (e == hipErrorNotReady) return 1;
  RWH_HIP(e);
  return RWH_OK;
}
int rwh_event_elapsed_ms(float* ms, rwh_event_t start, rwh_event_t stop) {
  RWH_CHECK(ms && start && stop, "null argument");
  RWH_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
  return RWH_OK;
}

// ---- workspace ---------------------------------------------------------------
int rwh_workspace_create(uint32_t max_points, uint32_t max_segments, rwh_workspace** out) {
  RWH_CHECK(out, "null out pointer");
  *out = nullptr;
  RWH_CHECK(max_points <= kMaxPoints, "max_points above 2^30");
  RWH_CHECK(max_segments >= 1 && max_segments <= (1u << 20), "max_segments out of range");
  auto* ws = new rwh_workspace();
  const WsLayout L = ws_layout(max_points, max_segments);
  char* base = nullptr;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&base), L.bytes);
  if (e != hipSuccess) {
    delete ws;
    return fail(e == hipErrorNoDevice ? RWH_ERR_NO_DEVICE : RWH_ERR_HIP,
                std::string("hipMalloc(workspace): ") + hipGetErrorString(e));
  }
  ws->block = base;
  bind_workspace(ws, base, L, max_points, max_segments);
  // Look-back words, histograms and counters must start at zero; later
  // launches keep them consistent by themselves.
  e = hipMemset(base, 0, L.bytes);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&ws->h_stage),
                                         sizeof(uint32_t) * ws->table_words(), hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&ws->h_counts),
                                         sizeof(uint32_t) * 3ull * max_segments, hipHostMallocDefault);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ws->upload_done, hipEventDisableTiming);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    rwh_workspace_destroy(ws);
    return fail(RWH_ERR_HIP, std::string("workspace setup: ") + hipGetErrorString(e));
  }
  *out = ws;
  return RWH_OK;
}

int rwh_workspace_destroy(rwh_workspace* ws) {
  if (!ws) return RWH_OK;
  if (ws->upload_done) (void)hipEventDestroy(ws->upload_done);
  if (ws->h_stage) (void)hipHostFree(ws->h_stage);
  if (ws->h_counts) (void)hipHostFree(ws->h_counts);
  const hipError_t e = ws->block ? hipFree(ws->block) : hipSuccess;
  delete ws;
  RWH_HIP(e);
  return RWH_OK;
}

size_t rwh_workspace_bytes(const rwh_workspace* ws) { return ws ? ws->bytes : 0; }

int rwh_sort_variant_count(void) { return kNumSortVariants; }

int rwh_sort_variant_shape(int variant, uint32_t* block, uint32_t* keys_per_thread) {
  RWH_CHECK(variant >= 0 && variant < kNumSortVariants, "no such sort variant");
  RWH_CHECK(block && keys_per_thread, "null argument");
  *block = static_cast<uint32_t>(kSortVariants[variant].block);
  *keys_per_thread = static_cast<uint32_t>(kSortVariants[variant].kpt);
  return RWH_OK;
}

int rwh_workspace_set_sort_variant(rwh_workspace* ws, int variant) {
  RWH_CHECK(ws, "null workspace");
  RWH_CHECK(variant >= 0 && variant < kNumSortVariants, "no such sort variant");
  ws->sort_variant = variant;
  return RWH_OK;
}
int rwh_sort_plan_count(void) { return kNumSortPlans; }

int rwh_sort_plan_widths(int plan, uint32_t* widths) {
  RWH_CHECK(plan >= 0 && plan < kNumSortPlans, "no such sort plan");
  RWH_CHECK(widths, "null argument");
  for (uint32_t p = 0; p < kMaxPasses; ++p) widths[p] = kPlanWidths[plan][p];
  return RWH_OK;
}

int rwh_workspace_set_sort_plan(rwh_workspace* ws, int plan) {
  RWH_CHECK(ws, "null workspace");
  RWH_CHECK(plan >= 0 && plan < kNumSortPlans, "no such sort plan");
  ws->sort_plan = plan;
  return RWH_OK;
}

int rwh_workspace_set_bucket_sort(rwh_workspace* ws, int enable) {
  RWH_CHECK(ws, "null workspace");
  RWH_CHECK(enable == 0 || enable == 1, "enable must be 0 or 1");
  ws->bucket_sort = enable == 1;
  return RWH_OK;
}

int rwh_workspace_sort_used_buckets(rwh_workspace* ws, rwh_stream_t stream, int* used) {
  RWH_CHECK(ws && used, "null argument");
  *used = 0;
  if (!ws->bucket_tried) return RWH_OK;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  RWH_HIP(hipMemcpyAsync(ws->h_counts, ws->bucket_ctl, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  RWH_HIP(hipStreamSynchronize(st));
  *used = ws->h_counts[0] != 0 ? 1 : 0;
  return RWH_OK;
}

uint32_t* rwh_workspace_counts(rwh_workspace* ws) { return ws ? ws->counts : nullptr; }
uint32_t rwh_workspace_max_segments(const rwh_workspace* ws) { return ws ? ws->max_segs : 0; }

// ---- stages -------------------------------------------------------------------
int rwh_morton(const float* d_xyzw, uint32_t n, float min_coord, float range, uint32_t* d_codes,
               rwh_stream_t stream) {
  if (injected(1)) return RWH_ERR_INJECTED;
  RWH_CHECK(n <= kMaxPoints, "n above 2^30");
  if (n == 0) return RWH_OK;
  RWH_CHECK(d_xyzw && d_codes, "null buffer");
  RWH_CHECK(aligned16(d_xyzw), "points must be 16-byte aligned");
  launch_morton(reinterpret_cast<const float4*>(d_xyzw), d_codes, single_seg(n), n, min_coord,
                range, nullptr, static_cast<hipStream_t>(stream));
  return launched();
}

int rwh_sort_u32(rwh_workspace* ws, const uint32_t* d_in, uint32_t* d_out, uint32_t n,
                 uint32_t key_bits, rwh_stream_t stream) {
  if (injected(2)) return RWH_ERR_INJECTED;
  RWH_TRY(check_ws(ws, n));
  RWH_CHECK(key_bits >= 1 && key_bits <= 32, "key_bits must be in [1, 32]");
  if (n == 0) return RWH_OK;
  RWH_CHECK(d_in && d_out, "null buffer");
  RWH_CHECK(aligned16(d_in) && aligned16(d_out), "keys must be 16-byte aligned");
  RWH_CHECK(d_in != d_out, "in-place sort is not supported");
  return enqueue_sort(ws, d_in, d_out, single_seg(n), 1, n, key_bits, true,
                      static_cast<hipStream_t>(stream));
}

int rwh_unique_u32(rwh_workspace* ws, const uint32_t* d_sorted, uint32_t n, uint32_t* d_unique,
                   uint32_t* d_n_unique, rwh_stream_t stream) {
  if (injected(3)) return RWH_ERR_INJECTED;
  RWH_TRY(check_ws(ws, n));
  RWH_CHECK(d_n_unique, "null count pointer");
  const hipStream_t st = static_cast<hipStream_t>(stream);
  if (n == 0) {
    launch_fill_u32(d_n_unique, 0, 1, st);
    return launched();
  }
  RWH_CHECK(d_sorted && d_unique, "null buffer");
  RWH_CHECK(aligned16(d_sorted), "keys must be 16-byte aligned");
  launch_unique(d_sorted, d_unique, single_seg(n), n, d_n_unique, ws->lb_unique,
                ws->ctrl + 1 * kCtrlWords, st);
  return launched();
}

int rwh_radix_tree(rwh_workspace* ws, const uint32_t* d_codes, const uint32_t* d_n_unique, uint32_t n_max,
                   uint8_t* d_prefix_n, uint8_t* d_has_leaf_left, uint8_t* d_has_leaf_right,
                   int32_t* d_left_child, int32_t* d_parents, rwh_stream_t stream) {
  if (injected(4)) return RWH_ERR_INJECTED;
  RWH_CHECK(n_max <= kMaxPoints, "n_max above 2^30");
  RWH_CHECK(ws, "null workspace");
  RWH_CHECK(n_max <= ws->max_points, "n_max above the workspace's capacity");
  if (n_max == 0) return RWH_OK;
  RWH_CHECK(d_codes && d_n_unique && d_prefix_n && d_has_leaf_left && d_has_leaf_right &&
                d_left_child && d_parents,
            "null buffer");
  RWH_CHECK(aligned16(d_prefix_n) && aligned16(d_has_leaf_left) && aligned16(d_has_leaf_right) &&
                aligned16(d_left_child) && aligned16(d_parents),
            "node buffers must be 16-byte aligned");
  TreeArgs a{d_codes, d_prefix_n, d_has_leaf_left, d_has_leaf_right, d_left_child, d_parents};
  a.status = ws->status();
  const hipStream_t st = static_cast<hipStream_t>(stream);
  launch_fill_u32(a.status, 0, 1, st);
  launch_radix_tree(a, single_seg_dev(d_n_unique), n_max, ws->tree_list, st);
  return launched();
}

int rwh_edge_count(const uint8_t* d_prefix_n, const int32_t* d_parents, const uint32_t* d_n_unique,
                   uint32_t n_max, int32_t* d_edge_count, rwh_stream_t stream) {
  if (injected(5)) return RWH_ERR_INJECTED;
  RWH_CHECK(n_max <= kMaxPoints, "n_max above 2^30");
  if (n_max == 0) return RWH_OK;
  RWH_CHECK(d_prefix_n && d_parents && d_n_unique && d_edge_count, "null buffer");
  launch_edge_count(d_prefix_n, d_parents, d_edge_count, single_seg_dev(d_n_unique), n_max,
                    static_cast<hipStream_t>(stream));
  return launched();
}

int rwh_prefix_sum(rwh_workspace* ws, const int32_t* d_edge_count, const uint32_t* d_n_unique,
                   uint32_t n_max, int32_t* d_edge_offset, uint32_t* d_n_octree_nodes,
                   rwh_stream_t stream) {
  if (injected(6)) return RWH_ERR_INJECTED;
  RWH_TRY(check_ws(ws, n_max));
  RWH_CHECK(d_n_unique && d_n_octree_nodes, "null count pointer");
  const hipStream_t st = static_cast<hipStream_t>(stream);
  if (n_max == 0) {
    launch_fill_u32(d_n_octree_nodes, 0, 1, st);
    return launched();
  }
  RWH_CHECK(d_edge_count && d_edge_offset, "null buffer");
  RWH_CHECK(aligned16(d_edge_count) && aligned16(d_edge_offset), "buffers must be 16-byte aligned");
  launch_prefix_sum(d_edge_count, d_edge_offset, single_seg_dev(d_n_unique), n_max,
                    d_n_octree_nodes, ws->lb_scan, ws->ctrl + 2 * kCtrlWords, st);
  return launched();
}

int rwh_octree(rwh_workspace* ws, const rwh_tree_buffers* bufs, const uint32_t* d_n_unique,
               const uint32_t* d_n_octree_nodes, uint32_t n_max, float min_coord, float range,
               uint32_t* d_status, rwh_stream_t stream) {
  if (injected(7)) return RWH_ERR_INJECTED;
  RWH_TRY(check_bufs(bufs));
  RWH_CHECK(d_n_unique && d_n_octree_nodes, "null count pointer");
  RWH_CHECK(n_max <= kMaxPoints, "n_max above 2^30");
  if (ws) RWH_TRY(check_ws(ws, n_max));
  if (n_max == 0) return RWH_OK;
  OctArgs a = oct_args(bufs, min_coord, range);
  a.n_oct = d_n_octree_nodes;
  a.status = d_status;
  if (ws) a.far_item = ws->far_item;
  launch_octree(a, single_seg_dev(d_n_unique), n_max, true, static_cast<hipStream_t>(stream));
  return launched();
}

int rwh_link_leaf(const rwh_tree_buffers* bufs, const uint32_t* d_n_unique,
                  const uint32_t* d_n_octree_nodes, uint32_t n_max, uint32_t* d_status,
                  rwh_stream_t stream) {
  RWH_TRY(check_bufs(bufs));
  RWH_CHECK(d_n_unique && d_n_octree_nodes, "null count pointer");
  RWH_CHECK(n_max <= kMaxPoints, "n_max above 2^30");
  if (n_max == 0) return RWH_OK;
  OctArgs a = oct_args(bufs, 0.0f, 1.0f);
  a.n_oct = d_n_octree_nodes;
  a.status = d_status;
  launch_link_leaf(a, single_seg_dev(d_n_unique), n_max, static_cast<hipStream_t>(stream));
  return launched();
}

int rwh_build_octree(rwh_workspace* ws, const rwh_tree_buffers* bufs, uint32_t n, float min_coord,
                     float range, rwh_stream_t stream) {
  RWH_TRY(check_ws(ws, n));
  return build_single(ws, bufs, n, min_coord, range, ws->n_unique(), ws->n_oct(), ws->status(),
                      static_cast<hipStream_t>(stream));
}

int rwh_build_octree_marked(rwh_workspace* ws, const rwh_tree_buffers* bufs, uint32_t n, float min_coord,
                            float range, const rwh_event_t* marks, rwh_stream_t stream) {
  RWH_TRY(check_ws(ws, n));
  RWH_CHECK(marks, "null marks");
  for (int k = 0; k < RWH_BUILD_MARKS; ++k) RWH_CHECK(marks[k], "null event in marks");
  return build_single(ws, bufs, n, min_coord, range, ws->n_unique(), ws->n_oct(), ws->status(),
                      static_cast<hipStream_t>(stream), marks);
}

int rwh_build_octree_noalloc(const rwh_tree_buffers* bufs, uint32_t n, float min_coord, float range,
                             uint32_t* d_counts, rwh_stream_t stream) {
  RWH_CHECK(d_counts, "null counts pointer");
  if (n > RWH_NOALLOC_MAX_POINTS)
    return fail(RWH_ERR_CAPACITY, "n = " + std::to_string(n) + " exceeds RWH_NOALLOC_MAX_POINTS (" +
                                      std::to_string(RWH_NOALLOC_MAX_POINTS) + ")");
  RWH_TRY(check_bufs(bufs));
  int dev = 0;
  RWH_HIP(hipGetDevice(&dev));
  RWH_CHECK(dev >= 0 && dev < kNoallocDevices, "device index above the library's table");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  NoallocDevice& d = g_noalloc[dev];
  if (!d.base) {
    // set-up, once per device with that device current: the scratch array's
    // address on this device
    void* p = nullptr;
    RWH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    d.base = static_cast<char*>(p);
  }
  rwh_workspace ws;  // a view of the scratch array: nothing allocated
  bind_workspace(&ws, d.base, kNoallocLayout, RWH_NOALLOC_MAX_POINTS, 1);
  ws.sort_variant = kNoallocSortVariant;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  if (d.dirty) {
    // an earlier call stopped between launches: its histogram copies may
    // hold counts the digit scan never cleared
    launch_fill_u32(ws.hist, 0, kHistCopies * kMaxPasses * 256, st);
    RWH_TRY(launched());
    d.dirty = false;
  }
  const int rc = build_single(&ws, bufs, n, min_coord, range, d_counts, d_counts + 1, d_counts + 2, st);
  if (rc != RWH_OK) d.dirty = true;
  return rc;
}

}  // extern "C"

namespace {

int build_single(rwh_workspace* ws, const rwh_tree_buffers* bufs, uint32_t n, float min_coord, float range,
                 uint32_t* d_nu, uint32_t* d_no, uint32_t* d_status, hipStream_t st,
                 const rwh_event_t* marks) {
  RWH_TRY(check_ws(ws, n));
  RWH_TRY(check_bufs(bufs));
  // the caller's stage-boundary events (measurement only)
  const auto mark = [&](int k) -> int {
    if (marks) RWH_HIP(hipEventRecord(static_cast<hipEvent_t>(marks[k]), st));
    return RWH_OK;
  };
  RWH_TRY(mark(0));
  // S1 fused with the sort's histogram sweep, then the digit passes
  if (injected(1) || injected(2)) return RWH_ERR_INJECTED;
  RWH_CHECK(bufs->morton_keys != bufs->morton_keys_alt, "the two key buffers must differ");
  const BucketPlan bp = bucket_plan(n);
  ws->bucket_tried = n > 0 && ws->bucket_sort && bp.ok;
  if (ws->bucket_tried) {
    // S1 with the bucket counts, the bucket sort, and the digit-pass sort
    // (its own histogram sweep first) gated on its flag: it runs only when a
    // bucket is too large for LDS
    launch_morton_bucket(reinterpret_cast<const float4*>(bufs->points), bufs->morton_keys, n, min_coord, range,
                         bp, ws->bucket_table, d_status, st);
    RWH_TRY(launched());
    RWH_TRY(mark(1));
    const BucketLaunch B{bufs->morton_keys, bufs->morton_keys_alt, ws->keys_tmp, n, bp, ws->bucket_table,
                         ws->bucket_start + kBucketMaxBins + 1, ws->bucket_start, ws->bucket_ctl};
    launch_bucket_sort(B, st);
    RWH_TRY(launched());
    RWH_TRY(enqueue_sort(ws, bufs->morton_keys, bufs->morton_keys_alt, single_seg(n), 1, n, 30, true,
                         st, false, ws->bucket_ctl));
  } else if (n > 0) {
    launch_morton_hist(reinterpret_cast<const float4*>(bufs->points), bufs->morton_keys, single_seg(n), n,
                       kPlainHistUnit, min_coord, range, make_plan(30, ws->sort_plan), ws->hist,
                       ws->sort_variant, ws->tile_counts, d_status, st);
    RWH_TRY(launched());
    RWH_TRY(mark(1));
    RWH_TRY(enqueue_sort(ws, bufs->morton_keys, bufs->morton_keys_alt, single_seg(n), 1, n, 30, true,
                         st, true));
  } else {
    RWH_TRY(mark(1));
  }
  RWH_TRY(mark(2));
  RWH_TRY(rwh_unique_u32(ws, bufs->morton_keys_alt, n, bufs->morton_keys, d_nu, st));
  RWH_TRY(mark(3));
  if (injected(4) || injected(5)) return RWH_ERR_INJECTED;
  if (n == 0) {
    launch_fill_u32(d_no, 0, 1, st);
    launch_fill_u32(d_status, 0, 1, st);
    RWH_TRY(launched());
  } else {
    // S4 with S5: the tree kernels write the edge counts beside the tree
    // (Morton codes are below 2^31) and their sums per tree tile; S6 scans
    // the counts, each tile's prefix from those sums
    TreeArgs ta{bufs->morton_keys, bufs->prefix_n, bufs->has_leaf_left, bufs->has_leaf_right,
                bufs->left_child, bufs->parents, bufs->edge_count, ws->tree_sums, d_status};
    launch_radix_tree(ta, single_seg_dev(d_nu), n, ws->tree_list, st);
    RWH_TRY(launched());
    RWH_TRY(mark(4));
    RWH_TRY(mark(5));
    launch_prefix_sum_tiles(bufs->edge_count, bufs->edge_offset, single_seg_dev(d_nu), n, d_no, ws->tree_sums, st);
    RWH_TRY(launched());
  }
  if (n == 0) {
    RWH_TRY(mark(4));
    RWH_TRY(mark(5));
  }
  RWH_TRY(mark(6));
  // S7 on the tree just built (its kernels checked their own bounds)
  OctArgs oa = oct_args(bufs, min_coord, range);
  oa.n_oct = d_no;
  oa.status = d_status;
  oa.far_item = ws->far_item;
  if (injected(7)) return RWH_ERR_INJECTED;
  if (n > 0) launch_octree(oa, single_seg_dev(d_nu), n, false, st);
  RWH_TRY(launched());
  return mark(7);
}

}  // namespace

extern "C" {

int rwh_read_counts(rwh_workspace* ws, uint32_t oct_capacity, uint32_t* counts3,
                    rwh_stream_t stream) {
  RWH_CHECK(ws && counts3, "null argument");
  const hipStream_t st = static_cast<hipStream_t>(stream);
  RWH_HIP(hipMemcpyAsync(ws->h_counts, ws->n_unique(), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  RWH_HIP(hipMemcpyAsync(ws->h_counts + 1, ws->n_oct(), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  RWH_HIP(hipMemcpyAsync(ws->h_counts + 2, ws->status(), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  RWH_HIP(hipStreamSynchronize(st));
  counts3[0] = ws->h_counts[0];
  counts3[1] = ws->h_counts[0] > 0 ? ws->h_counts[0] - 1 : 0;
  counts3[2] = ws->h_counts[1];
  if (ws->h_counts[2] != 0)
    return fail(RWH_ERR_OVERFLOW, "build status " + std::to_string(ws->h_counts[2]) + ": " +
                                      status_text(ws->h_counts[2]));
  if (counts3[2] > oct_capacity)
    return fail(RWH_ERR_CAPACITY, "n_octree_nodes = " + std::to_string(counts3[2]) +
                                      " exceeds the octree capacity " + std::to_string(oct_capacity));
  return RWH_OK;
}

// ---- batched build ------------------------------------------------------------
int rwh_batch_layout(const uint32_t* n, uint32_t n_clouds, uint32_t* point_base, uint32_t* oct_base,
                     uint32_t* oct_cap, uint32_t* index_space, uint32_t* oct_space) {
  RWH_CHECK(n && point_base && oct_base && oct_cap && index_space && oct_space, "null argument");
  uint64_t p = 0, o = 0;
  for (uint32_t s = 0; s < n_clouds; ++s) {
    RWH_CHECK(n[s] <= kMaxPoints, "cloud above 2^30 points");
    point_base[s] = static_cast<uint32_t>(p);
    const uint64_t chunks = std::max<uint64_t>(1, (n[s] + kChunk - 1) / kChunk);
    p += chunks * kChunk;
    // the reference sizes the octree at n * 0.6f (app_data.hpp:13, 28)
    const uint32_t cap = static_cast<uint32_t>(static_cast<float>(n[s]) * 0.6f);
    oct_base[s] = static_cast<uint32_t>(o);
    oct_cap[s] = cap;
    o += align_up(cap, 64);
    RWH_CHECK(p <= kMaxPoints && o <= kMaxPoints, "batch index space above 2^30");
  }
  *index_space = static_cast<uint32_t>(p);
  *oct_space = static_cast<uint32_t>(o);
  return RWH_OK;
}

namespace {

int enqueue_batch(rwh_workspace* ws, const rwh_tree_buffers* bufs, float min_coord, float range,
                  hipStream_t st) {
  const uint32_t n_clouds = ws->batch_clouds, n_index = ws->batch_index;
  SegArgs pts{};
  pts.n_segs = n_clouds;
  pts.base = ws->seg_base();
  pts.count = ws->seg_n();
  pts.chunk_seg = ws->chunk_seg();
  SegArgs uni = pts;
  uni.count = ws->n_unique();

  // S1 fused with the sort's histogram sweep (chunk-sized units, each in one
  // cloud), then the digit passes
  launch_morton_hist(reinterpret_cast<const float4*>(bufs->points), bufs->morton_keys, pts, n_index, kChunk,
                     min_coord, range, make_plan(30, ws->sort_plan), ws->hist, ws->sort_variant, nullptr,
                     ws->status(), st);
  RWH_TRY(launched());
  RWH_TRY(enqueue_sort(ws, bufs->morton_keys, bufs->morton_keys_alt, pts, n_clouds, n_index, 30,
                       false, st, true));
  launch_unique(bufs->morton_keys_alt, bufs->morton_keys, pts, n_index, ws->n_unique(),
                ws->lb_unique, ws->ctrl + 1 * kCtrlWords, st);
  RWH_TRY(launched());
  TreeArgs ta{bufs->morton_keys, bufs->prefix_n, bufs->has_leaf_left, bufs->has_leaf_right,
              bufs->left_child, bufs->parents, bufs->edge_count, ws->tree_sums, ws->status()};
  launch_radix_tree(ta, uni, n_index, ws->tree_list, st);
  launch_prefix_sum_tiles(bufs->edge_count, bufs->edge_offset, uni, n_index, ws->n_oct(), ws->tree_sums, st);
  OctArgs oa = oct_args(bufs, min_coord, range);
  oa.n_oct = ws->n_oct();
  oa.oct_base = ws->oct_base();
  oa.oct_cap = ws->oct_cap();
  oa.status = ws->status();
  oa.far_item = ws->far_item;
  launch_octree(oa, uni, n_index, false, st);
  return launched();
}

// Validates a batch layout and stages its tables; the copy to the device is
// enqueued on st (sync: also waited for).
int stage_batch(rwh_workspace* ws, const uint32_t* n, const uint32_t* point_base,
                const uint32_t* oct_base, const uint32_t* oct_cap, uint32_t n_clouds,
                hipStream_t st, bool sync) {
  RWH_CHECK(ws, "null workspace");
  RWH_CHECK(n && point_base && oct_base && oct_cap, "null layout array");
  RWH_CHECK(n_clouds >= 1, "empty batch");
  if (n_clouds > ws->max_segs)
    return fail(RWH_ERR_CAPACITY, "batch of " + std::to_string(n_clouds) +
                                      " clouds exceeds the workspace's " + std::to_string(ws->max_segs));
  uint64_t space = 0;
  for (uint32_t s = 0; s < n_clouds; ++s) {
    RWH_CHECK(point_base[s] % kChunk == 0, "point_base must be a multiple of 8192");
    RWH_CHECK(s == 0 || point_base[s] > point_base[s - 1], "point_base must increase");
    const uint64_t chunks = std::max<uint64_t>(1, (n[s] + kChunk - 1) / kChunk);
    RWH_CHECK(s + 1 == n_clouds || point_base[s] + chunks * kChunk <= point_base[s + 1],
              "clouds overlap in the index space");
    space = std::max<uint64_t>(space, point_base[s] + chunks * kChunk);
  }
  if (space > ws->max_points)
    return fail(RWH_ERR_CAPACITY, "batch index space " + std::to_string(space) +
                                      " exceeds the workspace capacity " + std::to_string(ws->max_points));
  const uint32_t n_index = static_cast<uint32_t>(space);
  // the pinned staging area is reused only after the previous upload completed
  if (ws->upload_pending) RWH_HIP(hipEventSynchronize(ws->upload_done));
  uint32_t* h = ws->h_stage;
  const uint32_t S = ws->max_segs;
  for (uint32_t s = 0; s < n_clouds; ++s) {
    h[s] = point_base[s];
    h[S + s] = n[s];
    h[2 * S + s] = oct_base[s];
    h[3 * S + s] = oct_cap[s];
  }
  uint32_t* cs = h + 4 * S;
  const uint32_t n_chunks = n_index / kChunk;
  for (uint32_t s = 0; s < n_clouds; ++s) {
    const uint32_t c0 = point_base[s] / kChunk;
    const uint32_t c1 = (s + 1 < n_clouds) ? point_base[s + 1] / kChunk : n_chunks;
    for (uint32_t c = c0; c < c1; ++c) cs[c] = s;
  }
  RWH_HIP(hipMemcpyAsync(ws->tables, h, sizeof(uint32_t) * 4 * S, hipMemcpyHostToDevice, st));
  RWH_HIP(hipMemcpyAsync(ws->chunk_seg(), cs, sizeof(uint32_t) * n_chunks, hipMemcpyHostToDevice, st));
  RWH_HIP(hipEventRecord(ws->upload_done, st));
  ws->upload_pending = true;
  if (sync) RWH_HIP(hipEventSynchronize(ws->upload_done));
  ws->batch_clouds = n_clouds;
  ws->batch_index = n_index;
  return RWH_OK;
}

}  // namespace

int rwh_build_octree_batch(rwh_workspace* ws, const rwh_tree_buffers* bufs, const uint32_t* n,
                           const uint32_t* point_base, const uint32_t* oct_base,
                           const uint32_t* oct_cap, uint32_t n_clouds, float min_coord,
                           float range, rwh_stream_t stream) {
  RWH_TRY(check_bufs(bufs));
  const hipStream_t st = static_cast<hipStream_t>(stream);
  RWH_TRY(stage_batch(ws, n, point_base, oct_base, oct_cap, n_clouds, st, false));
  return enqueue_batch(ws, bufs, min_coord, range, st);
}

int rwh_batch_prepare(rwh_workspace* ws, const uint32_t* n, const uint32_t* point_base,
                      const uint32_t* oct_base, const uint32_t* oct_cap, uint32_t n_clouds,
                      rwh_stream_t stream) {
  return stage_batch(ws, n, point_base, oct_base, oct_cap, n_clouds,
                     static_cast<hipStream_t>(stream), true);
}

int rwh_build_octree_batch_enqueue(rwh_workspace* ws, const rwh_tree_buffers* bufs, float min_coord,
                                   float range, rwh_stream_t stream) {
  RWH_CHECK(ws, "null workspace");
  RWH_TRY(check_bufs(bufs));
  RWH_CHECK(ws->batch_clouds > 0, "no batch prepared on this workspace");
  return enqueue_batch(ws, bufs, min_coord, range, static_cast<hipStream_t>(stream));
}

int rwh_batch_graph_create(rwh_workspace* ws, const rwh_tree_buffers* bufs, float min_coord,
                           float range, rwh_stream_t stream, rwh_graph_t* graph) {
  RWH_CHECK(ws && graph, "null argument");
  RWH_TRY(check_bufs(bufs));
  RWH_CHECK(ws->batch_clouds > 0, "no batch prepared on this workspace");
  *graph = nullptr;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  RWH_CHECK(st != nullptr, "graph capture needs a non-default stream");
  RWH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_batch(ws, bufs, min_coord, range, st);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (rc != RWH_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  RWH_HIP(e);
  hipGraphExec_t ex = nullptr;
  const hipError_t e2 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  RWH_HIP(e2);
  *graph = ex;
  return RWH_OK;
}

int rwh_graph_launch(rwh_graph_t graph, rwh_stream_t stream) {
  RWH_CHECK(graph, "null graph");
  RWH_HIP(hipGraphLaunch(static_cast<hipGraphExec_t>(graph), static_cast<hipStream_t>(stream)));
  return RWH_OK;
}

int rwh_graph_destroy(rwh_graph_t graph) {
  if (graph) RWH_HIP(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph)));
  return RWH_OK;
}

int rwh_read_batch_counts(rwh_workspace* ws, uint32_t n_clouds, uint32_t* counts,
                          rwh_stream_t stream) {
  RWH_CHECK(ws && counts, "null argument");
  RWH_CHECK(n_clouds <= ws->max_segs, "n_clouds above the workspace's segments");
  const hipStream_t st = static_cast<hipStream_t>(stream);
  RWH_HIP(hipMemcpyAsync(ws->h_counts, ws->n_unique(), sizeof(uint32_t) * n_clouds,
                         hipMemcpyDeviceToHost, st));
  RWH_HIP(hipMemcpyAsync(ws->h_counts + ws->max_segs, ws->n_oct(), sizeof(uint32_t) * n_clouds,
                         hipMemcpyDeviceToHost, st));
  RWH_HIP(hipMemcpyAsync(ws->h_counts + 2 * ws->max_segs, ws->status(), sizeof(uint32_t) * n_clouds,
                         hipMemcpyDeviceToHost, st));
  RWH_HIP(hipStreamSynchronize(st));
  std::memcpy(counts, ws->h_counts, sizeof(uint32_t) * n_clouds);
  std::memcpy(counts + n_clouds, ws->h_counts + ws->max_segs, sizeof(uint32_t) * n_clouds);
  for (uint32_t s = 0; s < n_clouds; ++s) {
    const uint32_t bits = ws->h_counts[2 * ws->max_segs + s];
    if (bits != 0)
      return fail(RWH_ERR_OVERFLOW, "cloud " + std::to_string(s) + ": build status " + std::to_string(bits) +
                                        ": " + status_text(bits));
  }
  return RWH_OK;
}

// ---- packed result ----

uint64_t rwh_result_bytes(uint32_t n_input, uint32_t n_unique, uint32_t n_octree_nodes,
                          uint32_t oct_capacity, uint32_t sections) {
  if ((sections & ~RWH_SECTIONS_ALL) != 0) return 0;
  return result_layout(n_input, std::min(n_unique, n_input), std::min(n_octree_nodes, oct_capacity), sections)
      .total;
}

int rwh_pack_result(const rwh_tree_buffers* bufs, const uint32_t* d_counts, uint32_t counts_stride,
                    uint32_t n_input, uint32_t sections, void* d_out, uint64_t out_bytes,
                    rwh_stream_t stream) {
  RWH_CHECK(bufs && d_counts && d_out, "null argument");
  RWH_CHECK(sections != 0 && (sections & ~RWH_SECTIONS_ALL) == 0, "sections outside RWH_SECTIONS_ALL");
  RWH_CHECK(counts_stride >= 1, "counts_stride must be at least 1");
  RWH_CHECK(n_input <= kMaxPoints, "n_input above 2^30");
  RWH_CHECK(aligned16(d_out), "d_out must be 16-byte aligned");
  const uint32_t cap = bufs->oct_capacity;
  const uint64_t need = result_layout(n_input, n_input, cap, sections).total;
  if (out_bytes < need)
    return fail(RWH_ERR_CAPACITY, "d_out holds " + std::to_string(out_bytes) + " bytes; a result of " +
                                      std::to_string(n_input) + " points can need " + std::to_string(need));
  const void* src[RWH_SEC_COUNT] = {bufs->morton_keys_alt, bufs->morton_keys,  bufs->prefix_n,
                                    bufs->has_leaf_left,   bufs->has_leaf_right, bufs->left_child,
                                    bufs->parents,         bufs->edge_count,   bufs->edge_offset,
                                    bufs->children,        bufs->corner,       bufs->cell_size,
                                    bufs->child_node_mask, bufs->child_leaf_mask};
  PackArgs a{};
  for (int s = 0; s < RWH_SEC_COUNT; ++s) {
    if (!((sections >> s) & 1u)) continue;
    RWH_CHECK(src[s] != nullptr, "a selected section's buffer is null");
    RWH_CHECK(aligned16(src[s]), "a selected section's buffer is not 16-byte aligned");
    a.src[s] = static_cast<const unsigned char*>(src[s]);
  }
  a.counts = d_counts;
  a.stride = counts_stride;
  a.n_input = n_input;
  a.oct_capacity = cap;
  a.sections = sections;
  a.out = static_cast<unsigned char*>(d_out);
  launch_pack_result(a, need, static_cast<hipStream_t>(stream));
  return launched();
}

}  // extern "C"

namespace {

// The format checks shared by rwh_result_check and rwh_result_section.
int check_result_format(const void* buf, uint64_t bytes, rwh_result_header* out) {
  RWH_CHECK(buf, "null buffer");
  RWH_CHECK(bytes >= RWH_RESULT_HEADER_BYTES, "buffer shorter than a result header");
  rwh_result_header h;
  std::memcpy(&h, buf, sizeof h);
  RWH_CHECK(h.magic == RWH_RESULT_MAGIC, "not a packed result (magic)");
  RWH_CHECK(h.version == RWH_RESULT_VERSION,
            "packed result version " + std::to_string(h.version) + "; this library reads version " +
                std::to_string(RWH_RESULT_VERSION));
  RWH_CHECK(h.header_bytes == RWH_RESULT_HEADER_BYTES && h.reserved == 0, "header size or reserved word");
  RWH_CHECK(h.sections != 0 && (h.sections & ~RWH_SECTIONS_ALL) == 0, "sections outside RWH_SECTIONS_ALL");
  RWH_CHECK(h.n_unique <= h.n_input && h.n_brt == (h.n_unique > 0 ? h.n_unique - 1 : 0),
            "counts disagree (n_unique, n_brt, n_input)");
  RWH_CHECK(h.oct_rows == std::min(h.n_octree_nodes, h.oct_capacity) &&
                h.flags == (h.n_octree_nodes > h.oct_capacity ? RWH_RESULT_TRUNCATED : 0u),
            "octree rows or flags disagree with the counts");
  const ResultLayout L = result_layout(h.n_input, h.n_unique, h.oct_rows, h.sections);
  RWH_CHECK(h.total_bytes == L.total, "total_bytes disagrees with the counts");
  RWH_CHECK(h.total_bytes <= bytes, "buffer shorter than total_bytes (" + std::to_string(bytes) + " < " +
                                        std::to_string(h.total_bytes) + ")");
  for (int s = 0; s < RWH_SEC_COUNT; ++s) {
    const bool on = (h.sections >> s) & 1u;
    RWH_CHECK(h.offset[s] == L.offset[s] && h.count[s] == L.count[s] &&
                  h.entry_bytes[s] == (on ? result_entry_bytes(s) : 0u),
              "section " + std::to_string(s) + ": offset, count or entry size disagrees with the layout");
  }
  *out = h;
  return RWH_OK;
}

}  // namespace

extern "C" {

int rwh_result_check(const void* buf, uint64_t bytes) {
  rwh_result_header h;
  RWH_TRY(check_result_format(buf, bytes, &h));
  if (h.status != 0)
    return fail(RWH_ERR_OVERFLOW, "build status " + std::to_string(h.status) + ": " + status_text(h.status));
  if (h.flags & RWH_RESULT_TRUNCATED)
    return fail(RWH_ERR_CAPACITY, "octree of " + std::to_string(h.n_octree_nodes) + " nodes above its capacity " +
                                      std::to_string(h.oct_capacity) + "; rows cut");
  return RWH_OK;
}

int rwh_result_section(const void* buf, uint64_t bytes, int section, const void** data,
                       uint32_t* count) {
  RWH_CHECK(data && count, "null argument");
  RWH_CHECK(section >= 0 && section < RWH_SEC_COUNT, "section id outside [0, RWH_SEC_COUNT)");
  rwh_result_header h;
  RWH_TRY(check_result_format(buf, bytes, &h));
  RWH_CHECK((h.sections >> section) & 1u, "section " + std::to_string(section) + " not in this result");
  *data = static_cast<const unsigned char*>(buf) + h.offset[section];
  *count = h.count[section];
  return RWH_OK;
}

int rwh_result_header_get(const void* buf, uint64_t bytes, rwh_result_header* out) {
  RWH_CHECK(out, "null argument");
  return check_result_format(buf, bytes, out);
}

}  // extern "C"
