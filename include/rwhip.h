/*
 * rwhip — C ABI of the MI355X (gfx950) octree-build backend.
 *
 * The hot path of the reference's tree app, as seven stages on device memory:
 *   S1 Morton encode, S2 radix sort, S3 duplicate removal, S4 Karras radix
 *   tree, S5 edge count, S6 exclusive prefix sum, S7 octree nodes.
 * Every entry point takes plain device pointers and sizes, enqueues work on
 * the given stream (NULL = the default stream) and returns a status code
 * (RWH_OK = 0).  Counts that the reference reads back on the host between
 * stages (n_unique after S3, n_octree_nodes after S6) stay in device memory
 * here and are passed to later stages as device pointers, so a whole build is
 * enqueued without a host round trip.
 *
 * Each stage entry point replaces one reference dispatcher; the C++ layer in
 * redwood-comp_amd/host (redwood::hip::run_stage1..7 returning
 * std::future<void>) is the drop-in for them and calls only this ABI.
 *
 * Thread safety: a workspace serves one stream at a time.  Different
 * workspaces may run concurrently on different streams.
 */
#ifndef RWHIP_H
#define RWHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define RWH_OK 0
#define RWH_ERR_INVALID_ARG 1 /* bad pointer, size, alignment */
#define RWH_ERR_HIP 2         /* a HIP runtime call failed (detail string) */
#define RWH_ERR_CAPACITY 3    /* size above the workspace, or octree above capacity */
#define RWH_ERR_NO_DEVICE 4   /* no usable GPU */
#define RWH_ERR_INJECTED 5    /* failure injected by rwh_debug_fail_next */
#define RWH_ERR_OVERFLOW 6    /* a kernel stopped at one of its bounds: malformed input (status bits below) */

/* Status word of a build (one per cloud, next to its counts): a kernel ORs a
 * bit in when it stops at a bound instead of finishing its work.  Each bound
 * holds for every tree built from sorted unique codes; a set bit means the
 * outputs of that cloud are not valid.  RWH_ERR_OVERFLOW reports any of them. */
#define RWH_STATUS_PAIR_CAP 0x1u   /* S4: a wave's queue of leaf pairs overflowed */
#define RWH_STATUS_CLIMB 0x2u      /* S4: a climb stopped at its iteration bound */
#define RWH_STATUS_CROSS_CAP 0x4u  /* S4: a tile had more edge-crossing nodes than its list holds */
#define RWH_STATUS_KEY_ORDER 0x8u  /* S4: keys not strictly increasing (the tile was skipped) */
#define RWH_STATUS_PULL 0x10u      /* S7: a link walk ended at its step or stack bound, or left [0, n_brt) */
#define RWH_STATUS_LEAF 0x20u      /* leaf linking: a leaf or parent index outside the tree */

typedef void* rwh_stream_t; /* a hipStream_t */
typedef void* rwh_event_t;  /* a hipEvent_t */
typedef struct rwh_workspace rwh_workspace;

/* Status text; the detail of the calling thread's last error. */
const char* rwh_error_string(int status);
const char* rwh_last_error(void);
/* Library version (major*10000 + minor*100 + patch). */
int rwh_version(void);

/* ---- device, memory, streams, events (plumbing) ---------------------- */
int rwh_device_count(int* count);
int rwh_device_name(char* buf, size_t len);
int rwh_malloc(void** ptr, size_t bytes);
int rwh_free(void* ptr);
int rwh_host_alloc(void** ptr, size_t bytes); /* pinned host memory */
int rwh_host_free(void* ptr);
int rwh_memcpy_h2d(void* dst, const void* src, size_t bytes, rwh_stream_t stream);
int rwh_memcpy_d2h(void* dst, const void* src, size_t bytes, rwh_stream_t stream);
int rwh_memcpy_d2d(void* dst, const void* src, size_t bytes, rwh_stream_t stream);
int rwh_memset(void* dst, int value, size_t bytes, rwh_stream_t stream);
int rwh_stream_create(rwh_stream_t* stream);
int rwh_stream_destroy(rwh_stream_t stream);
int rwh_stream_sync(rwh_stream_t stream);
int rwh_device_sync(void);
int rwh_event_create(rwh_event_t* event);
int rwh_event_destroy(rwh_event_t event);
int rwh_event_record(rwh_event_t event, rwh_stream_t stream);
int rwh_event_sync(rwh_event_t event);
int rwh_event_query(rwh_event_t event); /* RWH_OK when complete, 1 while pending */
int rwh_event_elapsed_ms(float* ms, rwh_event_t start, rwh_event_t stop);

/* ---- workspace (the reference's cuda::ImStorage, im_storage.cu:29-40) ---
 * Temporary device storage for the sort / unique / scan look-back state and
 * the device-resident counts of up to max_segments clouds.  max_points is the
 * largest index space a call will use (a single cloud: its point count). */
int rwh_workspace_create(uint32_t max_points, uint32_t max_segments, rwh_workspace** ws);
int rwh_workspace_destroy(rwh_workspace* ws);
size_t rwh_workspace_bytes(const rwh_workspace* ws);
/* Device array of per-cloud counts written by S3 and S6: n_unique of cloud s
 * at [s], n_octree_nodes at [max_segments + s], the status word (RWH_STATUS_*)
 * at [2 * max_segments + s] (a single build is cloud 0). */
uint32_t* rwh_workspace_counts(rwh_workspace* ws);
uint32_t rwh_workspace_max_segments(const rwh_workspace* ws);

/* Tile shapes of the sort's digit passes: variant v runs workgroups of
 * block threads with keys_per_thread keys each.  A workspace uses variant 0
 * until told otherwise; the host layer's autotuner picks one per size.
 * Shapes whose tile exceeds 8192 keys serve single-cloud sorts; a batched
 * sort with one of them selected uses the 1024 x 8 shape. */
int rwh_sort_variant_count(void);
int rwh_sort_variant_shape(int variant, uint32_t* block, uint32_t* keys_per_thread);
int rwh_workspace_set_sort_variant(rwh_workspace* ws, int variant);

/* Digit plans of the sort of 30-bit keys (the Morton codes of a build, or
 * rwh_sort_u32 with key_bits = 30): plan p runs four passes whose digit
 * widths, least significant first, rwh_sort_plan_widths gives (they sum to
 * 30; plan 0 is 8, 8, 8, 6).  A workspace uses plan 0 until told otherwise;
 * the host layer's autotuner picks one per size.  Sorts of keys of any other
 * width run 8-bit passes with the remaining bits in the last one. */
int rwh_sort_plan_count(void);
int rwh_sort_plan_widths(int plan, uint32_t* widths4);
int rwh_workspace_set_sort_plan(rwh_workspace* ws, int plan);

/* The sort of a whole single-cloud build of n points, 2^20 < n <= 2^24
 * (same output as the reference's sort, cu_dispatcher.cu:528-591): with
 * enable = 1 (a workspace's default) a bucket sort, two passes over the keys
 * (the top 8..11 bits scatter the codes into buckets of about 8192 keys, one
 * workgroup sorts each bucket in LDS), which hands over to the digit passes
 * above on the device, at no host round trip, when a bucket of the input
 * exceeds 9216 keys (clustered points); enable = 0 always runs the digit
 * passes.  Other sizes, batches and rwh_sort_u32 run the digit passes. */
int rwh_workspace_set_bucket_sort(rwh_workspace* ws, int enable);
/* *used = 1 when the last whole build on ws sorted with the bucket sort, 0
 * when it ran the digit passes (bucket sort off, n out of its range, or a
 * bucket too large).  Waits for the stream. */
int rwh_workspace_sort_used_buckets(rwh_workspace* ws, rwh_stream_t stream, int* used);

/* ---- S1  (reference: cpu::run_stage1 host_dispatchers.cpp:17-41,
 *          cuda::run_stage1 cu_dispatcher.cu:501-522)
 * codes[i] = Morton code of float4 point i (x,y,z used), 10 bits per axis.
 * d_xyzw: 16-byte aligned. */
int rwh_morton(const float* d_xyzw, uint32_t n, float min_coord, float range, uint32_t* d_codes,
               rwh_stream_t stream);

/* ---- S2  (reference: cpu::run_stage2 host_dispatchers.cpp:135-177,
 *          cuda::run_stage2 cu_dispatcher.cu:528-591)
 * Ascending sort of n keys: d_in is left unchanged, d_out gets the result.
 * key_bits: significant bits of every key (30 for Morton codes, 32 for any
 * u32); it sets the number of 8-bit passes.  16-byte aligned buffers. */
int rwh_sort_u32(rwh_workspace* ws, const uint32_t* d_in, uint32_t* d_out, uint32_t n,
                 uint32_t key_bits, rwh_stream_t stream);

/* ---- S3  (reference: cpu::run_stage3 host_dispatchers.cpp:183-194,
 *          cuda::run_stage3 cu_dispatcher.cu:597-627)
 * First key of every run of equal keys, in order; *d_n_unique (device) gets
 * the count. */
int rwh_unique_u32(rwh_workspace* ws, const uint32_t* d_sorted, uint32_t n, uint32_t* d_unique,
                   uint32_t* d_n_unique, rwh_stream_t stream);

/* ---- S4  (reference: cpu::run_stage4 host_dispatchers.cpp:200-228,
 *          cuda::run_stage4 cu_dispatcher.cu:633-648)
 * Radix-tree nodes i in [0, n_unique-1) over the unique codes; n_unique is
 * read from device memory; n_max bounds it (buffers hold n_max entries, and
 * n_max is at most the workspace's max_points).  The workspace holds the
 * per-tile lists of nodes that cross a tile edge.  The five node buffers are
 * 16-byte aligned.  The codes must be sorted and unique: cloud 0's status
 * word (cleared here) gets RWH_STATUS_KEY_ORDER otherwise, and
 * rwh_read_counts reports it. */
int rwh_radix_tree(rwh_workspace* ws, const uint32_t* d_codes, const uint32_t* d_n_unique, uint32_t n_max,
                   uint8_t* d_prefix_n, uint8_t* d_has_leaf_left, uint8_t* d_has_leaf_right,
                   int32_t* d_left_child, int32_t* d_parents, rwh_stream_t stream);

/* ---- S5  (reference: cpu::run_stage5 host_dispatchers.cpp:234-257,
 *          cuda::run_stage5 cu_dispatcher.cu:654-666) */
int rwh_edge_count(const uint8_t* d_prefix_n, const int32_t* d_parents,
                   const uint32_t* d_n_unique, uint32_t n_max, int32_t* d_edge_count,
                   rwh_stream_t stream);

/* ---- S6  (reference: cpu::run_stage6 host_dispatchers.cpp:263-297,
 *          cuda::run_stage6 cu_dispatcher.cu:672-686)
 * Exclusive scan of the n_unique-1 edge counts; *d_n_octree_nodes (device)
 * gets their sum.  16-byte aligned buffers. */
int rwh_prefix_sum(rwh_workspace* ws, const int32_t* d_edge_count, const uint32_t* d_n_unique,
                   uint32_t n_max, int32_t* d_edge_offset, uint32_t* d_n_octree_nodes,
                   rwh_stream_t stream);

/* Device buffers of one build, laid out as the reference's AppData
 * (apps/tree/app_data.hpp:48-96).  oct_capacity = floor(0.6 * n) there.
 * Every array is device memory and 16-byte aligned (rwh_malloc gives 256). */
typedef struct rwh_tree_buffers {
  const float* points;        /* n x (x,y,z,w) */
  uint32_t* morton_keys;      /* n: S1 codes, then S3 unique keys */
  uint32_t* morton_keys_alt;  /* n: S2 sorted keys */
  uint8_t* prefix_n;          /* n */
  uint8_t* has_leaf_left;     /* n */
  uint8_t* has_leaf_right;    /* n */
  int32_t* left_child;        /* n */
  int32_t* parents;           /* n */
  int32_t* edge_count;        /* n */
  int32_t* edge_offset;       /* n */
  int32_t* children;          /* 8 x oct_capacity */
  float* corner;              /* 4 x oct_capacity */
  float* cell_size;           /* oct_capacity */
  int32_t* child_node_mask;   /* oct_capacity */
  int32_t* child_leaf_mask;   /* oct_capacity */
  uint32_t oct_capacity;
} rwh_tree_buffers;

/* ---- S7  (reference: cpu::run_stage7 host_dispatchers.cpp:303-336,
 *          cuda::run_stage7 cu_dispatcher.cu:692-752)
 * Octree nodes for brt nodes [1, n_unique-1); rows [0, min(n_oct, cap)) of
 * children / masks are rewritten, corner / cell_size of those slots are set.
 * d_status: a device word that gets RWH_STATUS_PULL OR-ed in when the radix
 * tree is malformed (NULL: not reported; the caller clears it).
 * ws (may be NULL; else of at least n_max points): scratch for the few link
 * walks that leave a tile's window, which a second short kernel then runs;
 * with NULL they run inside the node kernel (same output, slower). */
int rwh_octree(rwh_workspace* ws, const rwh_tree_buffers* bufs, const uint32_t* d_n_unique,
               const uint32_t* d_n_octree_nodes, uint32_t n_max, float min_coord, float range,
               uint32_t* d_status, rwh_stream_t stream);

/* Leaf linking: the reference's k_LinkLeafNodes (07_octree.cu:228-258), which
 * it never launches.  Opt-in; runs after rwh_octree.  d_status as for
 * rwh_octree (RWH_STATUS_LEAF). */
int rwh_link_leaf(const rwh_tree_buffers* bufs, const uint32_t* d_n_unique,
                  const uint32_t* d_n_octree_nodes, uint32_t n_max, uint32_t* d_status,
                  rwh_stream_t stream);

/* ---- whole build (the reference's stage loop, pipe_main_baseline.cpp:23-31)
 * Enqueues S1..S7 for n points on the stream; the counts of cloud 0 go to
 * rwh_workspace_counts(ws). */
int rwh_build_octree(rwh_workspace* ws, const rwh_tree_buffers* bufs, uint32_t n, float min_coord,
                     float range, rwh_stream_t stream);
/* The same build with the caller's events recorded on the stream at its stage
 * boundaries, for per-stage timing of the build as it runs (fused kernels and
 * the sort it picks included); the reference times its stages with one
 * timer per stage call (apps/tree/pipe_main_baseline.cpp:23-31).
 * marks[0] before S1; marks[k] after stage k (k = 1..7).  The bucket-sort
 * path counts its buckets inside S1 (the codes' pass); S5's edge counts are
 * written by the S4 kernels, so marks[4] and marks[5] are recorded together
 * (S5 then takes no time of its own). */
#define RWH_BUILD_MARKS 8
int rwh_build_octree_marked(rwh_workspace* ws, const rwh_tree_buffers* bufs, uint32_t n, float min_coord,
                            float range, const rwh_event_t* marks, rwh_stream_t stream);
/* ---- the same build without allocating (no counterpart in the reference)
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The build of
 * rwh_build_octree, with its workspace in a static device array of this
 * library (about 20 MiB per GPU, held until the process ends), so at most
 * RWH_NOALLOC_MAX_POINTS points (larger n: RWH_ERR_CAPACITY, nothing
 * enqueued).  d_counts (device, 3 words) gets {n_unique, n_octree_nodes,
 * status word (RWH_STATUS_*)}; compare n_octree_nodes with bufs->oct_capacity.  Calls on one GPU share that
 * array: calls that overlap on one GPU (two streams, two threads) race.  The
 * sort runs its 1024 x 16 tile shape. */
#define RWH_NOALLOC_MAX_POINTS (1u << 22)
int rwh_build_octree_noalloc(const rwh_tree_buffers* bufs, uint32_t n, float min_coord, float range,
                             uint32_t* d_counts, rwh_stream_t stream);

/* Waits for the stream and copies {n_unique, n_brt, n_octree_nodes} of cloud
 * 0 to host; RWH_ERR_OVERFLOW if cloud 0's status word is set (the message
 * names the bounds), else RWH_ERR_CAPACITY if n_octree_nodes exceeds the
 * capacity given. */
int rwh_read_counts(rwh_workspace* ws, uint32_t oct_capacity, uint32_t* counts3,
                    rwh_stream_t stream);

/* ---- batched build: many clouds in one pass ----------------------------
 * Cloud s has n[s] points.  rwh_batch_layout fills, for every cloud, the
 * first index in the shared point index space (point_base, aligned to 8192)
 * and in the octree slot space (oct_base, capacity oct_cap = floor(0.6 n));
 * *index_space and *oct_space are the buffer sizes to allocate. */
int rwh_batch_layout(const uint32_t* n, uint32_t n_clouds, uint32_t* point_base, uint32_t* oct_base,
                     uint32_t* oct_cap, uint32_t* index_space, uint32_t* oct_space);
/* Enqueues S1..S7 for all clouds at once (segmented kernels).  Points of cloud
 * s sit at bufs->points[4*point_base[s]..]; outputs use the same bases.  Host
 * arrays n, point_base, oct_base, oct_cap as from rwh_batch_layout.  Counts
 * go to rwh_workspace_counts(ws) per cloud. */
int rwh_build_octree_batch(rwh_workspace* ws, const rwh_tree_buffers* bufs, const uint32_t* n,
                           const uint32_t* point_base, const uint32_t* oct_base,
                           const uint32_t* oct_cap, uint32_t n_clouds, float min_coord,
                           float range, rwh_stream_t stream);
/* The same build split in two, for callers that replay it: rwh_batch_prepare
 * validates the layout and uploads the per-cloud tables to the workspace
 * (waiting for the copy); rwh_build_octree_batch_enqueue then enqueues only
 * kernels (no copy, no host wait) for the last prepared layout, so it can be
 * captured into a graph.  Replays see the tables' current contents. */
int rwh_batch_prepare(rwh_workspace* ws, const uint32_t* n, const uint32_t* point_base,
                      const uint32_t* oct_base, const uint32_t* oct_cap, uint32_t n_clouds,
                      rwh_stream_t stream);
int rwh_build_octree_batch_enqueue(rwh_workspace* ws, const rwh_tree_buffers* bufs,
                                   float min_coord, float range, rwh_stream_t stream);

/* HIP graph of rwh_build_octree_batch_enqueue for the prepared layout
 * (captured on a non-default stream), its launch and its release.  The
 * kernels reset their own look-back state, so the graph replays as is. */
typedef void* rwh_graph_t; /* a hipGraphExec_t */
int rwh_batch_graph_create(rwh_workspace* ws, const rwh_tree_buffers* bufs, float min_coord,
                           float range, rwh_stream_t stream, rwh_graph_t* graph);
int rwh_graph_launch(rwh_graph_t graph, rwh_stream_t stream);
int rwh_graph_destroy(rwh_graph_t graph);

/* Waits and copies the counts of every cloud: n_unique to counts[s],
 * n_octree_nodes to counts[n_clouds + s]; RWH_ERR_OVERFLOW if a cloud's
 * status word is set (the counts are copied either way). */
int rwh_read_batch_counts(rwh_workspace* ws, uint32_t n_clouds, uint32_t* counts,
                          rwh_stream_t stream);

/* ---- packed result: one build's outputs in one buffer (format version 1)
 * The reference exports nothing beyond printed counts (apps/tree/main.cpp:12-16);
 * its outputs are AppData's arrays (apps/tree/app_data.hpp:48-96) at the
 * lengths its count getters give (app_data.hpp:110-133).  A packed result
 * holds those arrays at exactly those lengths, so a client needs one copy to
 * the host, and says what it holds, so a reader can check it.
 *
 * Layout (little-endian, as host and GPU are): a 256-byte rwh_result_header,
 * then each selected section in section order, starting at a multiple of 256
 * bytes from the start of the buffer, `count` entries of `entry_bytes` each,
 * zero bytes up to the next section.  Counts: sorted keys n_input, unique
 * keys n_unique, radix-tree and edge arrays n_brt = max(n_unique - 1, 0),
 * octree arrays oct_rows = min(n_octree_nodes, oct_capacity). */
#define RWH_RESULT_MAGIC 0x424f5752u /* the bytes "RWOB" */
#define RWH_RESULT_VERSION 1u
#define RWH_RESULT_HEADER_BYTES 256u
#define RWH_RESULT_ALIGN 256u
enum rwh_result_section {
  RWH_SEC_SORTED_KEYS = 0,      /* u32: S2 output (u_morton_keys_alt) */
  RWH_SEC_UNIQUE_KEYS = 1,      /* u32: S3 output (u_morton_keys) */
  RWH_SEC_PREFIX_N = 2,         /* u8 */
  RWH_SEC_HAS_LEAF_LEFT = 3,    /* u8 */
  RWH_SEC_HAS_LEAF_RIGHT = 4,   /* u8 */
  RWH_SEC_LEFT_CHILD = 5,       /* i32 */
  RWH_SEC_PARENT = 6,           /* i32 */
  RWH_SEC_EDGE_COUNT = 7,       /* i32 */
  RWH_SEC_EDGE_OFFSET = 8,      /* i32 */
  RWH_SEC_CHILDREN = 9,         /* i32[8] per octree node */
  RWH_SEC_CORNER = 10,          /* f32[4] per octree node */
  RWH_SEC_CELL_SIZE = 11,       /* f32 */
  RWH_SEC_CHILD_NODE_MASK = 12, /* i32 */
  RWH_SEC_CHILD_LEAF_MASK = 13, /* i32 */
  RWH_SEC_COUNT = 14
};
#define RWH_SECTIONS_ALL 0x3fffu    /* every section */
#define RWH_SECTIONS_OCTREE 0x3e00u /* the five octree arrays */
#define RWH_RESULT_TRUNCATED 0x1u   /* flags: n_octree_nodes above oct_capacity (rows cut) */
typedef struct rwh_result_header {
  uint32_t magic;          /* RWH_RESULT_MAGIC */
  uint32_t version;        /* RWH_RESULT_VERSION */
  uint32_t header_bytes;   /* RWH_RESULT_HEADER_BYTES */
  uint32_t sections;       /* bit s set: section s present */
  uint32_t status;         /* the build's status word (RWH_STATUS_*); 0 = valid */
  uint32_t flags;          /* RWH_RESULT_* */
  uint32_t n_input, n_unique, n_brt, n_octree_nodes, oct_rows, oct_capacity;
  uint64_t total_bytes;    /* header + sections + padding */
  uint64_t offset[RWH_SEC_COUNT];      /* from the start of the buffer; 0 if absent */
  uint32_t count[RWH_SEC_COUNT];       /* entries; 0 if absent */
  uint16_t entry_bytes[RWH_SEC_COUNT]; /* 0 if absent */
  uint32_t reserved;                   /* 0 */
} rwh_result_header;

/* Size of the packed result of a build with these counts (what a client
 * copies once it knows them); the largest size any build of n_input points
 * can give is rwh_result_bytes(n_input, n_input, oct_capacity, oct_capacity,
 * sections).  0 for sections outside RWH_SECTIONS_ALL. */
uint64_t rwh_result_bytes(uint32_t n_input, uint32_t n_unique, uint32_t n_octree_nodes,
                          uint32_t oct_capacity, uint32_t sections);
/* Enqueues one kernel that writes the packed result of a finished build
 * (enqueued before it on the stream) to device memory d_out (16-byte
 * aligned, out_bytes at least the largest size above).  bufs: the cloud's
 * arrays (for cloud s of a batch, offset by point_base[s] and oct_base[s],
 * oct_capacity = oct_cap[s]); its counts are read on the device:
 * n_unique = d_counts[0], n_octree_nodes = d_counts[counts_stride], status =
 * d_counts[2 * counts_stride] (the workspace counts: d_counts =
 * rwh_workspace_counts(ws) + s, counts_stride = max_segments; the
 * allocation-free build: its 3 words, counts_stride = 1).  No host wait. */
int rwh_pack_result(const rwh_tree_buffers* bufs, const uint32_t* d_counts, uint32_t counts_stride,
                    uint32_t n_input, uint32_t sections, void* d_out, uint64_t out_bytes,
                    rwh_stream_t stream);
/* Host-side reading (no GPU needed).  rwh_result_check returns RWH_OK for a
 * well-formed result of a valid build: RWH_ERR_INVALID_ARG if the buffer is
 * not a packed result of this version or its layout disagrees with its
 * counts (rwh_last_error says what), RWH_ERR_OVERFLOW if the build's status
 * word is set, RWH_ERR_CAPACITY if the octree rows were cut.
 * rwh_result_section gives a present section's first entry and its entry
 * count after the format checks (RWH_ERR_INVALID_ARG if absent), whatever the
 * build's status.  The buffer may sit at any address (the header is read by
 * copy); a section's entries are aligned as the buffer is, plus a multiple of
 * 256 bytes. */
int rwh_result_check(const void* buf, uint64_t bytes);
int rwh_result_section(const void* buf, uint64_t bytes, int section, const void** data,
                       uint32_t* count);
/* The header of a well-formed result (the format checks above), whatever the
 * build's status. */
int rwh_result_header_get(const void* buf, uint64_t bytes, rwh_result_header* out);

/* ---- testing hook: the next call of stage k (1..7; 0 = any) fails with
 * RWH_ERR_INJECTED before enqueuing anything.  Thread-local, one shot. */
void rwh_debug_fail_next(int stage);

#ifdef __cplusplus
}
#endif

#endif
