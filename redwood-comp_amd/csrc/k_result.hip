// The packed result of one build (include/rwhip.h, format version 1): the
// outputs of AppData (apps/tree/app_data.hpp:48-96) at their live lengths,
// gathered on the device into one buffer behind a self-describing header, so
// a client copies a build to the host in one transfer.  The reference has no
// export of its results (apps/tree/main.cpp:12-16 prints the counts).
//
// A plain HBM-bound gather: each work-item moves 16-byte units (one vector
// load, one vector store) at stride blockDim, four units per work-item.
// Every array is 16-byte aligned and every section starts at a multiple of
// 256 bytes, so all full units are aligned vector accesses; the one partial
// unit at a section's end is read byte by byte (never past the array) and
// the bytes after it, up to the next section, are written as zeros.
#include "rw_internal.hpp"

namespace rw {

namespace {

constexpr int kPackBlock = 256;
constexpr int kPackSteps = 4;  // 16-byte units per work-item
constexpr uint64_t kPackUnitsPerBlock = static_cast<uint64_t>(kPackBlock) * kPackSteps;

__global__ __launch_bounds__(kPackBlock) void k_pack_result(PackArgs a) {
  __shared__ uint64_t s_off[RWH_SEC_COUNT];  // section offsets (present sections only)
  __shared__ uint64_t s_len[RWH_SEC_COUNT];  // live bytes
  __shared__ const unsigned char* s_src[RWH_SEC_COUNT];
  __shared__ uint32_t s_present;
  __shared__ uint64_t s_total;
  __shared__ uint4 s_head[RWH_RESULT_HEADER_BYTES / 16];

  if (threadIdx.x == 0) {
    // counts of the build (clamped, so no section reads past its array)
    const uint32_t raw_unique = a.counts[0];
    const uint32_t n_unique = raw_unique < a.n_input ? raw_unique : a.n_input;
    const uint32_t n_oct = a.counts[a.stride];
    const uint32_t rows = n_oct < a.oct_capacity ? n_oct : a.oct_capacity;
    const ResultLayout L = result_layout(a.n_input, n_unique, rows, a.sections);
    uint32_t present = 0;
    for (int s = 0; s < RWH_SEC_COUNT; ++s) {
      if (!((a.sections >> s) & 1u)) continue;
      s_off[present] = L.offset[s];
      s_len[present] = static_cast<uint64_t>(L.count[s]) * result_entry_bytes(s);
      s_src[present] = a.src[s];
      ++present;
    }
    s_present = present;
    s_total = L.total;
    if (blockIdx.x == 0) {
      rwh_result_header h{};
      h.magic = RWH_RESULT_MAGIC;
      h.version = RWH_RESULT_VERSION;
      h.header_bytes = RWH_RESULT_HEADER_BYTES;
      h.sections = a.sections;
      h.status = a.counts[2 * a.stride];
      h.flags = n_oct > a.oct_capacity ? RWH_RESULT_TRUNCATED : 0u;
      h.n_input = a.n_input;
      h.n_unique = n_unique;
      h.n_brt = n_unique > 0 ? n_unique - 1 : 0;
      h.n_octree_nodes = n_oct;
      h.oct_rows = rows;
      h.oct_capacity = a.oct_capacity;
      h.total_bytes = L.total;
      for (int s = 0; s < RWH_SEC_COUNT; ++s) {
        h.offset[s] = L.offset[s];
        h.count[s] = L.count[s];
        h.entry_bytes[s] = ((a.sections >> s) & 1u) ? static_cast<uint16_t>(result_entry_bytes(s)) : 0;
      }
      __builtin_memcpy(s_head, &h, sizeof h);
    }
  }
  __syncthreads();

  const uint64_t units = s_total / 16;
  const uint32_t present = s_present;
  const uint64_t first = static_cast<uint64_t>(blockIdx.x) * kPackUnitsPerBlock + threadIdx.x;
  uint4* out = reinterpret_cast<uint4*>(a.out);
  for (int k = 0; k < kPackSteps; ++k) {
    const uint64_t u = first + static_cast<uint64_t>(k) * kPackBlock;
    if (u >= units) return;
    const uint64_t b = u * 16;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (b < RWH_RESULT_HEADER_BYTES) {
      v = s_head[u];
    } else {
      // the last present section starting at or before b (the first starts
      // right after the header; an empty section shares the next one's offset)
      uint32_t i = 0;
      for (uint32_t j = 1; j < present; ++j)
        if (s_off[j] <= b) i = j;
      const uint64_t off = b - s_off[i], len = s_len[i];
      const unsigned char* src = s_src[i] + off;
      if (off + 16 <= len) {
        v = *reinterpret_cast<const uint4*>(src);
      } else if (off < len) {
        uint32_t w[4] = {0, 0, 0, 0};
        const uint32_t tail = static_cast<uint32_t>(len - off);
        for (uint32_t t = 0; t < tail; ++t) w[t >> 2] |= static_cast<uint32_t>(src[t]) << (8 * (t & 3));
        v = make_uint4(w[0], w[1], w[2], w[3]);
      }
    }
    out[u] = v;
  }
}

}  // namespace

void launch_pack_result(const PackArgs& a, uint64_t max_bytes, hipStream_t stream) {
  const uint64_t units = max_bytes / 16;
  const uint64_t blocks = (units + kPackUnitsPerBlock - 1) / kPackUnitsPerBlock;
  hipLaunchKernelGGL(k_pack_result, dim3(static_cast<uint32_t>(blocks)), dim3(kPackBlock), 0, stream, a);
}

}  // namespace rw
