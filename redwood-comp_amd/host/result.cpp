#include "result.hpp"

#include "hip_engine.hpp"

namespace redwood::result {

using redwood::hip::check;

int PackedResult::check() const { return rwh_result_check(bytes_.data(), bytes_.size()); }

rwh_result_header PackedResult::header() const {
  rwh_result_header h;
  redwood::hip::check(rwh_result_header_get(bytes_.data(), bytes_.size(), &h), "PackedResult::header");
  return h;
}

uint32_t PackedResult::section_raw(int s, const void** p, uint32_t* count) const {
  redwood::hip::check(rwh_result_section(bytes_.data(), bytes_.size(), s, p, count), "PackedResult::section");
  return header().entry_bytes[s];
}

PackedResult fetch(const rwh_tree_buffers& bufs, const uint32_t* d_counts, uint32_t counts_stride,
                   uint32_t n_input, uint32_t sections, void* d_scratch, uint64_t scratch_bytes,
                   rwh_stream_t stream, uint64_t known_bytes) {
  check(rwh_pack_result(&bufs, d_counts, counts_stride, n_input, sections, d_scratch, scratch_bytes, stream),
        "rwh_pack_result");
  uint64_t total = known_bytes;
  std::vector<uint8_t> out;
  if (total == 0) {
    rwh_result_header h;
    check(rwh_memcpy_d2h(&h, d_scratch, sizeof h, stream), "rwh_memcpy_d2h");
    check(rwh_stream_sync(stream), "rwh_stream_sync");
    total = h.magic == RWH_RESULT_MAGIC && h.total_bytes <= scratch_bytes ? h.total_bytes : sizeof h;
  }
  out.resize(total);
  check(rwh_memcpy_d2h(out.data(), d_scratch, total, stream), "rwh_memcpy_d2h");
  check(rwh_stream_sync(stream), "rwh_stream_sync");
  return PackedResult(std::move(out));
}

}  // namespace redwood::result
