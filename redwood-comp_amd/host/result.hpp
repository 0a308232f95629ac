// A build's outputs in the packed result format of include/rwhip.h (format
// version 1): one host buffer, filled by one device-to-host copy of what
// rwh_pack_result wrote, read through views that are checked against the
// header.  The reference has no export of its results (apps/tree/main.cpp:
// 12-16 prints the counts); the sections are AppData's arrays
// (apps/tree/app_data.hpp:48-96) at their live lengths.
#pragma once

#include <cstdint>
#include <span>
#include <vector>

#include "rwhip.h"

namespace redwood::result {

class PackedResult {
 public:
  PackedResult() = default;
  explicit PackedResult(std::vector<uint8_t> bytes) : bytes_(std::move(bytes)) {}

  bool empty() const { return bytes_.empty(); }
  const std::vector<uint8_t>& bytes() const { return bytes_; }
  // RWH_OK for a well-formed result of a valid build (rwh_result_check).
  int check() const;
  // The header, after the format checks (redwood::hip::Error otherwise).
  rwh_result_header header() const;
  // Section s as elements of T: count x (entry size / sizeof(T)) of them,
  // e.g. section<int32_t>(RWH_SEC_CHILDREN) gives 8 per octree node.  Throws
  // redwood::hip::Error if the section is absent or the buffer malformed.
  template <class T>
  std::span<const T> section(int s) const {
    const void* p = nullptr;
    uint32_t count = 0;
    const uint32_t entry = section_raw(s, &p, &count);
    return {static_cast<const T*>(p), static_cast<size_t>(count) * (entry / sizeof(T))};
  }

 private:
  // the section's data and entry count; returns its entry size in bytes
  uint32_t section_raw(int s, const void** p, uint32_t* count) const;
  std::vector<uint8_t> bytes_;
};

// Packs a finished build (enqueued before on the stream) into d_scratch,
// which holds at least rwh_result_bytes(n_input, n_input, cap, cap,
// sections) bytes, and copies it to the host; waits for the stream.
// known_bytes: the result's size when the caller knows the counts
// (rwh_result_bytes of them), else 0 (the header is copied first).  Counts
// as for rwh_pack_result.
PackedResult fetch(const rwh_tree_buffers& bufs, const uint32_t* d_counts, uint32_t counts_stride,
                   uint32_t n_input, uint32_t sections, void* d_scratch, uint64_t scratch_bytes,
                   rwh_stream_t stream, uint64_t known_bytes = 0);

}  // namespace redwood::result
