// TEST INFRASTRUCTURE ONLY: runs the library's octree kernels (k_oct_nodes
// and, on request, k_link_leaf, compiled from redwood-comp_amd/csrc/
// k_octree.hip unmodified) on the CPU under the emulation header in shim/,
// with AddressSanitizer.  Node arrays hold exactly n_brt entries, the codes
// n_unique, the octree arrays exactly the capacity: an index past any of them
// or past an LDS array aborts the run.
//
//   octree_emul <in.bin> <out.bin>
// in.bin:  u32 n_unique, n_octree_nodes, capacity, mode (bit 0: run leaf
//          linking, bit 1: the kernel without checks, bit 2: queue the
//          walks that leave a tile's window for k_oct_far, as a whole build
//          does); f32 min, range;
//          codes (u32 x n_unique); prefix_n, has_leaf_left, has_leaf_right
//          (u8 x n_brt); left_child, parents, edge_count, edge_offset
//          (i32 x n_brt)
// out.bin: u32 status; children (i32 x 8 x capacity), corner (f32 x 4 x
//          capacity), cell_size (f32), child_node_mask, child_leaf_mask
//          (i32 x capacity); u32 walks queued for k_oct_far (all tiles)
#include "k_octree.hip"

#include <cstdio>
#include <vector>

namespace {
template <class T>
T* read_array(FILE* f, size_t n) {
  T* p = new T[n == 0 ? 1 : n]();
  if (n && std::fread(p, sizeof(T), n, f) != n) {
    std::fprintf(stderr, "short input\n");
    std::exit(2);
  }
  return p;
}
}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  uint32_t hdr[4];
  float box[2];
  if (std::fread(hdr, 4, 4, f) != 4 || std::fread(box, 4, 2, f) != 2) return 2;
  const uint32_t nu = hdr[0], n_oct = hdr[1], cap = hdr[2], leaf = hdr[3] & 1u;
  const bool check = (hdr[3] & 2u) == 0;
  const bool queue = (hdr[3] & 4u) != 0;
  const uint32_t nb = nu - 1;
  auto* codes = read_array<uint32_t>(f, nu);
  auto* prefix_n = read_array<uint8_t>(f, nb);
  auto* hll = read_array<uint8_t>(f, nb);
  auto* hlr = read_array<uint8_t>(f, nb);
  auto* lc = read_array<int32_t>(f, nb);
  auto* par = read_array<int32_t>(f, nb);
  auto* ec = read_array<int32_t>(f, nb);
  auto* eo = read_array<int32_t>(f, nb);
  std::fclose(f);
  auto* children = new int32_t[8ull * cap]();
  auto* corner = new float4[cap]();
  auto* cell = new float[cap]();
  auto* mask = new int32_t[cap]();
  auto* leaf_mask = new int32_t[cap]();
  auto* status = new uint32_t[1]();
  auto* d_nu = new uint32_t[1]{nu};
  auto* d_no = new uint32_t[1]{n_oct};

  rw::OctArgs a{};
  a.edge_offset = eo;
  a.edge_count = ec;
  a.codes = codes;
  a.prefix_n = prefix_n;
  a.parent = par;
  a.has_leaf_left = hll;
  a.has_leaf_right = hlr;
  a.left_child = lc;
  a.n_oct = d_no;
  a.single_cap = cap;
  a.min_coord = box[0];
  a.range = box[1];
  a.children = children;
  a.corner = corner;
  a.cell_size = cell;
  a.child_node_mask = mask;
  a.child_leaf_mask = leaf_mask;
  a.status = status;
  const uint32_t n_tiles = rw::oct_tiles(nu);
  const size_t n_items = static_cast<size_t>(n_tiles) * rw::kOctFarCap;
  auto* far_item = new uint2[n_items == 0 ? 1 : n_items]();
  if (queue) a.far_item = far_item;
  rw::SegArgs seg{};
  seg.n_segs = 1;
  seg.count = d_nu;
  rw::launch_octree(a, seg, nu, check, nullptr);
  if (leaf) rw::launch_link_leaf(a, seg, nu, nullptr);

  FILE* o = std::fopen(argv[2], "wb");
  if (!o) return 2;
  std::fwrite(status, 4, 1, o);
  std::fwrite(children, 4, 8ull * cap, o);
  std::fwrite(corner, 16, cap, o);
  std::fwrite(cell, 4, cap, o);
  std::fwrite(mask, 4, cap, o);
  std::fwrite(leaf_mask, 4, cap, o);
  uint32_t queued = 0;
  for (size_t i = 0; queue && i < n_items; ++i) queued += far_item[i].x != 0xffffffffu;
  std::fwrite(&queued, 4, 1, o);
  std::fclose(o);
  delete[] codes;
  delete[] prefix_n;
  delete[] hll;
  delete[] hlr;
  delete[] lc;
  delete[] par;
  delete[] ec;
  delete[] eo;
  delete[] children;
  delete[] corner;
  delete[] cell;
  delete[] mask;
  delete[] leaf_mask;
  delete[] status;
  delete[] d_nu;
  delete[] d_no;
  delete[] far_item;
  return 0;
}
