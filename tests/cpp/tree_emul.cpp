// TEST INFRASTRUCTURE ONLY: runs the library's radix-tree kernels
// (k_tree_local, k_tree_cross, compiled from redwood-comp_amd/csrc/k_stages.hip
// unmodified) on the CPU under the emulation header in shim/, with
// AddressSanitizer, over sorted keys read from a file.  Every buffer is
// allocated with exactly the size the C ABI documents (n_max = n_unique
// entries), so an index past any buffer or __shared__ array aborts the run.
//
//   tree_emul <keys.u32> <out.bin>
// out.bin: status word, then n_brt entries of prefix_n, has_leaf_left,
// has_leaf_right (u8), left_child, parents, edge_count (i32), and the tile
// sums (u32, one per tree tile).
#include "k_stages.hip"

#include <algorithm>
#include <cstdio>
#include <vector>

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s keys.u32 out.bin\n", argv[0]);
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  std::vector<uint32_t> keys;
  uint32_t k;
  while (std::fread(&k, 4, 1, f) == 1) keys.push_back(k);
  std::fclose(f);
  const uint32_t nu = static_cast<uint32_t>(keys.size());
  if (nu < 2) return 2;
  const uint32_t nb = nu - 1;
  // exact-size heap buffers (ASan guards both ends)
  auto* codes = new uint32_t[nu];
  std::copy(keys.begin(), keys.end(), codes);
  // outputs poisoned (bytes 0x5a): an entry no kernel writes shows as a mismatch
  auto* prefix_n = new uint8_t[nu];
  auto* hll = new uint8_t[nu];
  auto* hlr = new uint8_t[nu];
  auto* lc = new int32_t[nu];
  auto* par = new int32_t[nu];
  auto* ec = new int32_t[nu];
  std::fill_n(prefix_n, nu, uint8_t{0x5a});
  std::fill_n(hll, nu, uint8_t{0x5a});
  std::fill_n(hlr, nu, uint8_t{0x5a});
  std::fill_n(lc, nu, 0x5a5a5a5a);
  std::fill_n(par, nu, 0x5a5a5a5a);
  std::fill_n(ec, nu, 0x5a5a5a5a);
  const uint32_t tiles = rw::tree_tiles(nu);
  auto* list = new uint32_t[rw::tree_list_words(nu)]();
  auto* tile_sum = new uint32_t[tiles]();
  auto* status = new uint32_t[1]();
  auto* d_nu = new uint32_t[1]{nu};

  emu::set_domain(rw::k_tree_cross, rw::kCrossGroup);
  rw::TreeArgs a{codes, prefix_n, hll, hlr, lc, par, ec, tile_sum, status};
  rw::SegArgs seg{};
  seg.n_segs = 1;
  seg.count = d_nu;
  rw::launch_radix_tree(a, seg, nu, list, nullptr);

  FILE* o = std::fopen(argv[2], "wb");
  if (!o) return 2;
  std::fwrite(status, 4, 1, o);
  std::fwrite(prefix_n, 1, nb, o);
  std::fwrite(hll, 1, nb, o);
  std::fwrite(hlr, 1, nb, o);
  std::fwrite(lc, 4, nb, o);
  std::fwrite(par, 4, nb, o);
  std::fwrite(ec, 4, nb, o);
  std::fwrite(tile_sum, 4, tiles, o);
  std::fclose(o);
  delete[] codes;
  delete[] prefix_n;
  delete[] hll;
  delete[] hlr;
  delete[] lc;
  delete[] par;
  delete[] ec;
  delete[] list;
  delete[] tile_sum;
  delete[] status;
  delete[] d_nu;
  return 0;
}
