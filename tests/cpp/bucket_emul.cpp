// TEST INFRASTRUCTURE ONLY: runs the first two stages of a whole build as the
// build runs them on the bucket-sort path, with the library's own kernels and
// host launchers (redwood-comp_amd/csrc/k_sort.hip and k_bucket.hip compiled
// unmodified): stage 1 with the bucket counts (k_morton_bucket), the tile
// scan, the scatter and the in-bucket sorts, then, when the flag says the
// bucket sort ran, the digit-pass sort gated on it (its histogram sweep and
// 1024 x 16 tiles, the build's shape: every kernel returns, the scans after
// clearing the histogram copies).  The digit passes themselves are not
// emulated: their kernels run wave collectives in divergent code, which this
// emulation cannot model, so when the flag refuses the bucket sort the run
// stops there.  On the CPU under the emulation header in shim/, with
// AddressSanitizer: every buffer holds exactly what the build's workspace
// gives it, so an index past any of them or past an LDS array aborts the run.
//
//   bucket_emul <points.f32x4> <D> <out.bin>
// D: top bits per bucket (8..11; the build takes ceil(log2 n) - 13), which
// fixes the in-bucket passes: kBucketLowBits low bits, then 30 - D - that.
// out.bin: u32 status, flag, 1 when the histogram copies end all zero (flag 1
//          only); codes (u32 x n); sorted keys (u32 x n; 0x5a5a5a5a where
//          nothing was written); the scattered keys (u32 x n); bucket starts
//          (u32 x 2^D + 1)
#include "k_sort.hip"
#include "k_bucket.hip"

#include <cstdio>
#include <cstdlib>
#include <vector>

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s points.f32x4 D out.bin\n", argv[0]);
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  std::vector<float4> pts;
  float4 q;
  while (std::fread(&q, sizeof(q), 1, f) == 1) pts.push_back(q);
  std::fclose(f);
  const uint32_t n = static_cast<uint32_t>(pts.size());
  const uint32_t d = static_cast<uint32_t>(std::atoi(argv[2]));
  if (n == 0 || d < 8 || d > 11) return 2;
  rw::BucketPlan p{};
  p.ok = true;
  p.shift = 30 - d;
  p.n_bins = 1u << d;
  p.n_tiles = (n + rw::kBucketTile - 1) / rw::kBucketTile;
  p.hi_bits = p.shift - rw::kBucketLowBits;
  constexpr int kVariant = 5;  // 1024 x 16
  const rw::DigitPlan plan = rw::make_plan(30, 0);
  const uint32_t sort_tiles = (n + rw::kSortMinTile - 1) / rw::kSortMinTile;
  const size_t hist_words = size_t{rw::kHistCopies} * rw::kMaxPasses * 256;

  // exact-size heap buffers (ASan guards both ends)
  auto* in = new float4[n];
  std::copy(pts.begin(), pts.end(), in);
  auto* codes = new uint32_t[n];
  auto* tmp = new uint32_t[n];
  auto* out = new uint32_t[n];
  std::fill_n(codes, n, 0x5a5a5a5au);
  std::fill_n(tmp, n, 0x5a5a5a5au);
  std::fill_n(out, n, 0x5a5a5a5au);
  auto* table = new uint32_t[size_t{p.n_bins} * p.n_tiles];
  std::fill_n(table, size_t{p.n_bins} * p.n_tiles, 0x5a5a5a5au);
  auto* start = new uint32_t[rw::kBucketMaxBins + 1];
  std::fill_n(start, rw::kBucketMaxBins + 1, 0x5a5a5a5au);
  auto* count = new uint32_t[p.n_bins];
  std::fill_n(count, p.n_bins, 0x5a5a5a5au);
  auto* ctl = new uint32_t[rw::kBucketCtlWords]();
  auto* status = new uint32_t[1]{7};
  // the digit passes' workspace areas (zero, as a fresh workspace)
  auto* hist = new uint32_t[hist_words]();
  auto* digit_start = new uint32_t[rw::kMaxPasses * 256]();
  auto* lookback = new uint64_t[size_t{sort_tiles} * 256]();
  auto* tile_counts = new uint32_t[size_t{sort_tiles} * 256]();
  auto* ctrl = new uint32_t[rw::kCtrlWords]();

  rw::SegArgs seg{};
  seg.n_segs = 1;
  seg.single_count = n;
  rw::launch_morton_bucket(in, codes, n, 0.0f, 1024.0f, p, table, status, nullptr);
  const rw::BucketLaunch b{codes, out, tmp, n, p, table, count, start, ctl};
  rw::launch_bucket_sort(b, nullptr);
  rw::SortLaunch L{};
  L.keys_in = codes;
  L.keys_out = out;
  L.keys_tmp = tmp;
  L.seg = seg;
  L.n_segs = 1;
  L.variant = kVariant;
  L.n_index = n;
  L.hist_unit_keys = rw::kPlainHistUnit;
  L.n_hist_units = (n + rw::kPlainHistUnit - 1) / rw::kPlainHistUnit;
  L.plan = plan;
  L.hist = hist;
  L.digit_start = digit_start;
  L.lookback = lookback;
  L.ctrl = ctrl;
  L.hist_ready = false;  // the sort's own histogram sweep, gated too
  L.tile_counts = tile_counts;
  L.skip = ctl;
  uint32_t clear = 0;
  if (ctl[0] != 0) {
    rw::launch_sort(L, nullptr);
    clear = 1;
    for (size_t i = 0; i < hist_words; ++i)
      if (hist[i] != 0) clear = 0;
  }

  FILE* o = std::fopen(argv[3], "wb");
  if (!o) return 2;
  std::fwrite(status, 4, 1, o);
  std::fwrite(ctl, 4, 1, o);
  std::fwrite(&clear, 4, 1, o);
  std::fwrite(codes, 4, n, o);
  std::fwrite(out, 4, n, o);
  std::fwrite(tmp, 4, n, o);
  std::fwrite(start, 4, p.n_bins + 1, o);
  std::fclose(o);
  delete[] in;
  delete[] codes;
  delete[] tmp;
  delete[] out;
  delete[] table;
  delete[] start;
  delete[] count;
  delete[] ctl;
  delete[] status;
  delete[] hist;
  delete[] digit_start;
  delete[] lookback;
  delete[] tile_counts;
  delete[] ctrl;
  return 0;
}
