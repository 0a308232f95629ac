#include "autotune.hpp"

#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

#include "hip_engine.hpp"
#include "inputs.hpp"
#include "rwhip.h"

namespace redwood::tune {

namespace {
std::mutex g_mu;
std::map<uint32_t, int> g_best;  // bucket -> variant
bool g_loaded = false;

void load_cache() {
  if (g_loaded) return;
  g_loaded = true;
  std::ifstream f(cache_path());
  if (!f) return;
  // format: {"sort": {"<bucket>": <variant>, ...}}
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string t = ss.str();
  size_t p = t.find("\"sort\"");
  if (p == std::string::npos) return;
  p = t.find('{', p);
  const size_t end = t.find('}', p);
  while (p != std::string::npos && p < end) {
    const size_t q0 = t.find('"', p + 1);
    if (q0 == std::string::npos || q0 > end) break;
    const size_t q1 = t.find('"', q0 + 1);
    const size_t colon = t.find(':', q1);
    const uint32_t bucket = static_cast<uint32_t>(std::strtoul(t.c_str() + q0 + 1, nullptr, 10));
    const int v = static_cast<int>(std::strtol(t.c_str() + colon + 1, nullptr, 10));
    if (v >= 0 && v < rwh_sort_variant_count()) g_best[bucket] = v;
    p = t.find(',', colon);
  }
}

void save_cache() {
  const std::string path = cache_path();
  std::error_code ec;
  std::filesystem::create_directories(std::filesystem::path(path).parent_path(), ec);
  std::ofstream f(path + ".tmp");
  if (!f) return;
  f << "{\"sort\": {";
  bool first = true;
  for (const auto& [b, v] : g_best) {
    f << (first ? "" : ", ") << '"' << b << "\": " << v;
    first = false;
  }
  f << "}}\n";
  f.close();
  std::filesystem::rename(path + ".tmp", path, ec);
}
}  // namespace

uint32_t bucket_of(uint32_t n) {
  uint32_t b = 0;
  while ((1ull << b) < n) ++b;
  return b < 12 ? 12 : b;
}

std::string cache_path() {
  if (const char* e = std::getenv("RWH_TUNE_CACHE")) return e;
  const char* home = std::getenv("HOME");
  return std::string(home ? home : "/tmp") + "/.cache/redwood_hip/tune.json";
}

std::vector<SortTiming> measure_sort_variants(uint32_t n, int reps) {
  using redwood::hip::check;
  std::vector<SortTiming> out;
  std::vector<uint32_t> h(n);
  for (uint32_t i = 0; i < n; ++i) h[i] = redwood::inputs::bits32(0x7e57, 7, i) >> 2;
  void *d_in = nullptr, *d_out = nullptr;
  rwh_workspace* ws = nullptr;
  rwh_stream_t st = nullptr;
  rwh_event_t e0 = nullptr, e1 = nullptr;
  try {
    check(rwh_malloc(&d_in, 4ull * n), "rwh_malloc");
    check(rwh_malloc(&d_out, 4ull * n), "rwh_malloc");
    check(rwh_stream_create(&st), "rwh_stream_create");
    check(rwh_memcpy_h2d(d_in, h.data(), 4ull * n, st), "rwh_memcpy_h2d");
    check(rwh_workspace_create(n, 1, &ws), "rwh_workspace_create");
    check(rwh_event_create(&e0), "rwh_event_create");
    check(rwh_event_create(&e1), "rwh_event_create");
    for (int v = 0; v < rwh_sort_variant_count(); ++v) {
      check(rwh_workspace_set_sort_variant(ws, v), "set variant");
      const auto sort = [&] {
        check(rwh_sort_u32(ws, static_cast<uint32_t*>(d_in), static_cast<uint32_t*>(d_out), n, 30, st),
              "rwh_sort_u32");
      };
      sort();
      check(rwh_event_record(e0, st), "record");
      for (int r = 0; r < reps; ++r) sort();
      check(rwh_event_record(e1, st), "record");
      check(rwh_event_sync(e1), "sync");
      float ms = 0;
      check(rwh_event_elapsed_ms(&ms, e0, e1), "elapsed");
      out.push_back({v, static_cast<double>(ms) / reps});
    }
  } catch (...) {
    out.clear();
  }
  if (e0) (void)rwh_event_destroy(e0);
  if (e1) (void)rwh_event_destroy(e1);
  if (ws) (void)rwh_workspace_destroy(ws);
  if (st) (void)rwh_stream_destroy(st);
  if (d_in) (void)rwh_free(d_in);
  if (d_out) (void)rwh_free(d_out);
  return out;
}

int best_sort_variant(uint32_t n) {
  const uint32_t b = bucket_of(n);
  std::lock_guard<std::mutex> g(g_mu);
  load_cache();
  if (auto it = g_best.find(b); it != g_best.end()) return it->second;
  const uint32_t size = b >= 31 ? (1u << 30) : (1u << b);
  const auto t = measure_sort_variants(size, 3);
  if (t.empty()) return 0;
  int best = t[0].variant;
  double best_ms = t[0].ms;
  for (const auto& x : t)
    if (x.ms < best_ms) {
      best = x.variant;
      best_ms = x.ms;
    }
  g_best[b] = best;
  save_cache();
  return best;
}

}  // namespace redwood::tune
