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
// key: "<bucket>" for single clouds, "chunk/<bucket>" for batches
std::map<std::string, SortConfig> g_best;
bool g_loaded = false;

std::string key_of(uint32_t n, bool chunk) {
  return (chunk ? "chunk/" : "") + std::to_string(bucket_of(n));
}

bool valid(SortConfig c) {
  return c.variant >= 0 && c.variant < rwh_sort_variant_count() && c.plan >= 0 && c.plan < rwh_sort_plan_count();
}

// format: {"sort": {"<key>": [<variant>, <plan>], ...}}; an entry written as a
// bare <variant> (older files) reads as plan 0
void load_cache() {
  if (g_loaded) return;
  g_loaded = true;
  std::ifstream f(cache_path());
  if (!f) return;
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
    if (q1 == std::string::npos || colon == std::string::npos || colon > end) break;
    const std::string key = t.substr(q0 + 1, q1 - q0 - 1);
    SortConfig c;
    size_t v = t.find_first_not_of(" \t\n", colon + 1);
    if (v != std::string::npos && t[v] == '[') {
      char* e = nullptr;
      c.variant = static_cast<int>(std::strtol(t.c_str() + v + 1, &e, 10));
      const size_t comma = t.find(',', v);
      c.plan = comma != std::string::npos && comma < end ? static_cast<int>(std::strtol(t.c_str() + comma + 1, nullptr, 10)) : 0;
      p = t.find(']', v);
      if (p != std::string::npos) p = t.find(',', p);
    } else {
      c.variant = static_cast<int>(std::strtol(t.c_str() + colon + 1, nullptr, 10));
      p = t.find(',', colon);
    }
    if (valid(c)) g_best[key] = c;
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
  for (const auto& [k, c] : g_best) {
    f << (first ? "" : ", ") << '"' << k << "\": [" << c.variant << ", " << c.plan << "]";
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

std::vector<SortTiming> measure_sort_configs(uint32_t n, int reps, bool chunk) {
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
      uint32_t block = 0, kpt = 0;
      check(rwh_sort_variant_shape(v, &block, &kpt), "rwh_sort_variant_shape");
      if (chunk && block * kpt > 8192) continue;
      check(rwh_workspace_set_sort_variant(ws, v), "set variant");
      for (int p = 0; p < rwh_sort_plan_count(); ++p) {
        check(rwh_workspace_set_sort_plan(ws, p), "set plan");
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
        out.push_back({{v, p}, static_cast<double>(ms) / reps});
      }
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

bool lookup(uint32_t n, bool chunk, SortConfig* out) {
  std::lock_guard<std::mutex> g(g_mu);
  load_cache();
  const auto it = g_best.find(key_of(n, chunk));
  if (it == g_best.end()) return false;
  *out = it->second;
  return true;
}

void store(uint32_t n, bool chunk, SortConfig c) {
  std::lock_guard<std::mutex> g(g_mu);
  load_cache();
  if (!valid(c)) return;
  g_best[key_of(n, chunk)] = c;
  save_cache();
}

SortConfig best_sort_config(uint32_t n, bool chunk) {
  SortConfig c;
  if (lookup(n, chunk, &c)) return c;
  const uint32_t b = bucket_of(n);
  const uint32_t size = b >= 30 ? (1u << 30) : (1u << b);
  const auto t = measure_sort_configs(size, 3, chunk);
  if (t.empty()) return SortConfig{chunk ? 3 : 0, 0};
  SortTiming best = t[0];
  for (const auto& x : t)
    if (x.ms < best.ms) best = x;
  store(n, chunk, best.config);
  return best.config;
}

int best_sort_variant(uint32_t n) { return best_sort_config(n).variant; }

}  // namespace redwood::tune
