// The tree app's AppData: every buffer of one octree build and its counts,
// with the reference's field names, sizes and getters
// (reference: apps/tree/app_data.hpp:11-158).  Buffers are allocated through
// the resource given at construction: a HIP engine's resource puts them in
// HBM, std::pmr::new_delete_resource() in host memory.
#pragma once

#include <cstdint>
#include <limits>
#include <memory>
#include <stdexcept>

#include "redwood_base.hpp"

namespace redwood::tree {

// The octree is preallocated at 0.6 x n_input (app_data.hpp:11-13, 28).
constexpr float k_memory_ratio = 0.6f;

struct Vec4 {
  float x, y, z, w;
};

template <typename T>
using UsmVector = redwood::UsmVector<T>;

inline size_t octree_capacity(size_t n_input) {
  return static_cast<size_t>(static_cast<float>(n_input) * k_memory_ratio);
}

struct AppData final : public redwood::BaseAppData {
  static constexpr float min_coord = 0.0f;
  static constexpr float range = 1024.0f;
  static constexpr uint64_t default_seed = 114514;  // app_data.hpp:32

  // n_input points uniform in [0,1024)^3 from the counter-based generator.
  AppData(std::pmr::memory_resource* mr, size_t n_input, uint64_t seed = default_seed);
  // n_input points given as x,y,z,w host floats.
  AppData(std::pmr::memory_resource* mr, const float* xyzw, size_t n_input);
  ~AppData() override = default;

  const uint32_t n_input;
  uint32_t n_unique = std::numeric_limits<uint32_t>::max();
  uint32_t n_brt_nodes = std::numeric_limits<uint32_t>::max();
  uint32_t n_octree_nodes = std::numeric_limits<uint32_t>::max();

  UsmVector<Vec4> u_input_points;
  UsmVector<uint32_t> u_morton_keys;
  UsmVector<uint32_t> u_morton_keys_alt;
  UsmVector<int32_t> u_edge_count;
  UsmVector<int32_t> u_edge_offset;

  struct RadixTree {
    RadixTree(size_t n_nodes, std::pmr::memory_resource* mr)
        : u_prefix_n(n_nodes, mr),
          u_has_leaf_left(n_nodes, mr),
          u_has_leaf_right(n_nodes, mr),
          u_left_child(n_nodes, mr),
          u_parents(n_nodes, mr) {}
    UsmVector<uint8_t> u_prefix_n;
    UsmVector<uint8_t> u_has_leaf_left;
    UsmVector<uint8_t> u_has_leaf_right;
    UsmVector<int32_t> u_left_child;
    UsmVector<int32_t> u_parents;
  } brt;

  struct Octree {
    Octree(size_t n_nodes, std::pmr::memory_resource* mr)
        : u_children(n_nodes * 8, mr),
          u_corner(n_nodes, mr),
          u_cell_size(n_nodes, mr),
          u_child_node_mask(n_nodes, mr),
          u_child_leaf_mask(n_nodes, mr) {}
    UsmVector<int32_t> u_children;  // 8 per node
    UsmVector<Vec4> u_corner;
    UsmVector<float> u_cell_size;
    UsmVector<int32_t> u_child_node_mask;
    UsmVector<int32_t> u_child_leaf_mask;
  } oct;

  [[nodiscard]] uint32_t get_n_input() const { return n_input; }
  [[nodiscard]] uint32_t get_n_unique() const {
    if (n_unique == std::numeric_limits<uint32_t>::max()) throw std::runtime_error("n_unique is not set");
    return n_unique;
  }
  void set_n_unique(uint32_t v) { n_unique = v; }
  void set_n_brt_nodes(uint32_t v) { n_brt_nodes = v; }
  // n_unique - 1 (app_data.hpp:123-125), 0 for an empty input.
  [[nodiscard]] uint32_t get_n_brt_nodes() const {
    const uint32_t u = get_n_unique();
    return u > 0 ? u - 1 : 0;
  }
  [[nodiscard]] uint32_t get_n_octree_nodes() const {
    if (n_octree_nodes == std::numeric_limits<uint32_t>::max())
      throw std::runtime_error("n_octree_nodes is not set");
    return n_octree_nodes;
  }
  void set_n_octree_nodes(uint32_t v) { n_octree_nodes = v; }

  // After stage 2 the sorted keys are in u_morton_keys_alt; after stage 3 the
  // unique keys are in u_morton_keys (the contract of app_data.hpp:139-146).
  [[nodiscard]] uint32_t* get_sorted_morton_keys() { return u_morton_keys_alt.data(); }
  [[nodiscard]] const uint32_t* get_sorted_morton_keys() const { return u_morton_keys_alt.data(); }
  [[nodiscard]] uint32_t* get_unique_morton_keys() { return u_morton_keys.data(); }
  [[nodiscard]] const uint32_t* get_unique_morton_keys() const { return u_morton_keys.data(); }

  // Backend-private per-build state (a HIP engine keeps its workspace here).
  std::shared_ptr<void> backend_state;
};

}  // namespace redwood::tree
