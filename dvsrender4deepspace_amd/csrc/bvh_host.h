// Host-side BVH build, validation and SoA re-layout for the gfx950 kernels.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "dvs_hip.h"

namespace dvsamd {

// BVHBuildParams (include/bvh/build.h:9-13).
struct BuildOptions {
  int max_leaf_size = 8;
  float node_cost = 1.0f;
  float primitive_cost = 1.0f;
};

// Builds the reference-format node array (children before parents, root last)
// and reorders `indices` into leaf order.  Returns the root index or -1.
int build_bvh(const BuildOptions &opt, const float *vertices, int64_t vertex_count,
              uint32_t *indices, int64_t triangle_count, std::vector<dvs_bvh_node> &nodes);

// Empty string when the tree is reachable, acyclic and well formed.
std::string validate_bvh(const dvs_bvh_node *nodes, int64_t count, int32_t root);

// Device layout of one BVH.
//  pairs:      16 floats (64 B) per internal node: both child boxes then the two
//              child refs (int bits); a ref >= 0 names a pair record, a ref < 0
//              is ~leaf_index.
//  leaf_ranges: (start, count) per leaf.
//  triangles:  12 floats (48 B) per triangle record: v0, v1-v0, v2-v0, then
//              (int bits) the reference's BVH-order triangle index and the
//              [lo, hi) index range of the reference leaf holding it.
//
// device_leaf_size > 0 subdivides every reference leaf with more triangles
// into a device-only subtree (median splits, boxes inflated beyond rounding);
// the traversal then resolves equal-distance hits inside one reference leaf
// to the lowest reference index, as the reference's sequential leaf scan does.
struct SoaBvh {
  std::vector<float> pairs;
  std::vector<int32_t> leaf_ranges;
  std::vector<float> triangles;
  int32_t root_ref = 0;
  int32_t max_depth = 0;  // deepest leaf below the root
  // every reference subtree covers one triangle index range, left before
  // right (word 14 of each reference pair: the right subtree's first index)
  bool ordered_ranges = false;
  int64_t internal_nodes = 0, leaf_count = 0;
};

SoaBvh relayout(const dvs_bvh_node *nodes, int64_t count, int32_t root, const float *vertices,
                const uint32_t *indices, int64_t triangle_count, int device_leaf_size);

}  // namespace dvsamd
