// Host BVH builder interface (see bvh_build.cpp).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "dev_scene.h"

namespace phx {

struct BvhOptions {
  uint32_t max_leaf = 4;          // triangles per leaf (the reference uses <= 4, bvh.cpp:42)
  int max_binary_depth = 40;      // bound on the binary tree depth before collapsing
  // SAH cost of a binary split relative to one triangle test.  On gfx950 an
  // 8-wide node visit costs several triangle tests, so the split must save
  // more than one test to pay for itself.
  float traversal_cost = 1.0f;
  // Collapse to 8-wide nodes: 1 = cost-optimal (dynamic program over the
  // binary tree), 0 = greedy (open the largest-area inner child until 8).
  int collapse = 1;
  // Expected cost of a wide-node visit relative to one triangle test in the
  // cost-optimal collapse (both are one traversal iteration).
  float node_cost = 1.0f;
};

struct Bvh {
  std::vector<BvhNode> nodes;     // nodes[0] is the root
  std::vector<TriRec> tris;       // leaf order
  uint32_t depth = 0;             // levels of 8-wide nodes
  uint32_t leaves = 0;
};

int build_bvh(const float *positions, uint32_t vertex_count, const uint32_t *indices, uint32_t triangle_count,
              const BvhOptions &opt, Bvh &out, std::string &err);

}  // namespace phx
