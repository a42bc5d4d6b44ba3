// Host BVH builder: binned SAH binary tree, collapsed to a 4-wide BVH of
// 128-byte nodes with triangles in leaf order.
//
// Replaces the reference's Bvh::build_cpu (bvh.cpp:20-125: median split on the
// widest centroid axis, leaves of <= 4 triangles, 44-byte binary nodes).  The
// tree only changes which boxes are tested, never a hit result: traversal
// tests every triangle whose leaf box the ray reaches, with the same
// Moller-Trumbore arithmetic.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "bvh_build.h"

namespace phx {

namespace {

struct Box {
  float lo[3] = {INFINITY, INFINITY, INFINITY};
  float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  void grow(const float *p) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], p[a]);
      hi[a] = std::max(hi[a], p[a]);
    }
  }
  void grow(const Box &b) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], b.lo[a]);
      hi[a] = std::max(hi[a], b.hi[a]);
    }
  }
  float area() const {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (!(dx >= 0.f) || !(dy >= 0.f) || !(dz >= 0.f)) return 0.f;
    return 2.f * (dx * dy + dy * dz + dz * dx);
  }
};

struct BinNode {
  Box box;
  int left = -1, right = -1;  // inner: children; leaf: -1
  uint32_t first = 0, count = 0;
};

constexpr int kBins = 16;
constexpr float kTraversalCost = 1.0f;
constexpr float kTriangleCost = 1.0f;

struct Builder {
  const float *pos;
  const uint32_t *idx;
  std::vector<Box> pbox;
  std::vector<float> cent;  // 3 per prim
  std::vector<uint32_t> order;
  std::vector<BinNode> nodes;
  int max_depth;
  uint32_t max_leaf;

  int make_node(uint32_t first, uint32_t count, const Box &box) {
    BinNode n;
    n.box = box;
    n.first = first;
    n.count = count;
    nodes.push_back(n);
    return int(nodes.size() - 1);
  }

  // Splits [first, first+count) in place; returns the size of the left part,
  // 0 when the range should stay a leaf.
  uint32_t split(uint32_t first, uint32_t count, const Box &box, int depth) {
    Box cb;
    for (uint32_t i = first; i < first + count; ++i) cb.grow(&cent[size_t(order[i]) * 3]);
    int axis = 0;
    float ext[3];
    for (int a = 0; a < 3; ++a) ext[a] = cb.hi[a] - cb.lo[a];
    if (ext[1] > ext[axis]) axis = 1;
    if (ext[2] > ext[axis]) axis = 2;
    auto median_split = [&](int ax) {
      const uint32_t mid = count / 2;
      std::nth_element(order.begin() + first, order.begin() + first + mid, order.begin() + first + count,
                       [&](uint32_t a, uint32_t b) { return cent[size_t(a) * 3 + ax] < cent[size_t(b) * 3 + ax]; });
      return mid;
    };
    if (!(ext[axis] > 0.f)) {
      // All centroids coincide: split by index.
      return count > max_leaf ? count / 2 : 0;
    }
    // Depth budget: a range that could not reach leaf size within the
    // remaining levels by halving is split at the object median.
    int need = 0;
    for (uint32_t c = count; c > max_leaf; c = (c + 1) / 2) ++need;
    if (depth + need + 2 >= max_depth) return median_split(axis);

    float best_cost = INFINITY;
    int best_axis = -1, best_bin = -1;
    for (int a = 0; a < 3; ++a) {
      if (!(ext[a] > 0.f)) continue;
      Box bins[kBins];
      uint32_t cnt[kBins] = {0};
      const float scale = float(kBins) / ext[a];
      for (uint32_t i = first; i < first + count; ++i) {
        const uint32_t p = order[i];
        int b = int((cent[size_t(p) * 3 + a] - cb.lo[a]) * scale);
        b = std::min(std::max(b, 0), kBins - 1);
        bins[b].grow(pbox[p]);
        cnt[b]++;
      }
      float right_area[kBins];
      uint32_t right_cnt[kBins];
      Box acc;
      uint32_t n = 0;
      for (int b = kBins - 1; b > 0; --b) {
        acc.grow(bins[b]);
        n += cnt[b];
        right_area[b] = acc.area();
        right_cnt[b] = n;
      }
      Box lacc;
      uint32_t ln = 0;
      for (int b = 0; b < kBins - 1; ++b) {
        lacc.grow(bins[b]);
        ln += cnt[b];
        if (ln == 0 || right_cnt[b + 1] == 0) continue;
        const float cost = lacc.area() * float(ln) + right_area[b + 1] * float(right_cnt[b + 1]);
        if (cost < best_cost) {
          best_cost = cost;
          best_axis = a;
          best_bin = b;
        }
      }
    }
    const float parent_area = box.area();
    const float split_cost =
        kTraversalCost + kTriangleCost * (parent_area > 0.f ? best_cost / parent_area : float(count));
    const float leaf_cost = kTriangleCost * float(count);
    if (count <= max_leaf && (best_axis < 0 || leaf_cost <= split_cost)) return 0;
    if (best_axis < 0) return median_split(axis);
    const float scale = float(kBins) / ext[best_axis];
    auto mid = std::partition(order.begin() + first, order.begin() + first + count, [&](uint32_t p) {
      int b = int((cent[size_t(p) * 3 + best_axis] - cb.lo[best_axis]) * scale);
      b = std::min(std::max(b, 0), kBins - 1);
      return b <= best_bin;
    });
    const uint32_t left = uint32_t(mid - (order.begin() + first));
    if (left == 0 || left == count) return median_split(axis);
    return left;
  }

  void build(uint32_t nt) {
    struct Task {
      int node;
      int depth;
    };
    Box root;
    for (uint32_t i = 0; i < nt; ++i) root.grow(pbox[i]);
    nodes.reserve(size_t(nt) * 2 / std::max<uint32_t>(max_leaf / 2, 1) + 16);
    make_node(0, nt, root);
    std::vector<Task> stack{{0, 0}};
    while (!stack.empty()) {
      const Task t = stack.back();
      stack.pop_back();
      const uint32_t first = nodes[t.node].first, count = nodes[t.node].count;
      const uint32_t left = split(first, count, nodes[t.node].box, t.depth);
      if (left == 0) continue;
      Box lb, rb;
      for (uint32_t i = first; i < first + left; ++i) lb.grow(pbox[order[i]]);
      for (uint32_t i = first + left; i < first + count; ++i) rb.grow(pbox[order[i]]);
      const int l = make_node(first, left, lb);
      const int r = make_node(first + left, count - left, rb);
      nodes[t.node].left = l;
      nodes[t.node].right = r;
      stack.push_back({r, t.depth + 1});
      stack.push_back({l, t.depth + 1});
    }
  }
};

Box empty_box() {
  Box b;
  for (int a = 0; a < 3; ++a) b.lo[a] = b.hi[a] = 0.f;
  return b;
}

void set_child(Node4 &n, int slot, const Box &b, int code) {
  float *lo[3] = {&n.lox.x, &n.loy.x, &n.loz.x};
  float *hi[3] = {&n.hix.x, &n.hiy.x, &n.hiz.x};
  for (int a = 0; a < 3; ++a) {
    lo[a][slot] = b.lo[a];
    hi[a][slot] = b.hi[a];
  }
  (&n.child.x)[slot] = code;
}

}  // namespace

int build_bvh4(const float *positions, uint32_t vertex_count, const uint32_t *indices, uint32_t triangle_count,
               const Bvh4Options &opt, Bvh4 &out, std::string &err) {
  out = Bvh4{};
  if (triangle_count == 0) {
    err = "build_accel: the mesh has no triangles";
    return -1;
  }
  if (opt.max_leaf < 1 || opt.max_leaf > 8) {
    err = "build_accel: leaf size must be 1..8";
    return -1;
  }
  Builder b;
  b.pos = positions;
  b.idx = indices;
  b.max_leaf = opt.max_leaf;
  b.max_depth = opt.max_binary_depth;
  b.pbox.resize(triangle_count);
  b.cent.resize(size_t(triangle_count) * 3);
  b.order.resize(triangle_count);
  for (uint32_t t = 0; t < triangle_count; ++t) {
    Box bx;
    for (int k = 0; k < 3; ++k) {
      const uint32_t vi = indices[size_t(t) * 3 + k];
      if (vi >= vertex_count) {
        err = "build_accel: triangle index out of range";
        return -1;
      }
      bx.grow(positions + size_t(vi) * 3);
    }
    b.pbox[t] = bx;
    for (int a = 0; a < 3; ++a) b.cent[size_t(t) * 3 + a] = 0.5f * (bx.lo[a] + bx.hi[a]);
    b.order[t] = t;
  }
  b.build(triangle_count);

  // Triangle records in leaf order (the binary leaves' ranges of `order`).
  out.tris.resize(triangle_count);
  for (uint32_t i = 0; i < triangle_count; ++i) {
    const uint32_t p = b.order[i];
    const float *v0 = positions + size_t(indices[size_t(p) * 3]) * 3;
    const float *v1 = positions + size_t(indices[size_t(p) * 3 + 1]) * 3;
    const float *v2 = positions + size_t(indices[size_t(p) * 3 + 2]) * 3;
    TriRec &r = out.tris[i];
    r.v0 = make_float4(v0[0], v0[1], v0[2], 0.f);
    std::memcpy(&r.v0.w, &p, 4);
    r.e1 = make_float4(v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2], 0.f);
    r.e2 = make_float4(v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2], 0.f);
  }

  // Collapse: each 4-wide node takes its binary node's children and keeps
  // opening the inner child of largest surface area until it has 4.
  struct Job {
    int bin;
    int out;
    uint32_t depth;
  };
  auto leaf_code = [&](const BinNode &n) { return int(~((n.first << 3) | (n.count - 1))); };
  const BinNode &root = b.nodes[0];
  if (root.left < 0) {
    // A single leaf: wrap it in one node.
    Node4 n{};
    for (int s = 0; s < 4; ++s) set_child(n, s, empty_box(), kEmptyChild);
    set_child(n, 0, root.box, leaf_code(root));
    out.nodes.push_back(n);
    out.depth = 1;
    out.leaves = 1;
    return 0;
  }
  out.nodes.reserve(b.nodes.size() / 3 + 4);
  out.nodes.push_back(Node4{});
  std::vector<Job> jobs{{0, 0, 1}};
  uint32_t leaves = 0, depth = 1;
  while (!jobs.empty()) {
    const Job j = jobs.back();
    jobs.pop_back();
    depth = std::max(depth, j.depth);
    int kids[4];
    int nk = 0;
    kids[nk++] = b.nodes[j.bin].left;
    kids[nk++] = b.nodes[j.bin].right;
    while (nk < 4) {
      int best = -1;
      float best_area = -1.f;
      for (int k = 0; k < nk; ++k) {
        const BinNode &c = b.nodes[kids[k]];
        if (c.left >= 0 && c.box.area() > best_area) {
          best_area = c.box.area();
          best = k;
        }
      }
      if (best < 0) break;
      const int open = kids[best];
      kids[best] = b.nodes[open].left;
      kids[nk++] = b.nodes[open].right;
    }
    Node4 n{};
    for (int s = 0; s < 4; ++s) set_child(n, s, empty_box(), kEmptyChild);
    for (int k = 0; k < nk; ++k) {
      const BinNode &c = b.nodes[kids[k]];
      if (c.left < 0) {
        set_child(n, k, c.box, leaf_code(c));
        ++leaves;
      } else {
        const int idx = int(out.nodes.size());
        out.nodes.push_back(Node4{});
        set_child(n, k, c.box, idx);
        jobs.push_back({kids[k], idx, j.depth + 1});
      }
    }
    out.nodes[j.out] = n;
  }
  out.depth = depth;
  out.leaves = leaves;
  return 0;
}

}  // namespace phx
