// Host BVH builder: binned SAH binary tree, collapsed to an 8-wide BVH of
// 128-byte nodes with 16-bit quantised child boxes and triangles in leaf
// order.
//
// Replaces the reference's Bvh::build_cpu (bvh.cpp:20-125: median split on the
// widest centroid axis, leaves of <= 4 triangles, 44-byte binary nodes).  The
// tree only changes which boxes are tested, never a hit result: traversal
// tests every triangle whose leaf box the ray reaches, with the same
// Moller-Trumbore arithmetic.
#include <algorithm>
#include <functional>
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

// SAH bins per axis: on the 737K-triangle frame 16 bins gave 44.0 ms, 32
// bins 43.2 ms (1.6 % fewer nodes, +0.08 s of build), 64 bins 43.3 ms.
constexpr int kBins = 32;
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
  float traversal_cost;

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
        traversal_cost + kTriangleCost * (parent_area > 0.f ? best_cost / parent_area : float(count));
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

// Quantises the children's boxes into node n's frame: per axis the origin is
// the node box's low corner and the step a power of two 2^e with
// org + 255 * 2^e >= the node's high bound.  Each bound is rounded outward
// until the exact value org + q * 2^e contains the child's bound.
void quantize(BvhNode &n, const Box *kids, int nk) {
  Box nb;
  for (int k = 0; k < nk; ++k) nb.grow(kids[k]);
  for (int a = 0; a < 3; ++a) {
    const float org = nb.lo[a];
    const float ext = nb.hi[a] - org;
    int e = -100;
    if (ext > 0.f) e = std::max(-100, int(std::ceil(std::log2(double(ext) / 255.0))));
    // The exact value of a quantised plane, org + q * 2^e (exact in double).
    // Planes are rounded outward against it, which makes the float decode
    // (its rounding) and the traversal's fused slab test (which evaluates it
    // unrounded) conservative as well.
    auto dec = [&](uint32_t q, int ee) { return double(org) + double(q) * std::ldexp(1.0, ee); };
    while (dec(255u, e) < double(nb.hi[a])) ++e;
    n.org[a] = org;
    n.ex[a] = int8_t(e);
    const float s = std::ldexp(1.f, e);
    for (int k = 0; k < kNodeWidth; ++k) {
      if (k >= nk) {
        n.qlo[a][k] = 1;  // empty slots: lo > hi, never reported (k >= nchild)
        n.qhi[a][k] = 0;
        continue;
      }
      double ql = std::floor(double(kids[k].lo[a] - org) / double(s));
      double qh = std::ceil(double(kids[k].hi[a] - org) / double(s));
      uint32_t lo = uint32_t(std::min(std::max(ql, 0.0), 255.0));
      uint32_t hi = uint32_t(std::min(std::max(qh, 0.0), 255.0));
      while (lo > 0 && dec(lo, e) > double(kids[k].lo[a])) --lo;
      while (hi < 255u && dec(hi, e) < double(kids[k].hi[a])) ++hi;
      if (hi < lo) hi = lo;
      n.qlo[a][k] = uint8_t(lo);
      n.qhi[a][k] = uint8_t(hi);
    }
  }
}

}  // namespace

int build_bvh(const float *positions, uint32_t vertex_count, const uint32_t *indices, uint32_t triangle_count,
              const BvhOptions &opt, Bvh &out, std::string &err) {
  out = Bvh{};
  if (triangle_count == 0) {
    err = "build_accel: the mesh has no triangles";
    return -1;
  }
  if (opt.max_leaf < 1 || opt.max_leaf > 4) {
    err = "build_accel: leaf size must be 1..4";
    return -1;
  }
  Builder b;
  b.pos = positions;
  b.idx = indices;
  b.max_leaf = opt.max_leaf;
  b.traversal_cost = opt.traversal_cost;
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

  // Collapse.  Cost-optimal (opt.collapse = 1), the wide-BVH dynamic program
  // of Ylitie, Karras and Laine (HPG 2017): F[i][j] is the least expected cost
  // of representing binary subtree i by exactly j slots of its parent (j = 1:
  // one leaf, or one new wide node), each node visit and triangle test
  // weighted by the surface area of what is entered; a wide node made from i
  // takes the slot count m of least cost G[i][m].  Greedy (opt.collapse = 0):
  // each wide node takes its binary node's children and keeps opening the
  // inner child of largest surface area until it has 8.  Either way inner
  // children are numbered consecutively and the leaf children's triangles are
  // appended in slot order, so a node's leaves are one run of records.
  constexpr int W = kNodeWidth;
  const size_t nbin = b.nodes.size();
  std::vector<float> F, G;
  std::vector<uint8_t> split_of, node_slots;
  if (opt.collapse) {
    F.assign(nbin * (W + 1), INFINITY);
    G.assign(nbin * (W + 1), INFINITY);
    split_of.assign(nbin * (W + 1), 0);
    node_slots.assign(nbin, 0);
    for (size_t i = nbin; i-- > 0;) {  // children follow their parent: reverse order is post-order
      const BinNode &n = b.nodes[i];
      float *Fi = &F[i * (W + 1)], *Gi = &G[i * (W + 1)];
      if (n.left < 0) {
        Fi[1] = n.box.area() * kTriangleCost * float(n.count);
        continue;
      }
      const float *Fl = &F[size_t(n.left) * (W + 1)], *Fr = &F[size_t(n.right) * (W + 1)];
      for (int m = 2; m <= W; ++m)
        for (int j1 = 1; j1 < m; ++j1) {
          const float c = Fl[j1] + Fr[m - j1];
          if (c < Gi[m]) {
            Gi[m] = c;
            split_of[i * (W + 1) + m] = uint8_t(j1);
          }
        }
      int bm = 2;
      for (int m = 3; m <= W; ++m)
        if (Gi[m] < Gi[bm]) bm = m;
      node_slots[i] = uint8_t(bm);
      Fi[1] = n.box.area() * opt.node_cost + Gi[bm];
      for (int m = 2; m <= W; ++m) Fi[m] = Gi[m];
    }
  }
  // the binary subtrees that fill the j slots of subtree i, in order
  std::function<void(int, int, int *, int &)> slots_of = [&](int i, int j, int *kids, int &nk) {
    if (j == 1) {
      kids[nk++] = i;
      return;
    }
    const int j1 = split_of[size_t(i) * (W + 1) + j];
    slots_of(b.nodes[i].left, j1, kids, nk);
    slots_of(b.nodes[i].right, j - j1, kids, nk);
  };
  auto tri_rec = [&](uint32_t p) {
    const float *v0 = positions + size_t(indices[size_t(p) * 3]) * 3;
    const float *v1 = positions + size_t(indices[size_t(p) * 3 + 1]) * 3;
    const float *v2 = positions + size_t(indices[size_t(p) * 3 + 2]) * 3;
    TriRec r;
    r.v0 = make_float4(v0[0], v0[1], v0[2], 0.f);
    std::memcpy(&r.v0.w, &p, 4);
    r.e1 = make_float4(v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2], 0.f);
    r.e2 = make_float4(v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2], 0.f);
    return r;
  };
  struct Job {
    int bin;
    uint32_t out;
    uint32_t depth;
  };
  out.tris.reserve(triangle_count);
  out.nodes.reserve(b.nodes.size() / 6 + 4);
  out.nodes.push_back(BvhNode{});
  std::vector<Job> jobs{{0, 0, 1}};
  uint32_t leaves = 0, depth = 1;
  while (!jobs.empty()) {
    const Job j = jobs.back();
    jobs.pop_back();
    depth = std::max(depth, j.depth);
    int kids[kNodeWidth];
    int nk = 0;
    if (b.nodes[j.bin].left < 0) {
      kids[nk++] = j.bin;  // the root is a leaf
    } else if (opt.collapse) {
      slots_of(j.bin, node_slots[size_t(j.bin)], kids, nk);
    } else {
      kids[nk++] = b.nodes[j.bin].left;
      kids[nk++] = b.nodes[j.bin].right;
    }
    while (!opt.collapse && nk < kNodeWidth) {
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
    // inner children first, then leaves
    std::stable_partition(kids, kids + nk, [&](int k) { return b.nodes[k].left >= 0; });
    BvhNode n{};
    Box boxes[kNodeWidth];
    int ninner = 0;
    for (int k = 0; k < nk; ++k) {
      boxes[k] = b.nodes[kids[k]].box;
      if (b.nodes[kids[k]].left >= 0) ++ninner;
    }
    n.nchild = uint8_t(nk);
    n.ninner = uint8_t(ninner);
    n.child_base = uint32_t(out.nodes.size());
    n.tri_base = uint32_t(out.tris.size());
    for (int k = 0; k < ninner; ++k) {
      out.nodes.push_back(BvhNode{});
      jobs.push_back({kids[k], n.child_base + uint32_t(k), j.depth + 1});
    }
    for (int k = ninner; k < nk; ++k) {
      const BinNode &c = b.nodes[kids[k]];
      n.leaf_cnt = uint16_t(n.leaf_cnt | ((c.count - 1u) << (2 * k)));
      for (uint32_t i = c.first; i < c.first + c.count; ++i) out.tris.push_back(tri_rec(b.order[i]));
      ++leaves;
    }
    quantize(n, boxes, nk);
    out.nodes[j.out] = n;
  }
  out.depth = depth;
  out.leaves = leaves;
  return 0;
}

}  // namespace phx
