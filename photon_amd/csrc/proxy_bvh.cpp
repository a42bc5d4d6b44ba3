// Median-split binary BVH of the reference (bvh.cpp:20-125), the acceleration
// structure of the baseline proxy.  Same decisions: node bounds over the
// primitive boxes, split axis = widest extent of the centroid box, split at
// the middle element by std::nth_element, leaves of at most 4 primitives.
#include <algorithm>
#include <cmath>

#include "proxy_bvh.h"

namespace phx {

namespace {

struct Ref {
  uint32_t id;
  float lo[3], hi[3], c[3];
};

struct Builder {
  std::vector<Ref> refs;
  ProxyBvh *out;

  uint32_t node(uint32_t begin, uint32_t end) {
    const uint32_t id = uint32_t(out->nodes.size());
    out->nodes.push_back({});
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    float clo[3] = {1e30f, 1e30f, 1e30f}, chi[3] = {-1e30f, -1e30f, -1e30f};
    for (uint32_t i = begin; i < end; ++i)
      for (int a = 0; a < 3; ++a) {
        lo[a] = lo[a] < refs[i].lo[a] ? lo[a] : refs[i].lo[a];
        hi[a] = hi[a] > refs[i].hi[a] ? hi[a] : refs[i].hi[a];
        clo[a] = clo[a] < refs[i].c[a] ? clo[a] : refs[i].c[a];
        chi[a] = chi[a] > refs[i].c[a] ? chi[a] : refs[i].c[a];
      }
    ProxyNode n{};
    for (int a = 0; a < 3; ++a) {
      n.lo[a] = lo[a];
      n.hi[a] = hi[a];
    }
    n.count = end - begin;
    if (n.count <= 4) {
      n.is_leaf = 1;
      n.begin = uint32_t(out->prim_ids.size());
      for (uint32_t i = begin; i < end; ++i) out->prim_ids.push_back(refs[i].id);
      out->nodes[id] = n;
      return id;
    }
    n.begin = begin;
    const float ext[3] = {chi[0] - clo[0], chi[1] - clo[1], chi[2] - clo[2]};
    int axis = ext[1] > ext[0] ? 1 : 0;
    if (ext[2] > ext[axis]) axis = 2;
    const uint32_t mid = begin + (end - begin) / 2;
    std::nth_element(refs.begin() + begin, refs.begin() + mid, refs.begin() + end,
                     [axis](const Ref &a, const Ref &b) { return a.c[axis] < b.c[axis]; });
    const uint32_t l = node(begin, mid);
    const uint32_t r = node(mid, end);
    n.left = l;
    n.right = r;
    n.is_leaf = 0;
    out->nodes[id] = n;
    return id;
  }
};

}  // namespace

void build_proxy_bvh(const float *positions, const uint32_t *indices, uint32_t nt, ProxyBvh &out) {
  out = ProxyBvh{};
  Builder b;
  b.out = &out;
  b.refs.resize(nt);
  for (uint32_t t = 0; t < nt; ++t) {
    Ref &r = b.refs[t];
    r.id = t;
    const float *v[3];
    for (int k = 0; k < 3; ++k) v[k] = positions + size_t(indices[size_t(t) * 3 + k]) * 3;
    for (int a = 0; a < 3; ++a) {
      r.lo[a] = std::min({v[0][a], v[1][a], v[2][a]});
      r.hi[a] = std::max({v[0][a], v[1][a], v[2][a]});
      r.c[a] = (v[0][a] + v[1][a] + v[2][a]) * (1.f / 3.f);
    }
  }
  out.nodes.reserve(size_t(nt) * 2);
  out.prim_ids.reserve(nt);
  out.root = b.node(0, nt);
}

}  // namespace phx
