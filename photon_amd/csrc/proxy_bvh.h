// The reference's acceleration structure, kept as the baseline of the
// "same GPU, reference structure" comparison (the reference's Kokkos-HIP build
// cannot be built offline).  Binary BVH by median split on the widest centroid
// axis with leaves of <= 4 triangles and 44-byte nodes (bvh.cpp:20-125), and
// its traversal: a 64-entry per-thread stack, children tested when their
// parent is popped, nearer child first (mesh_intersector.h:57-124).
#pragma once

#include <stdint.h>

#include <vector>

#include "dev_trace.h"

namespace phx {

struct ProxyNode {  // 44 bytes, the reference's BvhNode (bvh.h:11-19)
  float lo[3], hi[3];
  uint32_t left, right, begin, count, is_leaf;
};
static_assert(sizeof(ProxyNode) == 44, "ProxyNode mirrors the 44-byte BvhNode");

struct ProxyBvh {
  std::vector<ProxyNode> nodes;
  std::vector<uint32_t> prim_ids;
  uint32_t root = 0;
};

void build_proxy_bvh(const float *positions, const uint32_t *indices, uint32_t triangle_count, ProxyBvh &out);

struct ProxyScene {
  const ProxyNode *nodes;
  const uint32_t *prim_ids;
  const uint32_t *indices;
  const float *positions;
  uint32_t root;
};

PHX_DEV float px_entry(const ProxyNode &n, vec3 o, vec3 inv, float tmin, float tmax) {
  const float tx0 = (n.lo[0] - o.x) * inv.x, tx1 = (n.hi[0] - o.x) * inv.x;
  const float ty0 = (n.lo[1] - o.y) * inv.y, ty1 = (n.hi[1] - o.y) * inv.y;
  const float tz0 = (n.lo[2] - o.z) * inv.z, tz1 = (n.hi[2] - o.z) * inv.z;
  const float a = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
  const float b = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
  return a <= b ? a : -1.f;
}

// mesh_intersector.h:57-124 with the reference's gathers: prim id, then the
// three vertex indices, then the three positions.
template <bool ANY>
PHX_DEV bool px_traverse(const ProxyScene &s, vec3 o, vec3 d, float tmin, float tmax, TraceHit &h) {
  h.tri = 0xffffffffu;
  h.prim = 0xffffffffu;
  h.t = tmax;
  h.u = h.v = 0.f;
  if (!(tmax > tmin)) return false;
  const vec3 inv{safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z)};
  uint32_t stack[64];
  int sp = 0;
  stack[sp++] = s.root;
  float best = tmax;
  while (sp > 0) {
    const ProxyNode n = s.nodes[stack[--sp]];
    if (px_entry(n, o, inv, tmin, best) < 0.f) continue;
    if (n.is_leaf) {
      for (uint32_t i = 0; i < n.count; ++i) {
        const uint32_t prim = s.prim_ids[n.begin + i];
        const vec3 v0 = ld3(s.positions + size_t(s.indices[prim * 3]) * 3);
        const vec3 v1 = ld3(s.positions + size_t(s.indices[prim * 3 + 1]) * 3);
        const vec3 v2 = ld3(s.positions + size_t(s.indices[prim * 3 + 2]) * 3);
        TriRec r;
        r.v0 = make_float4(v0.x, v0.y, v0.z, 0.f);
        r.e1 = make_float4(v1.x - v0.x, v1.y - v0.y, v1.z - v0.z, 0.f);
        r.e2 = make_float4(v2.x - v0.x, v2.y - v0.y, v2.z - v0.z, 0.f);
        float t, u, v;
        if (!tri_test(r, o, d, tmin, best, t, u, v)) continue;
        best = t;
        h.t = t;
        h.u = u;
        h.v = v;
        h.tri = prim;
        h.prim = prim;
        if (ANY) return true;
      }
    } else if (sp + 2 <= 64) {
      const float dl = px_entry(s.nodes[n.left], o, inv, tmin, best);
      const float dr = px_entry(s.nodes[n.right], o, inv, tmin, best);
      if (dl >= 0.f && dr >= 0.f) {
        if (dl < dr) {
          stack[sp++] = n.right;
          stack[sp++] = n.left;
        } else {
          stack[sp++] = n.left;
          stack[sp++] = n.right;
        }
      } else if (dl >= 0.f) {
        stack[sp++] = n.left;
      } else if (dr >= 0.f) {
        stack[sp++] = n.right;
      }
    }
  }
  return h.tri != 0xffffffffu;
}

}  // namespace phx
