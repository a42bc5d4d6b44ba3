// Software BVH traversal for gfx950 (no ray accelerators on CDNA4).
//
// One lane traces one ray through the 4-wide BVH.  The traversal stack is a
// short per-lane stack in LDS (stack[depth][lane], so a wave's 64 lanes touch
// 64 consecutive dwords at any depth: no bank conflicts) that spills to a
// per-thread region of global memory beyond kShortStack entries; the builder
// guarantees the total never exceeds kMaxStack (the reference's 64-entry
// stack, mesh_intersector.h:74).
//
// Hit semantics are those of the reference's intersect_mesh_bvh
// (mesh_intersector.h:57-124) and intersect_triangle (triangle_intersect.h:18-52):
// rays with !(tmax > tmin) are inactive; the hit interval [tmin, tmax] is
// inclusive; a triangle is rejected only when |det| < 1e-30.  The triangle
// test repeats the reference's float operations in the same order (no
// contraction), so t, u, v are bitwise those of the reference for the same
// triangle.  Equal-t ties between triangles go to the larger primitive id,
// which makes the result independent of the tree layout (the reference's own
// tie rule is "last visited", which depends on its tree).
#pragma once

#include "dev_scene.h"

namespace phx {

constexpr int kShortStack = 16;
constexpr int kMaxStack = 64;

struct TraceHit {
  float t;
  float u, v;
  uint32_t tri;   // leaf-order triangle index, 0xffffffff: none
  uint32_t prim;  // primitive id
};

PHX_DEV float safe_rcp(float x) { return 1.f / (fabsf(x) > 1e-20f ? x : (x < 0.f ? -1e-20f : 1e-20f)); }

// Per-lane stack: LDS for the first kShortStack entries, global spill beyond.
struct LaneStack {
  int *lds;        // &lds_base[lane], stride = block size
  int stride;
  int *spill;      // this thread's kMaxStack - kShortStack entries
  int sp;
  PHX_DEV void push(int v) {
    if (sp < kShortStack) lds[sp * stride] = v;
    else spill[sp - kShortStack] = v;
    ++sp;
  }
  PHX_DEV int pop() {
    --sp;
    return sp < kShortStack ? lds[sp * stride] : spill[sp - kShortStack];
  }
};

// Reference Moller-Trumbore (triangle_intersect.h:18-52) on a leaf record.
PHX_DEV bool tri_test(const TriRec &r, vec3 o, vec3 d, float tmin, float tmax, float &t, float &u, float &v) {
  const vec3 v0{r.v0.x, r.v0.y, r.v0.z};
  const vec3 e1{r.e1.x, r.e1.y, r.e1.z};
  const vec3 e2{r.e2.x, r.e2.y, r.e2.z};
  const vec3 p = cross(d, e2);
  const float det = dot(e1, p);
  if (fabsf(det) < 1e-30f) return false;
  const float inv = 1.f / det;
  const vec3 tv = o - v0;
  u = dot(tv, p) * inv;
  if (u < 0.f || u > 1.f) return false;
  const vec3 q = cross(tv, e1);
  v = dot(d, q) * inv;
  if (v < 0.f || (u + v) > 1.f) return false;
  t = dot(e2, q) * inv;
  return !(t < tmin || t > tmax);
}

// Slab test of the 4 children.  Entry distances are clamped to [tmin, tmax];
// the exit is widened by a few ulps so rounding never culls a box the exact
// test would enter.
PHX_DEV void box4(const Node4 &n, vec3 o, vec3 inv, float tmin, float tmax, float tn[4], bool hit[4]) {
  const float lox[4] = {n.lox.x, n.lox.y, n.lox.z, n.lox.w};
  const float loy[4] = {n.loy.x, n.loy.y, n.loy.z, n.loy.w};
  const float loz[4] = {n.loz.x, n.loz.y, n.loz.z, n.loz.w};
  const float hix[4] = {n.hix.x, n.hix.y, n.hix.z, n.hix.w};
  const float hiy[4] = {n.hiy.x, n.hiy.y, n.hiy.z, n.hiy.w};
  const float hiz[4] = {n.hiz.x, n.hiz.y, n.hiz.z, n.hiz.w};
  const int ch[4] = {n.child.x, n.child.y, n.child.z, n.child.w};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float tx0 = (lox[i] - o.x) * inv.x, tx1 = (hix[i] - o.x) * inv.x;
    const float ty0 = (loy[i] - o.y) * inv.y, ty1 = (hiy[i] - o.y) * inv.y;
    const float tz0 = (loz[i] - o.z) * inv.z, tz1 = (hiz[i] - o.z) * inv.z;
    const float a = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
    const float b = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
    tn[i] = a;
    hit[i] = a <= b * 1.0000004f && ch[i] != kEmptyChild;
  }
}

PHX_DEV void cswap(float &ka, int &va, float &kb, int &vb) {
  if (kb < ka) {
    const float k = ka;
    ka = kb;
    kb = k;
    const int v = va;
    va = vb;
    vb = v;
  }
}

// Closest hit (ANY = false) or any hit (ANY = true).  Returns true on a hit.
template <bool ANY>
PHX_DEV bool traverse(const DevScene &s, vec3 o, vec3 d, float tmin, float tmax, LaneStack &st, TraceHit &h) {
  h.tri = 0xffffffffu;
  h.prim = 0xffffffffu;
  h.t = tmax;
  h.u = h.v = 0.f;
  if (!(tmax > tmin)) return false;
  const vec3 inv{safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z)};
  float best = tmax;
  int node = 0;
  st.sp = 0;
  while (true) {
    if (node >= 0) {
      const Node4 n = s.nodes[node];
      float tn[4];
      bool hb[4];
      box4(n, o, inv, tmin, best, tn, hb);
      float k0 = hb[0] ? tn[0] : INFINITY, k1 = hb[1] ? tn[1] : INFINITY;
      float k2 = hb[2] ? tn[2] : INFINITY, k3 = hb[3] ? tn[3] : INFINITY;
      int c0 = n.child.x, c1 = n.child.y, c2 = n.child.z, c3 = n.child.w;
      const int cnt = int(hb[0]) + int(hb[1]) + int(hb[2]) + int(hb[3]);
      if (cnt > 0) {
        // sort the 4 (distance, child) pairs, misses last
        cswap(k0, c0, k1, c1);
        cswap(k2, c2, k3, c3);
        cswap(k0, c0, k2, c2);
        cswap(k1, c1, k3, c3);
        cswap(k1, c1, k2, c2);
        if (cnt > 3) st.push(c3);
        if (cnt > 2) st.push(c2);
        if (cnt > 1) st.push(c1);
        node = c0;
        continue;
      }
    } else {
      const uint32_t code = ~uint32_t(node);
      const uint32_t first = code >> 3, count = (code & 7u) + 1u;
      for (uint32_t k = 0; k < count; ++k) {
        const TriRec r = s.tris[first + k];
        float t, u, v;
        if (!tri_test(r, o, d, tmin, best, t, u, v)) continue;
        const uint32_t prim = __float_as_uint(r.v0.w);
        if (t < best || prim > h.prim || h.tri == 0xffffffffu) {
          best = t;
          h.t = t;
          h.u = u;
          h.v = v;
          h.tri = first + k;
          h.prim = prim;
          if (ANY) return true;
        }
      }
    }
    if (st.sp == 0) break;
    node = st.pop();
  }
  return h.tri != 0xffffffffu;
}

// ---- streaming traversal of a ray queue ---------------------------------------
// The persistent traversal kernels run one node (or one leaf) per loop
// iteration for every lane that holds a ray.  A lane whose ray is done takes
// the next ray at once, instead of idling until the wave's slowest ray is
// done: when at least `refill_min` lanes are idle, the wave hands them rays
// from its current chunk of the queue, taking a new chunk of `chunk` rays from
// `counter` (one atomic per wave) when the chunk runs out.  Per ray the result
// is exactly traverse<ANY>'s.
constexpr uint32_t kNoRay = 0xffffffffu;

// load(i, o, d, tmin, tmax, nodes, tris) reads ray i and the acceleration
// structure it is traced against; done(i, hit, h) consumes its result.
template <bool ANY, class Load, class Done>
PHX_DEV void traverse_stream(LaneStack &st, uint32_t *counter, uint32_t count, uint32_t chunk, uint32_t refill_min,
                             Load load, Done done) {
  const uint32_t lane = __lane_id();
  uint32_t next = 0, end = 0;  // the wave's current chunk [next, end)
  bool exhausted = false;
  uint32_t idx = kNoRay;
  vec3 o{0.f, 0.f, 0.f}, d{0.f, 0.f, 1.f}, inv{0.f, 0.f, 0.f};
  float tmin = 0.f, best = 0.f;
  int node = 0;
  const Node4 *nodes = nullptr;
  const TriRec *tris = nullptr;
  TraceHit h;
  h.tri = h.prim = 0xffffffffu;
  h.t = h.u = h.v = 0.f;
  while (true) {
    const uint64_t idle = __ballot(idx == kNoRay);
    const uint32_t nidle = uint32_t(__popcll(idle));
    if (nidle == 64u && exhausted) break;
    if (!exhausted && (nidle >= refill_min || nidle == 64u)) {
      const uint32_t avail = end - next;
      uint32_t nb = kNoRay;
      if (nidle > avail) {
        uint32_t b = 0;
        if (lane == 0) b = atomicAdd(counter, chunk);
        nb = __shfl(b, 0);
        if (nb >= count) nb = kNoRay;
      }
      if (idx == kNoRay) {
        const uint32_t rank = uint32_t(__popcll(idle & ((uint64_t(1) << lane) - 1)));
        uint32_t i = kNoRay;
        if (rank < avail) i = next + rank;
        else if (nb != kNoRay) i = nb + (rank - avail);
        if (i < count) {
          float tmax;
          load(i, o, d, tmin, tmax, nodes, tris);
          h.tri = h.prim = 0xffffffffu;
          h.t = tmax;
          h.u = h.v = 0.f;
          if (!(tmax > tmin)) {
            done(i, false, h);
          } else {
            idx = i;
            inv = vec3{safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z)};
            best = tmax;
            node = 0;
            st.sp = 0;
          }
        }
      }
      if (nidle <= avail) {
        next += nidle;
      } else if (nb != kNoRay) {
        next = nb + (nidle - avail);
        end = nb + chunk;
      } else {
        next = end;
        exhausted = true;
      }
      if (next >= count) {
        next = end;
        exhausted = true;
      }
    }
    if (idx != kNoRay) {
      bool finished = false;
      bool pop = true;
      if (node >= 0) {
        const Node4 n = nodes[node];
        float tn[4];
        bool hb[4];
        box4(n, o, inv, tmin, best, tn, hb);
        float k0 = hb[0] ? tn[0] : INFINITY, k1 = hb[1] ? tn[1] : INFINITY;
        float k2 = hb[2] ? tn[2] : INFINITY, k3 = hb[3] ? tn[3] : INFINITY;
        int c0 = n.child.x, c1 = n.child.y, c2 = n.child.z, c3 = n.child.w;
        const int cnt = int(hb[0]) + int(hb[1]) + int(hb[2]) + int(hb[3]);
        if (cnt > 0) {
          cswap(k0, c0, k1, c1);
          cswap(k2, c2, k3, c3);
          cswap(k0, c0, k2, c2);
          cswap(k1, c1, k3, c3);
          cswap(k1, c1, k2, c2);
          if (cnt > 3) st.push(c3);
          if (cnt > 2) st.push(c2);
          if (cnt > 1) st.push(c1);
          node = c0;
          pop = false;
        }
      } else {
        const uint32_t code = ~uint32_t(node);
        const uint32_t first = code >> 3, cnt = (code & 7u) + 1u;
        for (uint32_t k = 0; k < cnt; ++k) {
          const TriRec r = tris[first + k];
          float t, u, v;
          if (!tri_test(r, o, d, tmin, best, t, u, v)) continue;
          const uint32_t prim = __float_as_uint(r.v0.w);
          if (t < best || prim > h.prim || h.tri == 0xffffffffu) {
            best = t;
            h.t = t;
            h.u = u;
            h.v = v;
            h.tri = first + k;
            h.prim = prim;
            if (ANY) {
              finished = true;
              break;
            }
          }
        }
      }
      if (pop && !finished) {
        if (st.sp == 0) finished = true;
        else node = st.pop();
      }
      if (finished) {
        done(idx, h.tri != 0xffffffffu, h);
        idx = kNoRay;
      }
    }
  }
}

}  // namespace phx
