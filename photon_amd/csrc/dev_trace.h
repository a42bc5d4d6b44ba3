// Software BVH traversal for gfx950 (no ray accelerators on CDNA4).
//
// One lane traces one ray through the 8-wide quantised BVH.  The traversal stack is a
// short per-lane stack in LDS (stack[depth][lane], so a wave's 64 lanes touch
// 64 consecutive dwords at any depth: no bank conflicts) that spills to a
// per-thread region of global memory beyond kShortStack entries; the builder
// guarantees the total never exceeds kMaxStack (the reference's stack has 64
// entries and silently drops pushes beyond them, mesh_intersector.h:74).
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

#include "dev_queue.h"
#include "dev_scene.h"

namespace phx {

constexpr int kShortStack = 16;
// Worst case of an 8-wide tree of depth D: 7 D + 1 entries (the builder
// rejects trees deeper than that allows).
constexpr int kMaxStack = 128;

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
PHX_DEV bool tri_test(vec3 v0, vec3 e1, vec3 e2, vec3 o, vec3 d, float tmin, float tmax, float &t, float &u,
                      float &v) {
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
PHX_DEV bool tri_test(const TriRec &r, vec3 o, vec3 d, float tmin, float tmax, float &t, float &u, float &v) {
  return tri_test(vec3{r.v0.x, r.v0.y, r.v0.z}, vec3{r.e1.x, r.e1.y, r.e1.z}, vec3{r.e2.x, r.e2.y, r.e2.z}, o, d,
                  tmin, tmax, t, u, v);
}

// The header fields of a node that turn a child slot into its stack entry.
struct NodeView {
  uint32_t child_base;
  uint32_t leaf_base;  // tri_base - ninner
  uint32_t leaf_cnt;
  uint32_t ninner;
  // stack entry of child slot k: inner slots are nodes child_base + k; leaf
  // slot k starts after the triangles of the leaf slots before it (the 2-bit
  // counts of inner and unused slots are zero, so the sum of the 2-bit fields
  // below slot k counts the leaf slots' extra triangles)
  PHX_DEV int code(uint32_t k) const {
    if (k < ninner) return int(child_base + k);
    const uint32_t below = leaf_cnt & ((1u << (2u * k)) - 1u);
    const uint32_t first = leaf_base + k + __popc(below) + __popc(below & 0xaaaau);
    return int(~((first << 3) | ((leaf_cnt >> (2u * k)) & 3u)));
  }
};
// Packed child key: the entry distance's bits (clamped at 0, so the order of
// non-negative floats is that of their bits) with the slot in the low 3 bits;
// missed children hold kMissKey | slot, above every entered one.
constexpr uint32_t kMissKey = 0xfffffff8u;

// Slab test of a node's children against their quantised boxes.  A plane at
// org + q * 2^e is met at t = (org + q * 2^e - o) * inv = q * A + B with
// A = 2^e * inv (exact) and B = (org - o) * inv, one fused multiply-add per
// plane; the near and far planes of each axis are picked once per node from
// the sign of the direction.  Entry distances are clamped to [tmin, tmax]; the
// exit is widened by a few ulps so rounding never culls a box the exact test
// would enter.  `w` holds the node's 80 bytes (five 16-byte words).  The
// children's packed keys go to pk, nearest first when SORT, in slot order
// otherwise.
template <bool SORT>
PHX_DEV void visit_node(const uint4 w[5], vec3 o, vec3 inv, float tmin, float tmax, NodeView &nv,
                       uint32_t pk[kNodeWidth]) {
  const vec3 org{__uint_as_float(w[0].x), __uint_as_float(w[0].y), __uint_as_float(w[0].z)};
  const int ex = int(int8_t(w[0].w & 0xffu)), ey = int(int8_t((w[0].w >> 8) & 0xffu)),
            ez = int(int8_t((w[0].w >> 16) & 0xffu));
  nv = NodeView{w[1].x, w[1].y - (w[0].w >> 24), w[1].z & 0xffffu, w[0].w >> 24};
  const uint32_t nchild = (w[1].z >> 16) & 0xffu;
  const float Ax = ldexpf(inv.x, ex), Ay = ldexpf(inv.y, ey), Az = ldexpf(inv.z, ez);
  const float Bx = (org.x - o.x) * inv.x, By = (org.y - o.y) * inv.y, Bz = (org.z - o.z) * inv.z;
  const uint2 lx{w[2].x, w[2].y}, ly{w[2].z, w[2].w}, lz{w[3].x, w[3].y};
  const uint2 hx{w[3].z, w[3].w}, hy{w[4].x, w[4].y}, hz{w[4].z, w[4].w};
  const uint2 nxw = inv.x >= 0.f ? lx : hx, fxw = inv.x >= 0.f ? hx : lx;
  const uint2 nyw = inv.y >= 0.f ? ly : hy, fyw = inv.y >= 0.f ? hy : ly;
  const uint2 nzw = inv.z >= 0.f ? lz : hz, fzw = inv.z >= 0.f ? hz : lz;
  auto q16 = [](const uint2 &v, int k) {
    const uint32_t d = k < 4 ? v.x : v.y;
    return float((d >> (8 * (k & 3))) & 0xffu);
  };
#pragma unroll
  for (int k = 0; k < kNodeWidth; ++k) {
    const float a = fmaxf(fmaxf(fmaf(q16(nxw, k), Ax, Bx), fmaf(q16(nyw, k), Ay, By)),
                          fmaxf(fmaf(q16(nzw, k), Az, Bz), tmin));
    const float b = fminf(fminf(fmaf(q16(fxw, k), Ax, Bx), fmaf(q16(fyw, k), Ay, By)),
                          fminf(fmaf(q16(fzw, k), Az, Bz), tmax));
    const bool hit = k < int(nchild) && a <= b * 1.0000004f;
    // a negative entry distance (tmin < 0) is a negative int: clamped to 0
    pk[k] = (hit ? (uint32_t(max(int(__float_as_uint(a)), 0)) & ~7u) : kMissKey) | uint32_t(k);
  }
  if (SORT) {
    // Batcher's odd-even merge network on the packed keys: a min and a max
    // per exchange
    auto cx = [](uint32_t &x, uint32_t &y) {
      const uint32_t lo = min(x, y), hi = max(x, y);
      x = lo;
      y = hi;
    };
    cx(pk[0], pk[1]); cx(pk[2], pk[3]); cx(pk[4], pk[5]); cx(pk[6], pk[7]);
    cx(pk[0], pk[2]); cx(pk[1], pk[3]); cx(pk[4], pk[6]); cx(pk[5], pk[7]);
    cx(pk[1], pk[2]); cx(pk[5], pk[6]);
    cx(pk[0], pk[4]); cx(pk[1], pk[5]); cx(pk[2], pk[6]); cx(pk[3], pk[7]);
    cx(pk[2], pk[4]); cx(pk[3], pk[5]);
    cx(pk[1], pk[2]); cx(pk[3], pk[4]); cx(pk[5], pk[6]);
  }
}
template <bool SORT>
PHX_DEV void visit_node(const BvhNode *nodes, int node, vec3 o, vec3 inv, float tmin, float tmax, NodeView &nv,
                       uint32_t pk[kNodeWidth]) {
  // the node as five 16-byte loads, unpacked in registers
  const uint4 *src = reinterpret_cast<const uint4 *>(nodes + node);
  uint4 w[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) w[i] = src[i];
  visit_node<SORT>(w, o, inv, tmin, tmax, nv, pk);
}

// Pushes the entered children but the nearest, farthest first (pk sorted);
// the nearest becomes the next node.
PHX_DEV int descend(LaneStack &st, const NodeView &nv, const uint32_t pk[kNodeWidth]) {
#pragma unroll
  for (int k = kNodeWidth - 1; k >= 1; --k)
    if (pk[k] < kMissKey) st.push(nv.code(pk[k] & 7u));
  return nv.code(pk[0] & 7u);
}
// Keys in slot order: the nearest entered child becomes the next node, the
// other entered ones are pushed in slot order.
PHX_DEV int descend_unsorted(LaneStack &st, const NodeView &nv, const uint32_t pk[kNodeWidth], uint32_t best) {
#pragma unroll
  for (int k = kNodeWidth - 1; k >= 0; --k)
    if (pk[k] < kMissKey && pk[k] != best) st.push(nv.code(uint32_t(k)));
  return nv.code(best & 7u);
}
// The nearest entered child's key: kMissKey or above when none was entered.
template <bool SORT>
PHX_DEV uint32_t nearest_key(const uint32_t pk[kNodeWidth]) {
  if (SORT) return pk[0];
  uint32_t best = pk[0];
#pragma unroll
  for (int k = 1; k < kNodeWidth; ++k) best = min(best, pk[k]);
  return best;
}
template <bool SORT>
PHX_DEV int descend_any(LaneStack &st, const NodeView &nv, const uint32_t pk[kNodeWidth], uint32_t best) {
  return SORT ? descend(st, nv, pk) : descend_unsorted(st, nv, pk, best);
}

// Triangle `k` of the leaf order against the ray: updates the closest hit h
// (tie rule: equal t goes to the larger primitive id); returns true when it
// becomes the hit.
// `w` holds the record's 48 bytes: v0 with the primitive id, e1, e2.
PHX_DEV bool visit_tri(const uint4 w[3], uint32_t k, vec3 o, vec3 d, float tmin, float &best, TraceHit &h) {
  const vec3 v0{__uint_as_float(w[0].x), __uint_as_float(w[0].y), __uint_as_float(w[0].z)};
  const vec3 e1{__uint_as_float(w[1].x), __uint_as_float(w[1].y), __uint_as_float(w[1].z)};
  const vec3 e2{__uint_as_float(w[2].x), __uint_as_float(w[2].y), __uint_as_float(w[2].z)};
  float t, u, v;
  if (!tri_test(v0, e1, e2, o, d, tmin, best, t, u, v)) return false;
  const uint32_t prim = w[0].w;
  if (t < best || prim > h.prim || h.tri == 0xffffffffu) {
    best = t;
    h.t = t;
    h.u = u;
    h.v = v;
    h.tri = k;
    h.prim = prim;
    return true;
  }
  return false;
}
PHX_DEV bool visit_tri(const TriRec *tris, uint32_t k, vec3 o, vec3 d, float tmin, float &best, TraceHit &h) {
  const uint4 *src = reinterpret_cast<const uint4 *>(tris + k);
  const uint4 w[3] = {src[0], src[1], src[2]};
  return visit_tri(w, k, o, d, tmin, best, h);
}
// The triangles of a leaf entry against the ray; returns true on a hit when
// ANY (any hit ends the ray).
template <bool ANY>
PHX_DEV bool visit_leaf(const TriRec *tris, int node, vec3 o, vec3 d, float tmin, float &best, TraceHit &h) {
  const uint32_t code = ~uint32_t(node);
  const uint32_t first = code >> 3, count = (code & 7u) + 1u;
  for (uint32_t k = 0; k < count; ++k)
    if (visit_tri(tris, first + k, o, d, tmin, best, h) && ANY) return true;
  return false;
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
      NodeView nv;
      uint32_t pk[kNodeWidth];
      visit_node<!ANY>(s.nodes, node, o, inv, tmin, best, nv, pk);
      const uint32_t near = nearest_key<!ANY>(pk);
      if (near < kMissKey) {
        node = descend_any<!ANY>(st, nv, pk, near);
        continue;
      }
    } else if (visit_leaf<ANY>(s.tris, node, o, d, tmin, best, h)) {
      return true;
    }
    if (st.sp == 0) break;
    node = st.pop();
  }
  return h.tri != 0xffffffffu;
}

// ---- streaming traversal of a ray queue ---------------------------------------
// The persistent traversal kernels run one node (or one leaf triangle) per
// loop iteration for every lane that holds a ray.  A lane whose ray is done takes
// the next ray at once, instead of idling until the wave's slowest ray is
// done: when at least `refill_min` lanes are idle, the wave hands them rays
// from its current chunk of the queue, taking a new chunk of up to `chunk`
// rays from the fetch set `set` (fetch_chunk: one atomic per wave and chunk,
// spread over the queue's stripes) when the chunk runs out.  Per ray the
// result is exactly traverse<ANY>'s.
constexpr uint32_t kNoRay = 0xffffffffu;

// load(i, o, d, tmin, tmax, nodes, tris) reads ray i and the acceleration
// structure it is traced against; done(i, hit, h) consumes its result.  With
// ONE_ACCEL every ray is traced against `accel_nodes` / `accel_tris` (the
// launch's one scene: the pointers stay wave-uniform, in scalar registers)
// and load's last two outputs are ignored.  Every thread of the work-group
// must call it (it synchronises the work-group once, at the start).
template <bool ANY, bool ONE_ACCEL = false, class Load, class Done>
PHX_DEV void traverse_stream(LaneStack &st, uint32_t *set, uint32_t count, uint32_t chunk, uint32_t refill_min,
                             Load load, Done done, const BvhNode *accel_nodes = nullptr,
                             const TriRec *accel_tris = nullptr) {
  const uint32_t lane = __lane_id();
  uint32_t next = 0, end = 0;  // the wave's current chunk [next, end)
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t units = fetch_units(count, chunk);
  uint32_t stripe = wave % kStripes, left = fetch_walk(min(units, gridDim.x * blockDim.x / 64u));
  bool exhausted = wave >= units;  // waves beyond the fetchers take no part
  __shared__ uint32_t s_used;      // stripes the work-group's waves found used up
  if (threadIdx.x == 0) s_used = 0u;
  __syncthreads();
  uint32_t idx = kNoRay;
  vec3 o{0.f, 0.f, 0.f}, d{0.f, 0.f, 1.f}, inv{0.f, 0.f, 0.f};
  float tmin = 0.f, best = 0.f;
  int node = 0;
  const BvhNode *nodes = accel_nodes;
  const TriRec *tris = accel_tris;
  TraceHit h;
  h.tri = h.prim = 0xffffffffu;
  h.t = h.u = h.v = 0.f;
  while (true) {
    const uint64_t idle = __ballot(idx == kNoRay);
    const uint32_t nidle = uint32_t(__popcll(idle));
    if (nidle == 64u && exhausted) break;
    if (!exhausted && (nidle >= refill_min || nidle == 64u)) {
      const uint32_t avail = end - next;
      uint32_t nb = kNoRay, nbe = 0;  // a new chunk [nb, nbe)
      if (nidle > avail) fetch_chunk(set, &s_used, count, units, chunk, stripe, left, nb, nbe);
      if (idx == kNoRay) {
        const uint32_t rank = uint32_t(__popcll(idle & ((uint64_t(1) << lane) - 1)));
        uint32_t i = kNoRay;
        if (rank < avail) i = next + rank;
        else if (nb != kNoRay && rank - avail < nbe - nb) i = nb + (rank - avail);
        if (i != kNoRay) {
          float tmax;
          const BvhNode *ray_nodes = nullptr;
          const TriRec *ray_tris = nullptr;
          load(i, o, d, tmin, tmax, ray_nodes, ray_tris);
          if (!ONE_ACCEL) {
            nodes = ray_nodes;
            tris = ray_tris;
          }
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
        next = min(nb + (nidle - avail), nbe);
        end = nbe;
      } else {
        next = end;
        exhausted = true;
      }
    }
    if (idx != kNoRay) {
      bool finished = false;
      bool pop = true;
      // One fetch per iteration for node and triangle lanes alike, issued
      // before either test: a wave whose lanes hold both waits on memory once
      // per iteration, not once per kind.  A triangle lane reads its 48-B
      // record and the 32 B after it (the triangle array is padded past its
      // last record).
      const uint32_t code = ~uint32_t(node);
      const uint32_t first = code >> 3;
      const uint4 *src = node >= 0 ? reinterpret_cast<const uint4 *>(nodes + node)
                                   : reinterpret_cast<const uint4 *>(tris + first);
      uint4 w[5];
#pragma unroll
      for (int i = 0; i < 5; ++i) w[i] = src[i];
      asm volatile("" ::"v"(w[3].x), "v"(w[3].y), "v"(w[3].z), "v"(w[3].w), "v"(w[4].x), "v"(w[4].y), "v"(w[4].z),
                   "v"(w[4].w));
      if (node >= 0) {
        NodeView nv;
        uint32_t pk[kNodeWidth];
        visit_node<!ANY>(w, o, inv, tmin, best, nv, pk);
        const uint32_t near = nearest_key<!ANY>(pk);
        if (near < kMissKey) {
          node = descend_any<!ANY>(st, nv, pk, near);
          pop = false;
        }
      } else {
        // one triangle per iteration: the leaf entry shrinks in place until
        // its last triangle, so a wave whose lanes hold leaves of different
        // sizes pays one triangle test per iteration, not the largest leaf
        const uint32_t count = (code & 7u) + 1u;
        finished = visit_tri(w, first, o, d, tmin, best, h) && ANY;
        if (!finished && count > 1) {
          node = leaf_code(first + 1u, count - 1u);
          pop = false;
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
