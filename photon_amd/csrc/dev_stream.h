// Device helpers shared by the wavefront kernels (integrator.hip) and the
// path-replay adjoint (adjoint.hip): queue-entry packing, camera paths and
// the traversal scene of a lane.
#pragma once

#include <type_traits>

#include "dev_trace.h"
#include "integrator.h"

namespace phx {

constexpr uint32_t kPrimaryBit = 0x10000u;

PHX_DEV vec3 xyz(float4 v) { return {v.x, v.y, v.z}; }
PHX_DEV float4 f4(vec3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

PHX_DEV PathState unpack(float4 o, float4 d, float4 tp, uint4 st) {
  PathState p;
  p.ro = xyz(o);
  p.tmin = o.w;
  p.rd = xyz(d);
  p.tmax = d.w;
  p.tp = xyz(tp);
  p.prev_pdf = tp.w;
  p.depth = st.y & 0xffu;
  p.path_it = (st.y >> 8) & 0xffu;
  p.primary = (st.y & kPrimaryBit) != 0u;
  p.item = st.y >> kItemShift;
  p.kinds = st.z;
  p.path_dist = __uint_as_float(st.w);
  return p;
}
PHX_DEV uint32_t pack_flags(const PathState &p) {
  return (p.depth & 0xffu) | ((p.path_it & 0xffu) << 8) | (p.primary ? kPrimaryBit : 0u) | (p.item << kItemShift);
}

// Camera ray of sample s of a pixel (pathtracer.cpp:727-759).
PHX_DEV void camera_path(const ShadeParams &w, uint32_t pixel, uint32_t sample, PathState &p) {
  const uint32_t x = pixel % w.width, y = pixel / w.width;
  Rng rng = make_rng(pixel, sample, 0u);
  float u = (float(x) + rng.next()) / float(w.width);
  float v = (float(y) + rng.next()) / float(w.height);
  if (!w.bottom_up) v = 1.f - v;
  u = w.region[0] + (w.region[2] - w.region[0]) * u;
  v = w.region[1] + (w.region[3] - w.region[1]) * v;
  camera_ray(w.cam, u, v, rng, p.ro, p.rd);
  clip_range(w.cam, p.rd, p.tmin, p.tmax);
  p.tp = vec3{1.f, 1.f, 1.f};
  p.prev_pdf = -1.f;
  p.depth = 0;
  p.path_it = 0;
  p.kinds = 0;
  p.primary = true;
  p.path_dist = 0.f;
  p.item = 0;
}

// Traversal scene of a lane: the uniform kernel-argument scene, or the two
// pointers traversal reads, copied out of the scene-table entry of the lane's
// item.
template <class P>
PHX_DEV DevScene trace_scene(const P &v, uint32_t item) {
  if constexpr (std::is_same<P, OneItem>::value || std::is_same<P, ItemTable<true>>::value) {
    return v.scene;
  } else {
    const DevScene &sc = v.scene_of(v.item_at(item));
    DevScene ls{};
    ls.nodes = sc.nodes;
    ls.tris = sc.tris;
    return ls;
  }
}

}  // namespace phx
