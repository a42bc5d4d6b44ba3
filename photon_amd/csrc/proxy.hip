// The reference's algorithmic structure on this GPU: the baseline the
// streaming backend is measured against (the reference's Kokkos-HIP build
// needs Kokkos and cannot be built offline).  It follows PathTracer::render
// (pathtracer.cpp:564-1159) with KokkosBackend (kokkos_backend.cpp:21-86):
//   - binary median-split BVH, 44-byte nodes, 64-entry per-thread stack
//     (proxy_bvh.h, bvh.cpp:20-125, mesh_intersector.h:57-124);
//   - one sample of every pixel per pass; each pass runs max_depth + 9
//     iterations of trace_closest, pt_shade, trace_occluded and
//     pt_nee_accum_pack over all W*H slots, finished rays included;
//   - trace_closest writes the full 88-byte hit record (HitResult) per slot;
//   - the host reads the active count back after every iteration.
// Shading is the same device code as the streaming backend (shade_path), so
// both render the same image; only the structure differs.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "context.h"
#include "dev_queue.h"
#include "proxy_bvh.h"

namespace {

using namespace phx;
using namespace phx_detail;

struct ProxyState {
  float4 *o, *d, *tp;
  uint4 *st;
  phx_hit *hit;
  float4 *sh_o, *sh_d, *sh_c;
  uint32_t *nee_active, *occluded, *active;
  float4 *accum;       // per pixel (radiance, alpha)
  uint32_t *count;     // active paths after the iteration
  uint32_t *aov_s;
};

constexpr uint32_t kPrimary = 0x10000u;

PHX_DEV vec3 xyz(float4 v) { return {v.x, v.y, v.z}; }
PHX_DEV float4 f4(vec3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

__global__ void k_px_init(ShadeParams w, ProxyState s, uint32_t sample_index, uint32_t first) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= w.npix) return;
  if (first) s.accum[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
  const uint32_t x = idx % w.width, y = idx / w.width;
  Rng rng = make_rng(idx, w.sample_offset + sample_index, 0u);
  float u = (float(x) + rng.next()) / float(w.width);
  float v = (float(y) + rng.next()) / float(w.height);
  if (!w.bottom_up) v = 1.f - v;
  u = w.region[0] + (w.region[2] - w.region[0]) * u;
  v = w.region[1] + (w.region[3] - w.region[1]) * v;
  vec3 o, d;
  camera_ray(w.cam, u, v, rng, o, d);
  float tn, tf;
  clip_range(w.cam, d, tn, tf);
  s.o[idx] = f4(o, tn);
  s.d[idx] = f4(d, tf);
  s.tp[idx] = make_float4(1.f, 1.f, 1.f, -1.f);
  s.st[idx] = make_uint4(0u, kPrimary, 0u, 0u);
  s.active[idx] = 1u;
  s.sh_o[idx] = make_float4(0.f, 0.f, 0.f, 1.f);
  s.sh_d[idx] = make_float4(0.f, 0.f, 1.f, 0.f);
}

// trace_closest over all slots, full hit record (kokkos_backend.cpp:34-59).
__global__ void k_px_trace(ProxyScene ps, DevScene s, ProxyState st, uint32_t n) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const float4 o = st.o[idx], d = st.d[idx];
  TraceHit h;
  const bool hit = px_traverse<false>(ps, xyz(o), xyz(d), o.w, d.w, h);
  phx_hit &r = st.hit[idx];
  r.hit = hit ? 1u : 0u;
  if (!hit) return;
  const uint32_t prim = h.prim;
  const vec3 v0 = ld3(s.positions + size_t(s.indices[prim * 3]) * 3);
  const vec3 v1 = ld3(s.positions + size_t(s.indices[prim * 3 + 1]) * 3);
  const vec3 v2 = ld3(s.positions + size_t(s.indices[prim * 3 + 2]) * 3);
  const vec3 p = xyz(o) + xyz(d) * h.t;
  const vec3 nn = normalize(cross(v1 - v0, v2 - v0));
  vec3 sn = nn;
  const uint32_t i0 = s.indices[prim * 3], i1 = s.indices[prim * 3 + 1], i2 = s.indices[prim * 3 + 2];
  const float w0 = 1.f - h.u - h.v;
  if (s.normals) {
    sn = normalize(ld3(s.normals + size_t(i0) * 3) * w0 + ld3(s.normals + size_t(i1) * 3) * h.u +
                   ld3(s.normals + size_t(i2) * 3) * h.v);
    if (dot(sn, nn) < 0.f) sn = sn * -1.f;
  }
  r.t = h.t;
  r.position[0] = p.x; r.position[1] = p.y; r.position[2] = p.z;
  r.normal[0] = nn.x; r.normal[1] = nn.y; r.normal[2] = nn.z;
  r.shading_normal[0] = sn.x; r.shading_normal[1] = sn.y; r.shading_normal[2] = sn.z;
  r.bary[0] = h.u;
  r.bary[1] = h.v;
  r.prim_id = prim;
  r.material_id = s.material_ids ? s.material_ids[prim] : 0u;
}

// pt_shade over all slots (pathtracer.cpp:765-1090), shade_path on the slot.
__global__ void k_px_shade(DevScene s, ShadeParams w, ProxyState st, uint32_t it, uint32_t sample_index) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= w.npix) return;
  st.nee_active[idx] = 0u;
  if (st.active[idx] == 0u) return;
  const phx_hit hr = st.hit[idx];
  const float4 o = st.o[idx], d = st.d[idx], tp = st.tp[idx];
  const uint4 u = st.st[idx];
  PathState p;
  p.ro = xyz(o);
  p.tmin = o.w;
  p.rd = xyz(d);
  p.tmax = d.w;
  p.tp = xyz(tp);
  p.prev_pdf = tp.w;
  p.depth = u.y & 0xffu;
  p.path_it = it;
  p.primary = (u.y & kPrimary) != 0u;
  p.kinds = u.z;
  p.path_dist = __uint_as_float(u.w);
  p.item = 0;
  p.pixel = idx;
  Rng rng = make_rng(idx, w.sample_offset + sample_index, 1u + it);
  ShadeOut out;
  shade_path(s, w, p, hr.hit != 0u, hr.prim_id, hr.t, hr.bary[0], hr.bary[1], rng, out);
  if (out.add.x != 0.f || out.add.y != 0.f || out.add.z != 0.f || out.add_alpha != 0.f) {
    const float4 a = st.accum[idx];
    st.accum[idx] = make_float4(a.x + out.add.x, a.y + out.add.y, a.z + out.add.z, a.w + out.add_alpha);
  }
  if (out.primary_event) atomicMin(&st.aov_s[idx], sample_index);
  if (out.shadow) {
    st.nee_active[idx] = 1u;
    st.sh_o[idx] = f4(out.so, 0.f);
    st.sh_d[idx] = f4(out.sd, out.stmax);
    st.sh_c[idx] = f4(out.sc, 0.f);
  }
  if (out.cont) {
    const PathState &n = out.next;
    st.o[idx] = f4(n.ro, n.tmin);
    st.d[idx] = f4(n.rd, n.tmax);
    st.tp[idx] = f4(n.tp, n.prev_pdf);
    st.st[idx] = make_uint4(0u, (n.depth & 0xffu) | (n.primary ? kPrimary : 0u), n.kinds, __float_as_uint(n.path_dist));
  } else {
    st.active[idx] = 0u;
  }
}

// trace_occluded over all slots (kokkos_backend.cpp:74-84).
__global__ void k_px_occluded(ProxyScene ps, ProxyState st, uint32_t n) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const float4 o = st.sh_o[idx], d = st.sh_d[idx];
  TraceHit h;
  st.occluded[idx] = px_traverse<true>(ps, xyz(o), xyz(d), o.w, d.w, h) ? 1u : 0u;
}

// pt_nee_accum_pack (pathtracer.cpp:1095-1123): NEE, empty rays, active count.
__global__ void k_px_pack(DevScene s, ShadeParams w, ProxyState st, uint32_t it, uint32_t sample_index) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  bool act = false;
  if (idx < w.npix) {
    if (st.nee_active[idx] != 0u && st.occluded[idx] == 0u) {
      vec3 c = xyz(st.sh_c[idx]);
      if (s.vol_count > 0) {
        Rng rng = make_rng(idx, w.sample_offset + sample_index, 0x10000u + it);
        const float4 o = st.sh_o[idx], d = st.sh_d[idx];
        c = c * vol_transmittance(s, xyz(o), xyz(d), o.w, d.w, rng);
      }
      if (finite3(c)) {
        const float4 a = st.accum[idx];
        st.accum[idx] = make_float4(a.x + c.x, a.y + c.y, a.z + c.z, a.w);
      }
    }
    st.sh_o[idx] = make_float4(0.f, 0.f, 0.f, 1.f);
    st.sh_d[idx] = make_float4(0.f, 0.f, 1.f, 0.f);
    act = st.active[idx] != 0u;
    if (!act) {
      st.o[idx].w = 1.f;
      st.d[idx].w = 0.f;
    }
  }
  const uint64_t m = __ballot(act);
  if (lane_id() == 0 && m) atomicAdd(st.count, uint32_t(__popcll(m)));
}

__global__ void k_px_count_reset(uint32_t *count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *count = 0u;
}

__global__ void k_px_copy_accum(uint32_t npix, const float4 *src, float4 *dst) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < npix) dst[i] = src[i];
}

}  // namespace

extern "C" int phx_render_reference_structure(phx_context *c, const phx_camera *cam, const phx_render_params *p) {
  if (!c || !cam || !p) return PHX_ERR_INVALID;
  if (!c->have_scene) return fail(c, PHX_ERR_NO_SCENE, "render before set_scene");
  if (p->width == 0 || p->height == 0) return fail(c, PHX_ERR_INVALID, "render: empty frame");
  if (p->adaptive_sampling) return fail(c, PHX_ERR_UNSUPPORTED, "render: adaptive sampling is not offered");
  if (uint64_t(p->width) * p->height > (uint64_t(1) << 26)) return fail(c, PHX_ERR_INVALID, "render: frame too large");
  const uint32_t npix = p->width * p->height;
  const uint32_t spp = std::max<uint32_t>(p->samples_per_pixel, 1);
  const uint32_t max_depth = p->max_depth > 0 ? p->max_depth : 1u;
  const uint32_t max_it = max_depth + 1u + 8u;
  HIP_TRY(c, hipSetDevice(c->device));
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t stm = c->stream;
  int r = ensure_trace_scratch(c);
  if (r) return r;
  // The reference's BVH and prim-ordered triangle records, built once per scene.
  if (!c->px_built) {
    ProxyBvh bvh;
    build_proxy_bvh(c->host_positions.data(), c->host_indices.data(), c->triangle_count, bvh);
    std::vector<TriRec> recs(c->triangle_count);
    for (uint32_t t = 0; t < c->triangle_count; ++t) {
      const float *v0 = &c->host_positions[size_t(c->host_indices[t * 3]) * 3];
      const float *v1 = &c->host_positions[size_t(c->host_indices[t * 3 + 1]) * 3];
      const float *v2 = &c->host_positions[size_t(c->host_indices[t * 3 + 2]) * 3];
      float tbits;
      std::memcpy(&tbits, &t, 4);
      recs[t].v0 = make_float4(v0[0], v0[1], v0[2], tbits);
      recs[t].e1 = make_float4(v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2], 0.f);
      recs[t].e2 = make_float4(v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2], 0.f);
    }
    if ((r = upload(c, c->px_nodes, bvh.nodes.data(), bvh.nodes.size()))) return r;
    if ((r = upload(c, c->px_prims, bvh.prim_ids.data(), bvh.prim_ids.size()))) return r;
    if ((r = upload(c, c->px_tris, recs.data(), recs.size()))) return r;
    // the attribute records in the same (primitive) order
    std::vector<uint4> attr(c->triangle_count);
    for (uint32_t t = 0; t < c->triangle_count; ++t)
      attr[t] = make_uint4(c->host_indices[t * 3], c->host_indices[t * 3 + 1], c->host_indices[t * 3 + 2],
                           c->host_material_ids.empty() ? 0u : c->host_material_ids[t]);
    if ((r = upload(c, c->px_tri_attr, attr.data(), attr.size()))) return r;
    c->px_root = bvh.root;
    HIP_TRY(c, hipStreamSynchronize(stm));
    c->px_built = true;
  }
  ProxyScene ps{c->px_nodes.as<ProxyNode>(), c->px_prims.as<uint32_t>(), c->indices.as<uint32_t>(),
                c->positions.as<float>(), c->px_root};
  DevScene ds = c->dev;
  ds.tris = c->px_tris.as<TriRec>();
  ds.tri_attr = c->px_tri_attr.as<uint4>();
  ds.tri_ng = nullptr;  // the normals come from the primitive-order records

  // per-slot state (pathtracer.cpp:505-562)
  const size_t n = npix;
  if (c->px_slots < npix) {
    HIP_TRY(c, c->px_o.reserve(n * 16));
    HIP_TRY(c, c->px_d.reserve(n * 16));
    HIP_TRY(c, c->px_tp.reserve(n * 16));
    HIP_TRY(c, c->px_st.reserve(n * 16));
    HIP_TRY(c, c->px_hit.reserve(n * sizeof(phx_hit)));
    HIP_TRY(c, c->px_sh_o.reserve(n * 16));
    HIP_TRY(c, c->px_sh_d.reserve(n * 16));
    HIP_TRY(c, c->px_sh_c.reserve(n * 16));
    HIP_TRY(c, c->px_flags.reserve(n * 12));
    c->px_slots = npix;
  }
  HIP_TRY(c, c->px_accum.reserve(n * 16));
  HIP_TRY(c, c->px_count.reserve(64));
  HIP_TRY(c, c->aov_s.reserve(n * 4));
  if (!c->pinned) HIP_TRY(c, hipHostMalloc(reinterpret_cast<void **>(&c->pinned), 64, hipHostMallocDefault));
  if (c->f_w != p->width || c->f_h != p->height || !c->f_accum.p) {
    HIP_TRY(c, c->f_accum.reserve(n * 16));
    HIP_TRY(c, c->f_color.reserve(n * 12));
    HIP_TRY(c, c->f_alpha.reserve(n * 4));
    HIP_TRY(c, c->f_depth.reserve(n * 4));
    HIP_TRY(c, c->f_normal.reserve(n * 12));
    HIP_TRY(c, c->f_albedo.reserve(n * 12));
    c->f_w = p->width;
    c->f_h = p->height;
  }
  ProxyState st{};
  st.o = c->px_o.as<float4>();
  st.d = c->px_d.as<float4>();
  st.tp = c->px_tp.as<float4>();
  st.st = c->px_st.as<uint4>();
  st.hit = c->px_hit.as<phx_hit>();
  st.sh_o = c->px_sh_o.as<float4>();
  st.sh_d = c->px_sh_d.as<float4>();
  st.sh_c = c->px_sh_c.as<float4>();
  st.nee_active = c->px_flags.as<uint32_t>();
  st.occluded = st.nee_active + n;
  st.active = st.occluded + n;
  st.accum = c->px_accum.as<float4>();
  st.count = c->px_count.as<uint32_t>();
  st.aov_s = c->aov_s.as<uint32_t>();

  const ShadeParams sp = make_shade_params(*cam, *p, max_depth, max_it);
  const dim3 blk(256), grd((npix + 255) / 256);
  StreamQueues q{};
  q.aov_s = st.aov_s;
  q.spill = c->spill.as<int>();
  HIP_TRY(c, launch_aov_init(npix, q.aov_s, stm));
  for (uint32_t s = 0; s < spp; ++s) {
    hipLaunchKernelGGL(k_px_init, grd, blk, 0, stm, sp, st, s, s == 0 ? 1u : 0u);
    HIP_TRY(c, hipGetLastError());
    for (uint32_t it = 0; it < max_it; ++it) {
      hipLaunchKernelGGL(k_px_trace, grd, blk, 0, stm, ps, ds, st, npix);
      hipLaunchKernelGGL(k_px_shade, grd, blk, 0, stm, ds, sp, st, it, s);
      hipLaunchKernelGGL(k_px_occluded, grd, blk, 0, stm, ps, st, npix);
      hipLaunchKernelGGL(k_px_count_reset, dim3(1), dim3(64), 0, stm, st.count);
      hipLaunchKernelGGL(k_px_pack, grd, blk, 0, stm, ds, sp, st, it, s);
      HIP_TRY(c, hipGetLastError());
      // the host reads the active count every iteration (pathtracer.cpp:1095-1126)
      HIP_TRY(c, hipMemcpyAsync(c->pinned, st.count, 4, hipMemcpyDeviceToHost, stm));
      HIP_TRY(c, hipStreamSynchronize(stm));
      if (c->pinned[0] == 0) break;
    }
  }
  hipLaunchKernelGGL(k_px_copy_accum, grd, blk, 0, stm, npix, st.accum, c->f_accum.as<float4>());
  HIP_TRY(c, hipGetLastError());
  const FrameBuffers fb{c->f_accum.as<float4>(), c->f_color.as<float>(), c->f_alpha.as<float>(),
                        c->f_depth.as<float>(), c->f_normal.as<float>(), c->f_albedo.as<float>()};
  HIP_TRY(c, launch_clear_u32(c->fetch.as<uint32_t>(), 1, stm));
  HIP_TRY(c, launch_resolve(c->dev, sp, q, fb, spp, c->fetch.as<uint32_t>(), c->trace_grid, stm));
  HIP_TRY(c, hipStreamSynchronize(stm));
  const auto t1 = std::chrono::steady_clock::now();
  c->stats.closest_rays = c->stats.shadow_rays = c->stats.shade_events = 0;
  c->stats.samples = uint64_t(npix) * spp;
  c->stats.pixels = npix;
  c->stats.last_render_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  c->stats.extend_launches = c->stats.shade_launches = c->stats.shadow_launches = 0;
  c->stats.extend_ms = c->stats.shade_ms = c->stats.shadow_ms = 0.0;
  c->have_result = true;
  c->f_counted = false;
  return PHX_OK;
}
