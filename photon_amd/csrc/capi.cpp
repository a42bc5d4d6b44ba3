// C ABI of the gfx950 path-tracing backend (include/phx.h): context, scene
// upload, light tables, the wavefront render loop and the render-bridge
// accumulation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/phx.h"
#include "bvh_build.h"
#include "context.h"
#include "dev_trace.h"
#include "integrator.h"

namespace {

using namespace phx;
using namespace phx_detail;

float length3(const float *v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// Emissive-triangle test of the render bridge (material_may_emit,
// material.h:97-103): also true when the emission has a source.
bool may_emit(const phx_material &m) {
  if (m.emission_strength <= 0.f) return false;
  return m.emission_src.kind != PHX_SOURCE_CONSTANT || m.emission_tex >= 0 || m.emission[0] > 0.f ||
         m.emission[1] > 0.f || m.emission[2] > 0.f;
}
bool has_source(const phx_material &m) {
  for (const phx_param_source &p : {m.base_color_src, m.alpha_src, m.metallic_src, m.roughness_src, m.emission_src,
                                    m.transmission_src, m.normal_src})
    if (p.kind != PHX_SOURCE_CONSTANT) return true;
  return false;
}

// Marginal/conditional CDFs of an equirect map (environment_map.cpp:8-65),
// single precision like the reference.
void env_cdf(const float *px, uint32_t w, uint32_t h, std::vector<float> &marg, std::vector<float> &cond) {
  const float PI = 3.14159265358979323846f;
  marg.assign(h + 1, 0.f);
  cond.assign(size_t(h) * (w + 1), 0.f);
  std::vector<float> roww(h, 0.f);
  for (uint32_t y = 0; y < h; ++y) {
    float *c = &cond[size_t(y) * (w + 1)];
    const float theta = (float(y) + 0.5f) * (PI / float(h));
    const float st = std::sin(theta);
    float sum = 0.f;
    for (uint32_t x = 0; x < w; ++x) {
      const float *p = px + (size_t(y) * w + x) * 3;
      sum += 0.2126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2];
      c[x + 1] = sum;
    }
    if (sum > 0.f) {
      const float inv = 1.f / sum;
      for (uint32_t x = 0; x <= w; ++x) c[x] *= inv;
      c[w] = 1.f;
    } else {
      for (uint32_t x = 0; x <= w; ++x) c[x] = float(x) / float(w);
    }
    roww[y] = sum * st;
  }
  float total = 0.f;
  for (uint32_t y = 0; y < h; ++y) {
    total += roww[y];
    marg[y + 1] = total;
  }
  if (total > 0.f) {
    const float inv = 1.f / total;
    for (uint32_t y = 0; y <= h; ++y) marg[y] *= inv;
    marg[h] = 1.f;
  } else {
    for (uint32_t y = 0; y <= h; ++y) marg[y] = float(y) / float(h);
  }
}

// The CDF search of env_sample (largest i in [0, n) with cdf[i] <= u) and its
// guide: the answer at u = b / kEnvGuide for b = 0..kEnvGuide.
uint32_t cdf_search_host(const float *cdf, uint32_t n, float u) {
  uint32_t lo = 0, hi = n;
  while (lo + 1 < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (cdf[mid] <= u) lo = mid;
    else hi = mid;
  }
  return lo;
}
void cdf_guide(const float *cdf, uint32_t n, uint32_t *guide) {
  for (uint32_t b = 0; b <= kEnvGuide; ++b) guide[b] = cdf_search_host(cdf, n, float(b) / float(kEnvGuide));
}

bool ray_args_ok(const float *a, const float *b, const float *c, const float *d) { return a && b && c && d; }

}  // namespace

namespace phx_detail {

// pt::Material{} (material.h:35-82).
phx_material default_material() {
  phx_material m{};
  m.base_color[0] = m.base_color[1] = m.base_color[2] = 0.8f;
  m.roughness = 0.5f;
  m.ior = 1.5f;
  m.specular = 0.5f;
  m.clearcoat_roughness = 0.03f;
  m.sheen_tint = 0.5f;
  m.subsurface_color[0] = m.subsurface_color[1] = m.subsurface_color[2] = 1.f;
  m.subsurface_radius = 1.f;
  m.alpha = 1.f;
  m.base_color_tex = m.normal_tex = m.roughness_tex = m.metallic_tex = m.emission_tex = m.alpha_tex = -1;
  m.alpha_mode = PHX_ALPHA_BLEND;
  m.alpha_cutoff = 0.5f;
  return m;
}

// convert_source / convert_material (RenderBridge.cpp:314-370).
phx_material convert_bridge_material(const phx_bridge_material &in) {
  auto source = [](phx_param_source p) {
    phx_param_source o{};
    o.kind = p.kind == 1 ? PHX_SOURCE_ATTRIBUTE : (p.kind == 2 ? PHX_SOURCE_SAMPLER : PHX_SOURCE_CONSTANT);
    o.index = p.index;
    return o;
  };
  phx_material m = default_material();
  for (int k = 0; k < 3; ++k) {
    m.base_color[k] = in.base_color[k];
    m.emission[k] = in.emission[k];
  }
  m.metallic = in.metallic;
  m.roughness = in.roughness * in.roughness;  // the BSDF takes the GGX alpha
  m.ior = in.ior;
  m.transmission = in.transmission;
  m.specular = 0.5f * in.specular;  // 0.5 is the full dielectric reflectance
  m.clearcoat = in.clearcoat;
  m.clearcoat_roughness = in.clearcoat_roughness;
  m.sheen = in.sheen;
  m.emission_strength = 1.f;
  m.alpha = in.opacity;
  m.base_color_src = source(in.base_color_src);
  m.alpha_src = source(in.opacity_src);
  m.metallic_src = source(in.metallic_src);
  m.roughness_src = source(in.roughness_src);
  m.emission_src = source(in.emission_src);
  m.transmission_src = source(in.transmission_src);
  m.normal_src = source(in.normal_src);
  m.roughness_src_squared = 1;
  m.base_color_alpha = in.base_color_alpha ? 1 : 0;
  if (m.alpha_src.kind != PHX_SOURCE_CONSTANT) m.alpha = 1.f;  // a sourced opacity scales the constant
  m.alpha_mode = in.alpha_mode == 0 ? PHX_ALPHA_BLEND : (in.alpha_mode == 2 ? PHX_ALPHA_MASK : PHX_ALPHA_OPAQUE);
  m.alpha_cutoff = in.alpha_cutoff;
  m.thin = in.thin_walled ? 1 : 0;
  m.flags = PHX_MATERIAL_FLAG_REFRACTIVE;
  if (in.emitter_only) m.flags |= PHX_MATERIAL_FLAG_EMITTER_ONLY;
  if (in.hidden_to_camera) m.flags |= PHX_MATERIAL_FLAG_HIDDEN_TO_CAMERA;
  if (in.emission_side == 2)
    m.flags |= PHX_MATERIAL_FLAG_EMIT_BOTH_SIDES;
  else if (in.emission_side == 1)
    m.flags |= PHX_MATERIAL_FLAG_EMIT_BACK_ONLY;
  return m;
}

}  // namespace phx_detail

extern "C" {

const char *phx_backend_name(void) { return "hip"; }
int phx_abi_version(void) { return PHX_ABI_VERSION; }

int phx_create(phx_context **out, int device) {
  if (!out) return PHX_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return PHX_ERR_HIP;
  if (device < 0 || device >= n) return PHX_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return PHX_ERR_HIP;
  // the no-allocation entry point's static scratch on this device (once per device)
  if (phx_detail::noalloc_setup(device) != PHX_OK) return PHX_ERR_HIP;
  auto *c = new phx_context();
  c->device = device;
  if (hipStreamCreate(&c->stream) != hipSuccess || hipStreamCreate(&c->stream2) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return PHX_ERR_HIP;
  }
  c->lanes[0].st = c->stream;
  c->lanes[0].st2 = c->stream2;
  c->lanes[0].ev_fork = c->ev_fork;
  c->lanes[0].ev_join = c->ev_join;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
    c->cu_count = cus;
  c->tuning = phx_tuning{};
  c->tuning.use_graph = 1;
  *out = c;
  return PHX_OK;
}

void phx_destroy(phx_context *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  (void)hipStreamSynchronize(c->stream2);
  // DBuf members free their memory in the destructor.
  hipStream_t s = c->stream, s2 = c->stream2;
  for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
  (void)hipEventDestroy(c->ev_fork);
  (void)hipEventDestroy(c->ev_join);
  std::vector<hipStream_t> lane_streams;
  for (int l = 0; l < 2; ++l) {
    phx_context::Lane &ln = c->lanes[l];
    if (l > 0) {  // lane 0's streams and fork events are the context's own
      for (hipStream_t x : {ln.st, ln.st2})
        if (x) {
          (void)hipStreamSynchronize(x);
          lane_streams.push_back(x);
        }
      for (hipEvent_t e : {ln.ev_fork, ln.ev_join})
        if (e) (void)hipEventDestroy(e);
    }
    for (hipEvent_t e : {ln.ev_read, ln.ev_reduced})
      if (e) (void)hipEventDestroy(e);
    if (ln.graph) (void)hipGraphExecDestroy(ln.graph);
    if (ln.h_items) (void)hipHostFree(ln.h_items);
    if (ln.h_first) (void)hipHostFree(ln.h_first);
  }
  if (c->pinned) (void)hipHostFree(c->pinned);
  delete c;
  if (s) (void)hipStreamDestroy(s);
  if (s2) (void)hipStreamDestroy(s2);
  for (hipStream_t x : lane_streams) (void)hipStreamDestroy(x);
}

const char *phx_last_error(const phx_context *c) { return c ? c->error.c_str() : "no context"; }

int phx_set_tuning(phx_context *c, const phx_tuning *t) {
  if (!c || !t) return PHX_ERR_INVALID;
  if (t->trace_block != 0 && t->trace_block != 128 && t->trace_block != 256)
    return fail(c, PHX_ERR_UNSUPPORTED, "set_tuning: trace_block must be 128 or 256");
  if (t->trace_groups_per_cu > 32) return fail(c, PHX_ERR_INVALID, "set_tuning: at most 32 work-groups per CU");
  if (t->refill_lanes > 64) return fail(c, PHX_ERR_INVALID, "set_tuning: refill_lanes must be at most 64");
  if (t->queue_chunk != 0 && (t->queue_chunk % 64 != 0 || t->queue_chunk > 4096))
    return fail(c, PHX_ERR_INVALID, "set_tuning: queue_chunk must be a multiple of 64 up to 4096");
  if (t->paths_in_flight > (1u << 28)) return fail(c, PHX_ERR_INVALID, "set_tuning: more than 2^28 path slots");
  if (t->sort_by_material > kMaxGroups) return fail(c, PHX_ERR_INVALID, "set_tuning: at most 15 material groups");
  c->tuning = *t;
  return PHX_OK;
}
int phx_get_tuning(const phx_context *c, phx_tuning *t) {
  if (!c || !t) return PHX_ERR_INVALID;
  *t = c->tuning;
  return PHX_OK;
}
int phx_get_stats(const phx_context *c, phx_stats *s) {
  if (!c || !s) return PHX_ERR_INVALID;
  *s = c->stats;
  return PHX_OK;
}
int phx_synchronize(phx_context *c) {
  if (!c) return PHX_ERR_INVALID;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PHX_OK;
}

int phx_build_accel(phx_context *c, const phx_scene_desc *d) {
  if (!c || !d) return PHX_ERR_INVALID;
  if (!d->positions || !d->indices || d->vertex_count == 0 || d->triangle_count == 0)
    return fail(c, PHX_ERR_INVALID, "build_accel: empty mesh");
  if (d->triangle_count >= (1u << 28)) return fail(c, PHX_ERR_INVALID, "build_accel: more than 2^28 triangles");
  HIP_TRY(c, hipSetDevice(c->device));
  c->have_accel = false;
  c->have_scene = false;
  c->scene_bytes = 0;
  Bvh bvh;
  std::string err;
  BvhOptions opt;
  if (build_bvh(d->positions, d->vertex_count, d->indices, d->triangle_count, opt, bvh, err) != 0)
    return fail(c, PHX_ERR_INVALID, err);
  // Stack bound: a visit pushes at most kNodeWidth - 1 entries per level.
  if ((kNodeWidth - 1) * bvh.depth + 1 > uint32_t(kMaxStack))
    return fail(c, PHX_ERR_OVERFLOW, "build_accel: BVH too deep for the traversal stack");
  c->bvh_depth = bvh.depth;
  int r;
  if ((r = upload(c, c->nodes, bvh.nodes.data(), bvh.nodes.size()))) return r;
  // One zero record past the last triangle: the traversal fetches a node's
  // 80 bytes at a triangle record too (its 48 bytes and the 32 after them).
  bvh.tris.emplace_back();
  std::memset(&bvh.tris.back(), 0, sizeof(TriRec));
  if ((r = upload(c, c->tris, bvh.tris.data(), bvh.tris.size()))) return r;
  bvh.tris.pop_back();
  HIP_TRY(c, c->tri_ng.reserve(bvh.tris.size() * sizeof(float4)));
  HIP_TRY(c, launch_tri_normals(c->tris.as<TriRec>(), c->tri_ng.as<float4>(), uint32_t(bvh.tris.size()), c->stream));
  c->scene_bytes += bvh.tris.size() * sizeof(float4);
  if ((r = upload(c, c->indices, d->indices, size_t(d->triangle_count) * 3))) return r;
  if ((r = upload(c, c->positions, d->positions, size_t(d->vertex_count) * 3))) return r;
  if ((r = upload(c, c->normals, d->normals, d->normals ? size_t(d->vertex_count) * 3 : 0))) return r;
  if ((r = upload(c, c->texcoords, d->texcoords, d->texcoords ? size_t(d->vertex_count) * 2 : 0))) return r;
  if ((r = upload(c, c->colors, d->vertex_colors, d->vertex_colors ? size_t(d->vertex_count) * 3 : 0))) return r;
  std::vector<uint32_t> mids;
  if (d->material_ids) {
    mids.assign(d->material_ids, d->material_ids + d->triangle_count);
    const uint32_t mc = std::max<uint32_t>(d->material_count, 1);
    for (auto &m : mids)
      if (m >= mc) m = 0;  // out-of-range ids fall back to material 0 (RenderBridge.cpp:479-484)
    if ((r = upload(c, c->material_ids, mids.data(), mids.size()))) return r;
  } else {
    c->material_ids.release();
  }
  {
    std::vector<uint4> attr(bvh.tris.size());
    for (size_t k = 0; k < bvh.tris.size(); ++k) {
      uint32_t prim;
      std::memcpy(&prim, &bvh.tris[k].v0.w, sizeof prim);
      const uint32_t *ix = d->indices + size_t(prim) * 3;
      attr[k] = make_uint4(ix[0], ix[1], ix[2], mids.empty() ? 0u : mids[prim]);
    }
    if ((r = upload(c, c->tri_attr, attr.data(), attr.size()))) return r;
  }
  c->triangle_count = d->triangle_count;
  c->vertex_count = d->vertex_count;
  c->host_positions.assign(d->positions, d->positions + size_t(d->vertex_count) * 3);
  c->host_indices.assign(d->indices, d->indices + size_t(d->triangle_count) * 3);
  c->host_material_ids = std::move(mids);
  c->px_built = false;
  // Bounds of the triangles' vertices (the BVH root box).
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = 0; i < d->triangle_count * 3; ++i) {
    const float *p = d->positions + size_t(d->indices[i]) * 3;
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], p[a]);
      hi[a] = std::max(hi[a], p[a]);
    }
  }
  std::memcpy(c->bounds_lo, lo, sizeof lo);
  std::memcpy(c->bounds_hi, hi, sizeof hi);
  DevScene &s = c->dev;
  s = DevScene{};
  s.nodes = c->nodes.as<BvhNode>();
  s.tris = c->tris.as<TriRec>();
  s.tri_attr = c->tri_attr.as<uint4>();
  s.tri_ng = c->tri_ng.as<float4>();
  s.indices = c->indices.as<uint32_t>();
  s.positions = c->positions.as<float>();
  s.normals = c->normals.as<float>();
  s.texcoords = c->texcoords.as<float>();
  s.colors = c->colors.as<float>();
  s.material_ids = c->material_ids.as<uint32_t>();
  s.triangle_count = d->triangle_count;
  c->stats.bvh_nodes = bvh.nodes.size();
  c->stats.bvh_bytes = bvh.nodes.size() * sizeof(BvhNode) + bvh.tris.size() * (sizeof(TriRec) + 2 * sizeof(uint4));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->have_accel = true;
  return PHX_OK;
}

int phx_set_scene(phx_context *c, const phx_scene_desc *d) {
  if (!c || !d) return PHX_ERR_INVALID;
  int r = phx_build_accel(c, d);
  if (r) return r;
  if (d->material_count == 0 || !d->materials) return fail(c, PHX_ERR_INVALID, "set_scene: no materials");
  const uint32_t nt = d->triangle_count;
  std::vector<phx_material> mats(d->materials, d->materials + d->material_count);
  auto mat_of = [&](uint32_t prim) -> uint32_t {
    const uint32_t m = d->material_ids ? d->material_ids[prim] : 0u;
    return m < mats.size() ? m : 0u;
  };

  // Emissive triangles (RenderBridge.cpp:780-813) unless given.
  std::vector<uint32_t> em_ids;
  std::vector<float> em_area;
  if (d->emissive_prim_ids && d->emissive_prim_areas) {
    em_ids.assign(d->emissive_prim_ids, d->emissive_prim_ids + d->emissive_count);
    em_area.assign(d->emissive_prim_areas, d->emissive_prim_areas + d->emissive_count);
    for (uint32_t id : em_ids)
      if (id >= nt) return fail(c, PHX_ERR_INVALID, "set_scene: emissive primitive id out of range");
  } else {
    for (uint32_t t = 0; t < nt; ++t) {
      if (!may_emit(mats[mat_of(t)])) continue;
      const float *p0 = d->positions + size_t(d->indices[3 * t]) * 3;
      const float *p1 = d->positions + size_t(d->indices[3 * t + 1]) * 3;
      const float *p2 = d->positions + size_t(d->indices[3 * t + 2]) * 3;
      const float e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
      const float e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
      const float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
      const float area = 0.5f * length3(cr);
      if (!(area > 0.f)) continue;
      em_ids.push_back(t);
      em_area.push_back(area);
    }
  }
  const bool has_em = !em_ids.empty();

  // Light tables (pathtracer.cpp:411-503).
  std::vector<phx_light> nee;
  for (uint32_t i = 0; i < d->light_count; ++i) {
    const phx_light &l = d->lights[i];
    if (has_em && l.type == PHX_LIGHT_AREA && l.mesh_prim_count > 0) continue;
    nee.push_back(l);
  }
  std::vector<float> cdf, prim_pdf;
  uint32_t em_count = 0;
  if (has_em) {
    const uint32_t n = uint32_t(em_ids.size());
    std::vector<double> wts(n, 0.0);
    double total = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
      const phx_material &m = mats[mat_of(em_ids[i])];
      const float e[3] = {m.emission[0] * m.emission_strength, m.emission[1] * m.emission_strength,
                          m.emission[2] * m.emission_strength};
      const float power = 0.2126f * e[0] + 0.7152f * e[1] + 0.0722f * e[2];
      wts[i] = double(em_area[i]) * double(power > 1e-6f ? power : 1e-6f);
      if (!(em_area[i] > 0.f)) wts[i] = 0.0;
      total += wts[i];
    }
    if (total > 0.0) {
      em_count = n;
      cdf.assign(n, 0.f);
      prim_pdf.assign(nt, 0.f);
      double run = 0.0;
      for (uint32_t i = 0; i < n; ++i) {
        run += wts[i];
        cdf[i] = float(run / total);
        if (wts[i] > 0.0) prim_pdf[em_ids[i]] = float(wts[i] / total / double(em_area[i]));
      }
      cdf[n - 1] = 1.f;
    }
  }
  // Search bounds of the emissive CDF per xi bucket: the answer of the
  // search at each bucket edge b / kEmGuide (exact in float).
  std::vector<uint32_t> em_guide;
  if (em_count) {
    em_guide.resize(phx::kEmGuide + 1);
    for (uint32_t b = 0; b <= phx::kEmGuide; ++b) {
      const float u = float(b) / float(phx::kEmGuide);
      uint32_t lo = 0, hi = em_count - 1;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cdf[mid] <= u) lo = mid + 1;
        else hi = mid;
      }
      em_guide[b] = lo;
    }
  }

  if ((r = upload(c, c->materials, mats.data(), mats.size()))) return r;
  if ((r = upload(c, c->lights, nee.data(), nee.size()))) return r;
  if ((r = upload(c, c->em_ids, em_ids.data(), em_count ? em_ids.size() : 0))) return r;
  if ((r = upload(c, c->em_cdf, cdf.data(), cdf.size()))) return r;
  if ((r = upload(c, c->em_guide, em_guide.data(), em_guide.size()))) return r;
  HIP_TRY(c, c->em_recs.reserve(size_t(em_count) * 4 * sizeof(float4)));
  if ((r = upload(c, c->prim_light_pdf, prim_pdf.data(), prim_pdf.size()))) return r;

  DevScene &s = c->dev;
  s.materials = c->materials.as<phx_material>();
  s.material_count = uint32_t(mats.size());
  s.lights = c->lights.as<phx_light>();
  s.light_count = uint32_t(nee.size());
  s.em_ids = c->em_ids.as<uint32_t>();
  s.em_cdf = c->em_cdf.as<float>();
  s.em_guide = c->em_guide.as<uint32_t>();
  s.em_recs = c->em_recs.as<float4>();
  s.prim_light_pdf = c->prim_light_pdf.as<float>();
  s.em_count = em_count;

  // Environment map (RenderBridge.cpp:594-620): non-finite and negative
  // texels become 0, then the CDFs.
  s.has_env = 0;
  s.env_visible = d->env_visible ? 1 : 0;
  std::vector<float> envpx, marg, cond;
  if (d->env_width > 0 && d->env_height > 0 && d->env_pixels) {
    const size_t n = size_t(d->env_width) * d->env_height * 3;
    envpx.assign(d->env_pixels, d->env_pixels + n);
    for (float &v : envpx) v = (std::isfinite(v) && v > 0.f) ? v : 0.f;
    env_cdf(envpx.data(), d->env_width, d->env_height, marg, cond);
    if ((r = upload(c, c->env_pixels, envpx.data(), envpx.size()))) return r;
    if ((r = upload(c, c->env_marginal, marg.data(), marg.size()))) return r;
    if ((r = upload(c, c->env_cond, cond.data(), cond.size()))) return r;
    std::vector<uint32_t> gm(kEnvGuide + 1), gc(size_t(d->env_height) * (kEnvGuide + 1));
    cdf_guide(marg.data(), d->env_height, gm.data());
    for (uint32_t y = 0; y < d->env_height; ++y)
      cdf_guide(&cond[size_t(y) * (d->env_width + 1)], d->env_width, &gc[size_t(y) * (kEnvGuide + 1)]);
    if ((r = upload(c, c->env_guide_m, gm.data(), gm.size()))) return r;
    if ((r = upload(c, c->env_guide_c, gc.data(), gc.size()))) return r;
    s.has_env = 1;
    s.env_w = d->env_width;
    s.env_h = d->env_height;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        s.w2t[i * 3 + j] = d->env_world_to_tex[i * 3 + j];
        s.t2w[j * 3 + i] = d->env_world_to_tex[i * 3 + j];
      }
  }
  s.env_pixels = c->env_pixels.as<float>();
  s.env_marginal = c->env_marginal.as<float>();
  s.env_cond = c->env_cond.as<float>();
  s.env_guide_m = c->env_guide_m.as<uint32_t>();
  s.env_guide_c = c->env_guide_c.as<uint32_t>();

  // Texture atlas.
  s.tex_count = 0;
  if (d->texture_count && d->texture_infos && d->texture_pixels) {
    for (uint32_t i = 0; i < d->texture_count; ++i) {
      const phx_texture_info &ti = d->texture_infos[i];
      if (size_t(ti.offset) + size_t(ti.width) * ti.height > d->texture_pixel_count)
        return fail(c, PHX_ERR_INVALID, "set_scene: texture outside the atlas");
    }
    if ((r = upload(c, c->tex_pixels, d->texture_pixels, size_t(d->texture_pixel_count) * 3))) return r;
    if ((r = upload(c, c->tex_infos, d->texture_infos, d->texture_count))) return r;
    s.tex_count = d->texture_count;
  }
  c->tex_pixel_count = s.tex_count ? d->texture_pixel_count : 0;
  s.tex_pixels = c->tex_pixels.as<float>();
  s.tex_infos = c->tex_infos.as<phx_texture_info>();

  // Vertex attributes and samplers of the parameter sources.
  for (int k = 0; k < PHX_ATTRIBUTE_MESH_COUNT; ++k) {
    const int32_t off = d->attributes ? d->attribute_offset[k] : -1;
    if (off >= 0 && (!d->attributes || uint64_t(off) + c->vertex_count > d->attribute_count))
      return fail(c, PHX_ERR_INVALID, "set_scene: attribute slot outside the attribute records");
    s.attr_offset[k] = off >= 0 ? off : -1;
  }
  if ((r = upload(c, c->attributes, reinterpret_cast<const float4 *>(d->attributes),
                  d->attributes ? d->attribute_count : 0)))
    return r;
  s.attributes = c->attributes.as<float4>();
  s.sampler_count = 0;
  if (d->sampler_count && d->samplers) {
    for (uint32_t i = 0; i < d->sampler_count; ++i) {
      const phx_sampler_info &si = d->samplers[i];
      if (si.type > PHX_SAMPLER_TRANSFORM) return fail(c, PHX_ERR_INVALID, "set_scene: unknown sampler type");
      if (si.type == PHX_SAMPLER_TRANSFORM) continue;
      uint64_t n = 1;
      for (int k = 0; k < 3; ++k) {
        if (si.size[k] == 0 || si.wrap[k] > PHX_WRAP_MIRROR)
          return fail(c, PHX_ERR_INVALID, "set_scene: sampler size 0 or unknown wrap mode");
        n *= si.size[k];
      }
      if (!d->sampler_texels || uint64_t(si.offset) + n > d->sampler_texel_count)
        return fail(c, PHX_ERR_INVALID, "set_scene: sampler image outside the sampler texels");
    }
    if ((r = upload(c, c->samplers, d->samplers, d->sampler_count))) return r;
    if ((r = upload(c, c->sampler_texels, reinterpret_cast<const float4 *>(d->sampler_texels),
                    d->sampler_texels ? d->sampler_texel_count : 0)))
      return r;
    s.sampler_count = d->sampler_count;
  } else {
    c->samplers.release();
    c->sampler_texels.release();
  }
  s.samplers = c->samplers.as<phx_sampler_info>();
  s.sampler_texels = c->sampler_texels.as<float4>();
  s.has_sources = 0;
  for (const phx_material &m : mats)
    if (has_source(m)) s.has_sources = 1;

  // Volumes.
  s.vol_count = 0;
  s.vol_tf_count = 0;
  if (d->volume_count && d->volumes) {
    for (uint32_t i = 0; i < d->volume_count; ++i) {
      const phx_volume &v = d->volumes[i];
      // tracking skips a volume without a majorant or table and reads none of its data
      if (!(v.majorant > 0.f) || v.tf_size == 0) continue;
      for (int k = 0; k < 3; ++k)
        if (v.dims[k] == 0 || v.macro_dims[k] == 0)
          return fail(c, PHX_ERR_INVALID, "set_scene: volume with an empty grid or macro-cell grid");
      const size_t cells = size_t(v.dims[0]) * v.dims[1] * v.dims[2];
      if (v.data_offset + cells > d->volume_value_count || v.tf_offset + v.tf_size > d->volume_tf_count)
        return fail(c, PHX_ERR_INVALID, "set_scene: volume data out of range");
      const size_t mcells = size_t(v.macro_dims[0]) * v.macro_dims[1] * v.macro_dims[2];
      if (v.macro_offset + mcells > d->volume_majorant_count)
        return fail(c, PHX_ERR_INVALID, "set_scene: volume majorants out of range");
    }
    if ((r = upload(c, c->volumes, d->volumes, d->volume_count))) return r;
    if ((r = upload(c, c->vol_values, d->volume_values, d->volume_value_count))) return r;
    if ((r = upload(c, c->vol_tf, reinterpret_cast<const float4 *>(d->volume_tf), d->volume_tf_count))) return r;
    if ((r = upload(c, c->majorants, d->volume_majorants, d->volume_majorant_count))) return r;
    s.vol_count = d->volume_count;
    s.vol_tf_count = d->volume_tf_count;
  }
  s.volumes = c->volumes.as<phx_volume>();
  s.vol_values = c->vol_values.as<float>();
  s.vol_tf = c->vol_tf.as<float4>();
  s.majorants = c->majorants.as<float>();

  s.strategy_count = s.light_count + (em_count > 0 ? 1u : 0u) + (s.has_env ? 1u : 0u);
  // Ray offset (pathtracer.cpp:489-502).
  float eps = 1e-3f;
  if (d->ray_epsilon > 0.f) {
    eps = d->ray_epsilon;
  } else if (c->bounds_hi[0] >= c->bounds_lo[0]) {
    const float dg[3] = {c->bounds_hi[0] - c->bounds_lo[0], c->bounds_hi[1] - c->bounds_lo[1],
                         c->bounds_hi[2] - c->bounds_lo[2]};
    const float e = 1e-4f * length3(dg);
    eps = e < 1e-6f ? 1e-6f : (e > 1e-2f ? 1e-2f : e);
  }
  s.ray_eps = eps;
  HIP_TRY(c, launch_em_records(s, c->em_recs.as<float4>(), c->stream));
  c->stats.scene_bytes = c->scene_bytes;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->have_scene = true;
  return PHX_OK;
}

int phx_trace_closest_compact(phx_context *c, const float *o, const float *d, const float *tmin, const float *tmax,
                              uint32_t n, phx_hit_compact *out, void *stream) {
  if (!c || (n && (!ray_args_ok(o, d, tmin, tmax) || !out))) return PHX_ERR_INVALID;
  if (!c->have_accel) return fail(c, PHX_ERR_NO_SCENE, "trace before build_accel");
  if (n == 0) return PHX_OK;
  int r = ensure_trace_scratch(c);
  if (r) return r;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  HIP_TRY(c, launch_clear_u32(c->fetch.as<uint32_t>(), kFetchSetWords, st));
  HIP_TRY(c, launch_trace_compact(c->dev, o, d, tmin, tmax, n, out, c->fetch.as<uint32_t>(), c->spill.as<int>(),
                                  c->trace_grid, st));
  return PHX_OK;
}
int phx_trace_closest(phx_context *c, const float *o, const float *d, const float *tmin, const float *tmax,
                      uint32_t n, phx_hit *out, void *stream) {
  if (!c || (n && (!ray_args_ok(o, d, tmin, tmax) || !out))) return PHX_ERR_INVALID;
  if (!c->have_accel) return fail(c, PHX_ERR_NO_SCENE, "trace before build_accel");
  if (n == 0) return PHX_OK;
  int r = ensure_trace_scratch(c);
  if (r) return r;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  HIP_TRY(c, launch_clear_u32(c->fetch.as<uint32_t>(), kFetchSetWords, st));
  HIP_TRY(c, launch_trace_full(c->dev, o, d, tmin, tmax, n, out, c->fetch.as<uint32_t>(), c->spill.as<int>(),
                               c->trace_grid, st));
  return PHX_OK;
}
int phx_trace_occluded(phx_context *c, const float *o, const float *d, const float *tmin, const float *tmax,
                       uint32_t n, uint32_t *out, void *stream) {
  if (!c || (n && (!ray_args_ok(o, d, tmin, tmax) || !out))) return PHX_ERR_INVALID;
  if (!c->have_accel) return fail(c, PHX_ERR_NO_SCENE, "trace before build_accel");
  if (n == 0) return PHX_OK;
  int r = ensure_trace_scratch(c);
  if (r) return r;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  HIP_TRY(c, launch_clear_u32(c->fetch.as<uint32_t>(), kFetchSetWords, st));
  HIP_TRY(c, launch_trace_occluded(c->dev, o, d, tmin, tmax, n, out, c->fetch.as<uint32_t>(), c->spill.as<int>(),
                                   c->trace_grid, st));
  return PHX_OK;
}

int phx_convert_material(const phx_bridge_material *in, phx_material *out) {
  if (!in || !out) return PHX_ERR_INVALID;
  *out = convert_bridge_material(*in);
  return PHX_OK;
}

int phx_camera_perspective(phx_camera *out, const float f[3], const float at[3], const float up[3], float vfov,
                           float aspect, float aperture, float focus) {
  if (!out || !f || !at || !up) return PHX_ERR_INVALID;
  // Camera::make_perspective (camera.h:32-55).
  const float theta = vfov * 3.1415926535f / 180.f;
  const float h = std::tan(theta * 0.5f);
  const float vh = 2.f * h, vw = aspect * vh;
  float w[3] = {f[0] - at[0], f[1] - at[1], f[2] - at[2]};
  auto norm = [](float *v) {
    const float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const float inv = l > 0.f ? 1.f / l : 0.f;
    if (l > 0.f) {
      v[0] *= inv;
      v[1] *= inv;
      v[2] *= inv;
    } else {
      v[0] = v[1] = v[2] = 0.f;
    }
  };
  auto crs = [](const float *a, const float *b, float *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  norm(w);
  float u[3], v[3];
  crs(up, w, u);
  norm(u);
  crs(w, u, v);
  phx_camera c{};
  for (int i = 0; i < 3; ++i) {
    c.origin[i] = f[i];
    c.u[i] = u[i];
    c.v[i] = v[i];
    c.w[i] = w[i];
    c.horizontal[i] = u[i] * (vw * focus);
    c.vertical[i] = v[i] * (vh * focus);
  }
  for (int i = 0; i < 3; ++i)
    c.lower_left[i] = ((f[i] - c.horizontal[i] * 0.5f) - c.vertical[i] * 0.5f) - w[i] * focus;
  c.lens_radius = aperture * 0.5f;
  c.focus_dist = focus;
  c.is_ortho = 0;
  c.clip_near = 0.f;
  c.clip_far = 1e30f;
  *out = c;
  return PHX_OK;
}

int phx_camera_orthographic(phx_camera *out, const float f[3], const float at[3], const float up[3], float height,
                            float aspect) {
  if (!out || !f || !at || !up) return PHX_ERR_INVALID;
  // Camera::make_orthographic (camera.h:57-77).
  phx_camera c{};
  int r = phx_camera_perspective(&c, f, at, up, 90.f, aspect, 0.f, 1.f);
  if (r) return r;
  const float vh = height, vw = height * aspect;
  for (int i = 0; i < 3; ++i) {
    c.horizontal[i] = c.u[i] * vw;
    c.vertical[i] = c.v[i] * vh;
    c.lower_left[i] = (f[i] - c.horizontal[i] * 0.5f) - c.vertical[i] * 0.5f;
  }
  c.focus_dist = 1.f;
  c.lens_radius = 0.f;
  c.is_ortho = 1;
  *out = c;
  return PHX_OK;
}

int phx_result_device_color(phx_context *c, const float **d_color) {
  if (!c || !d_color) return PHX_ERR_INVALID;
  if (!c->have_result) return fail(c, PHX_ERR_NO_FRAME, "no frame has been rendered");
  *d_color = c->f_color.as<float>();
  return PHX_OK;
}

int phx_read_result(phx_context *c, float *color, float *alpha, float *depth, float *normal, float *albedo) {
  if (!c) return PHX_ERR_INVALID;
  if (!c->have_result) return fail(c, PHX_ERR_NO_FRAME, "no frame has been rendered");
  const size_t n = size_t(c->f_w) * c->f_h;
  hipStream_t st = c->stream;
  if (color) HIP_TRY(c, hipMemcpyAsync(color, c->f_color.p, n * 12, hipMemcpyDefault, st));
  if (alpha) HIP_TRY(c, hipMemcpyAsync(alpha, c->f_alpha.p, n * 4, hipMemcpyDefault, st));
  if (depth) HIP_TRY(c, hipMemcpyAsync(depth, c->f_depth.p, n * 4, hipMemcpyDefault, st));
  if (normal) HIP_TRY(c, hipMemcpyAsync(normal, c->f_normal.p, n * 12, hipMemcpyDefault, st));
  if (albedo) HIP_TRY(c, hipMemcpyAsync(albedo, c->f_albedo.p, n * 12, hipMemcpyDefault, st));
  HIP_TRY(c, hipStreamSynchronize(st));
  return PHX_OK;
}

int phx_read_sample_counts(phx_context *c, uint32_t *counts) {
  if (!c || !counts) return PHX_ERR_INVALID;
  if (!c->have_result) return fail(c, PHX_ERR_NO_FRAME, "no frame has been rendered");
  if (!c->f_counted)
    return fail(c, PHX_ERR_NO_FRAME, "the frame was not adaptively sampled: every pixel took samples_per_pixel");
  const size_t n = size_t(c->f_w) * c->f_h;
  HIP_TRY(c, hipMemcpyAsync(counts, c->f_nsamp.p, n * 4, hipMemcpyDefault, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PHX_OK;
}

}  // extern "C"
