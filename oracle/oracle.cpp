// ORACLE - test infrastructure only (see orc_core.h).  The reference's BVH
// build, traversal, light tables and the PathTracer::render integrator
// restated on the CPU, plus the extern "C" entry points the tests call.
#include <omp.h>

#include <cstdio>
#include <memory>
#include <numeric>

#include "orc_scene.h"

namespace orc {

// ---- BVH build: bvh.cpp:13-125 ----------------------------------------------
namespace {
struct PrimRef {
  uint32_t id;
  Box b;
  V3<float> c;
};
uint32_t build_node(std::vector<Node> &nodes, std::vector<PrimRef> &prims, std::vector<uint32_t> &ids, uint32_t begin,
                    uint32_t end) {
  const uint32_t node_id = uint32_t(nodes.size());
  nodes.push_back({});
  Box b, cb;
  for (uint32_t i = begin; i < end; ++i) {
    extend(b, prims[i].b.lo);
    extend(b, prims[i].b.hi);
    extend(cb, prims[i].c);
  }
  Node &n = nodes[node_id];
  n.b = b;
  const uint32_t pb = uint32_t(ids.size());
  for (uint32_t i = begin; i < end; ++i) ids.push_back(prims[i].id);
  n.begin = pb;
  n.count = end - begin;
  if (n.count <= 4) {  // bvh.cpp:42-48
    n.leaf = 1;
    return node_id;
  }
  ids.resize(pb);
  n.begin = begin;
  const V3<float> dg = cb.hi - cb.lo;
  int axis = 0;  // bvh.cpp:53-58
  if (dg.y > dg.x) axis = 1;
  if ((axis == 0 ? dg.z : dg.z) > (axis == 0 ? dg.x : dg.y)) axis = 2;
  auto key = [axis](const PrimRef &p) { return axis == 0 ? p.c.x : (axis == 1 ? p.c.y : p.c.z); };
  const uint32_t mid = begin + (end - begin) / 2;
  std::nth_element(prims.begin() + begin, prims.begin() + mid, prims.begin() + end,
                   [&](const PrimRef &a, const PrimRef &c) { return key(a) < key(c); });
  const uint32_t l = build_node(nodes, prims, ids, begin, mid);
  const uint32_t r = build_node(nodes, prims, ids, mid, end);
  nodes[node_id].left = l;
  nodes[node_id].right = r;
  nodes[node_id].leaf = 0;
  return node_id;
}
}  // namespace

void build_bvh(const Mesh &m, Bvh &out) {
  const uint32_t nt = m.ntri();
  std::vector<PrimRef> prims(nt);
  for (uint32_t i = 0; i < nt; ++i) {
    const V3<float> v0 = m.P(m.idx[i * 3]), v1 = m.P(m.idx[i * 3 + 1]), v2 = m.P(m.idx[i * 3 + 2]);
    prims[i].id = i;
    Box b;
    extend(b, v0);
    extend(b, v1);
    extend(b, v2);
    prims[i].b = b;
    prims[i].c = (v0 + v1 + v2) * (1.f / 3.f);
  }
  out.nodes.clear();
  out.prim_ids.clear();
  out.nodes.reserve(size_t(nt) * 2);
  out.prim_ids.reserve(nt);
  out.root = build_node(out.nodes, prims, out.prim_ids, 0, nt);
}

// mesh_intersector.h:57-124
ClosestHit trace_bvh(const Mesh &m, const Bvh &bvh, const V3<float> &o, const V3<float> &d, float tmin, float tmax,
                     bool any_hit) {
  ClosestHit best;
  best.t = tmax;
  if (!(tmax > tmin) || bvh.nodes.empty()) return best;
  const V3<float> inv = safe_inv(d);
  uint32_t stack[64];
  int sp = 0;
  stack[sp++] = bvh.root;
  while (sp > 0) {
    const Node &n = bvh.nodes[stack[--sp]];
    if (aabb_entry(n.b, o, inv, tmin, best.t) < 0.f) continue;
    if (n.leaf) {
      for (uint32_t i = 0; i < n.count; ++i) {
        const uint32_t prim = bvh.prim_ids[n.begin + i];
        const TriHit h = intersect_triangle(o, d, m.P(m.idx[prim * 3]), m.P(m.idx[prim * 3 + 1]), m.P(m.idx[prim * 3 + 2]),
                                            tmin, best.t);
        if (!h.hit) continue;
        best.t = h.t;
        best.prim = prim;
        best.hit = true;
        best.u = h.u;
        best.v = h.v;
        if (any_hit) return best;
      }
    } else if (sp + 2 <= 64) {
      const float dl = aabb_entry(bvh.nodes[n.left].b, o, inv, tmin, best.t);
      const float dr = aabb_entry(bvh.nodes[n.right].b, o, inv, tmin, best.t);
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
  return best;
}

ClosestHit trace_brute(const Mesh &m, const V3<float> &o, const V3<float> &d, float tmin, float tmax, bool any_hit) {
  ClosestHit best;
  best.t = tmax;
  if (!(tmax > tmin)) return best;
  const uint32_t nt = m.ntri();
  for (uint32_t p = 0; p < nt; ++p) {
    const TriHit h = intersect_triangle(o, d, m.P(m.idx[p * 3]), m.P(m.idx[p * 3 + 1]), m.P(m.idx[p * 3 + 2]), tmin, best.t);
    if (!h.hit) continue;
    best.t = h.t;
    best.prim = p;
    best.u = h.u;
    best.v = h.v;
    best.hit = true;
    if (any_hit) return best;
  }
  return best;
}

// ---- majorants: RenderBridge.cpp:656-778 ------------------------------------
uint32_t build_majorants(phx_volume &v, const float *values, const float *tf_rgba, uint32_t tf_size,
                         std::vector<float> &out) {
  out.clear();
  float max_op = 0.f;
  for (uint32_t t = 0; t < tf_size; ++t) max_op = std::max(max_op, tf_rgba[4 * t + 3]);
  for (int d = 0; d < 3; ++d) v.dims[d] = std::max<uint32_t>(v.dims[d], 1);
  v.majorant = max_op * v.density_scale;
  const bool complete = values != nullptr;
  if (!complete) v.majorant = 0.f;
  uint32_t sub[3] = {0, 0, 0};
  for (int d = 0; d < 3; ++d) {
    const uint32_t cells = v.dims[d] > 1 ? v.dims[d] - 1 : 1;
    if (cells <= 32) {
      sub[d] = std::min<uint32_t>(std::max<uint32_t>((32 + cells - 1) / cells, 1), 16);
      v.macro_dims[d] = cells * sub[d];
    } else {
      v.macro_dims[d] = std::min<uint32_t>(std::max<uint32_t>((v.dims[d] + 2) / 4, 1), 128);
    }
  }
  const float range = v.value_hi - v.value_lo;
  const bool refined = sub[0] > 0 && sub[1] > 0 && sub[2] > 0;
  if (!refined)
    for (int d = 0; d < 3; ++d)
      if (sub[d] > 0) v.macro_dims[d] = std::min<uint32_t>(std::max<uint32_t>((v.dims[d] + 2) / 4, 1), 128);
  auto node = [&](uint32_t x, uint32_t y, uint32_t z) {
    x = std::min(x, v.dims[0] - 1);
    y = std::min(y, v.dims[1] - 1);
    z = std::min(z, v.dims[2] - 1);
    return values[(size_t(z) * v.dims[1] + y) * v.dims[0] + x];
  };
  auto interp = [&](float gx, float gy, float gz) {
    const uint32_t x = uint32_t(gx), y = uint32_t(gy), z = uint32_t(gz);
    const float fx = gx - float(x), fy = gy - float(y), fz = gz - float(z);
    const float c00 = node(x, y, z) * (1.f - fx) + node(x + 1, y, z) * fx;
    const float c10 = node(x, y + 1, z) * (1.f - fx) + node(x + 1, y + 1, z) * fx;
    const float c01 = node(x, y, z + 1) * (1.f - fx) + node(x + 1, y, z + 1) * fx;
    const float c11 = node(x, y + 1, z + 1) * (1.f - fx) + node(x + 1, y + 1, z + 1) * fx;
    return (c00 * (1.f - fy) + c10 * fy) * (1.f - fz) + (c01 * (1.f - fy) + c11 * fy) * fz;
  };
  for (uint32_t mz = 0; mz < v.macro_dims[2]; ++mz)
    for (uint32_t my = 0; my < v.macro_dims[1]; ++my)
      for (uint32_t mx = 0; mx < v.macro_dims[0]; ++mx) {
        float bound = 0.f;
        if (complete && tf_size > 0) {
          const uint32_t mm[3] = {mx, my, mz};
          float vmin = std::numeric_limits<float>::max(), vmax = -std::numeric_limits<float>::max();
          if (refined) {
            for (int corner = 0; corner < 8; ++corner) {
              float g[3];
              for (int d = 0; d < 3; ++d) g[d] = float(mm[d] + ((corner >> d) & 1)) / float(sub[d]);
              const float val = interp(g[0], g[1], g[2]);
              vmin = std::min(vmin, val);
              vmax = std::max(vmax, val);
            }
          } else {
            uint32_t first[3], last[3];
            for (int d = 0; d < 3; ++d) {
              const float cells = float(v.dims[d] > 1 ? v.dims[d] - 1 : 1);
              const float lo = float(mm[d]) / float(v.macro_dims[d]) * cells;
              const float hi = float(mm[d] + 1) / float(v.macro_dims[d]) * cells;
              first[d] = uint32_t(std::max(std::floor(lo) - 1.f, 0.f));
              last[d] = std::min<uint32_t>(uint32_t(std::ceil(hi)) + 1, v.dims[d] - 1);
            }
            for (uint32_t z = first[2]; z <= last[2]; ++z)
              for (uint32_t y = first[1]; y <= last[1]; ++y)
                for (uint32_t x = first[0]; x <= last[0]; ++x) {
                  const float val = values[(size_t(z) * v.dims[1] + y) * v.dims[0] + x];
                  vmin = std::min(vmin, val);
                  vmax = std::max(vmax, val);
                }
          }
          auto tpos = [&](float val) {
            float x = range != 0.f ? (val - v.value_lo) / range : 0.f;
            x = std::min(std::max(x, 0.f), 1.f);
            return x * float(tf_size - 1);
          };
          const size_t i0 = size_t(std::floor(tpos(vmin)));
          const size_t i1 = std::min(size_t(std::ceil(tpos(vmax))), size_t(tf_size) - 1);
          for (size_t t = i0; t <= i1; ++t) bound = std::max(bound, tf_rgba[4 * t + 3]);
          if (!(vmin <= vmax)) bound = max_op;
        }
        out.push_back(bound * v.density_scale * 1.01f);
      }
  return uint32_t(out.size());
}

// ---- scene -----------------------------------------------------------------
struct Scene {
  Mesh mesh;
  Bvh bvh;
  std::vector<phx_material> mats;
  std::vector<phx_light> nee_lights;  // after build_light_tables
  std::vector<uint32_t> emissive_ids;
  std::vector<float> emissive_areas;
  std::vector<float> emissive_cdf;
  std::vector<float> prim_light_pdf;
  uint32_t emissive_count{0};
  EnvMap env;
  bool has_env{false}, env_visible{true};
  Atlas atlas;
  // MeshAttributes and SamplerTable (sampler.h:25-139)
  std::vector<float> attrs;  // RGBA records
  int attr_off[PHX_ATTRIBUTE_MESH_COUNT]{-1, -1, -1, -1, -1};
  std::vector<phx_sampler_info> samplers;
  std::vector<float> sampler_texels;  // RGBA
  Volumes vol;
  float ray_eps{1e-3f};
  bool brute{false};  // trace by brute force instead of the BVH
};

static void scene_load(Scene &s, const phx_scene_desc &d) {
  const uint32_t nv = d.vertex_count, nt = d.triangle_count;
  s.mesh.pos.assign(d.positions, d.positions + size_t(nv) * 3);
  s.mesh.idx.assign(d.indices, d.indices + size_t(nt) * 3);
  if (d.normals) s.mesh.nrm.assign(d.normals, d.normals + size_t(nv) * 3);
  if (d.texcoords) s.mesh.uv.assign(d.texcoords, d.texcoords + size_t(nv) * 2);
  if (d.vertex_colors) s.mesh.col.assign(d.vertex_colors, d.vertex_colors + size_t(nv) * 3);
  if (d.material_ids) s.mesh.mat.assign(d.material_ids, d.material_ids + nt);
  s.mats.assign(d.materials, d.materials + d.material_count);
  build_bvh(s.mesh, s.bvh);

  // Emissive list: given, or derived as RenderBridge.cpp:780-813 does.
  s.emissive_ids.clear();
  s.emissive_areas.clear();
  if (d.emissive_prim_ids) {
    s.emissive_ids.assign(d.emissive_prim_ids, d.emissive_prim_ids + d.emissive_count);
    s.emissive_areas.assign(d.emissive_prim_areas, d.emissive_prim_areas + d.emissive_count);
  } else {
    const uint32_t mc = std::max<uint32_t>(d.material_count, 1);
    for (uint32_t t = 0; t < nt; ++t) {
      const uint32_t mid = (!s.mesh.mat.empty() && s.mesh.mat[t] < mc) ? s.mesh.mat[t] : 0u;
      if (mid >= s.mats.size() || !material_may_emit(mat_from<float>(s.mats[mid]))) continue;
      const V3<float> p0 = s.mesh.P(s.mesh.idx[3 * t]), p1 = s.mesh.P(s.mesh.idx[3 * t + 1]), p2 = s.mesh.P(s.mesh.idx[3 * t + 2]);
      const float area = 0.5f * length(cross(p1 - p0, p2 - p0));
      if (!(area > 0.f)) continue;
      s.emissive_ids.push_back(t);
      s.emissive_areas.push_back(area);
    }
  }
  const bool has_em = !s.emissive_ids.empty();

  // build_light_tables: pathtracer.cpp:411-503
  s.nee_lights.clear();
  for (uint32_t i = 0; i < d.light_count; ++i) {
    const phx_light &l = d.lights[i];
    if (has_em && l.type == PHX_LIGHT_AREA && l.mesh_prim_count > 0) continue;
    s.nee_lights.push_back(l);
  }
  s.emissive_count = 0;
  s.emissive_cdf.clear();
  s.prim_light_pdf.clear();
  if (has_em) {
    const uint32_t n = uint32_t(s.emissive_ids.size());
    std::vector<double> w(n, 0.0);
    double total = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t prim = s.emissive_ids[i];
      const uint32_t mid = !s.mesh.mat.empty() ? s.mesh.mat[prim] : 0u;
      float power = 1.f;
      if (mid < s.mats.size()) power = luminance(material_emission(mat_from<float>(s.mats[mid])));
      w[i] = double(s.emissive_areas[i]) * double(power > 1e-6f ? power : 1e-6f);
      if (!(s.emissive_areas[i] > 0.f)) w[i] = 0.0;
      total += w[i];
    }
    if (total > 0.0) {
      s.emissive_count = n;
      s.emissive_cdf.assign(n, 0.f);
      s.prim_light_pdf.assign(nt, 0.f);
      double run = 0.0;
      for (uint32_t i = 0; i < n; ++i) {
        run += w[i];
        s.emissive_cdf[i] = float(run / total);
        if (w[i] > 0.0 && s.emissive_ids[i] < nt)
          s.prim_light_pdf[s.emissive_ids[i]] = float(w[i] / total / double(s.emissive_areas[i]));
      }
      s.emissive_cdf[n - 1] = 1.f;
    }
  }

  s.has_env = d.env_width > 0 && d.env_height > 0 && d.env_pixels;
  if (s.has_env) {
    s.env.width = d.env_width;
    s.env.height = d.env_height;
    s.env.pixels.assign(d.env_pixels, d.env_pixels + size_t(d.env_width) * d.env_height * 3);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        s.env.w2t[r * 3 + c] = d.env_world_to_tex[r * 3 + c];
        s.env.t2w[c * 3 + r] = d.env_world_to_tex[r * 3 + c];
      }
    s.env.build_cdf();
  }
  s.env_visible = d.env_visible != 0;
  if (d.texture_count) {
    s.atlas.pixels.assign(d.texture_pixels, d.texture_pixels + size_t(d.texture_pixel_count) * 3);
    s.atlas.infos.assign(d.texture_infos, d.texture_infos + d.texture_count);
  }
  for (int k = 0; k < PHX_ATTRIBUTE_MESH_COUNT; ++k) s.attr_off[k] = d.attributes ? d.attribute_offset[k] : -1;
  if (d.attributes) s.attrs.assign(d.attributes, d.attributes + size_t(d.attribute_count) * 4);
  if (d.sampler_count && d.samplers) {
    s.samplers.assign(d.samplers, d.samplers + d.sampler_count);
    if (d.sampler_texels) s.sampler_texels.assign(d.sampler_texels, d.sampler_texels + size_t(d.sampler_texel_count) * 4);
  }
  if (d.volume_count) {
    s.vol.vols.assign(d.volumes, d.volumes + d.volume_count);
    s.vol.values.assign(d.volume_values, d.volume_values + d.volume_value_count);
    s.vol.tf.assign(d.volume_tf, d.volume_tf + size_t(d.volume_tf_count) * 4);
    s.vol.majorants.assign(d.volume_majorants, d.volume_majorants + d.volume_majorant_count);
  }
  // pathtracer.cpp:489-502
  s.ray_eps = 1e-3f;
  if (d.ray_epsilon > 0.f) {
    s.ray_eps = d.ray_epsilon;
  } else if (!s.bvh.nodes.empty() && nt > 0) {
    const Box &b = s.bvh.nodes[s.bvh.root].b;
    if (b.hi.x >= b.lo.x) {
      const float eps = 1e-4f * length(b.hi - b.lo);
      s.ray_eps = eps < 1e-6f ? 1e-6f : (eps > 1e-2f ? 1e-2f : eps);
    }
  }
}

// Full hit record: mesh_intersector.h:126-178.
template <class T> struct Hit {
  bool hit{false};
  T t{0};
  V3<T> p, n, sn, color;
  V2<T> uv, bary;
  uint32_t prim{0}, mat{0};
  bool has_color{false};
};
template <class T> static Hit<T> closest(const Scene &s, const V3<T> &o, const V3<T> &d, T tmin, T tmax) {
  const V3<float> of(o), df(d);
  const ClosestHit ch = s.brute ? trace_brute(s.mesh, of, df, float(tmin), float(tmax), false)
                                : trace_bvh(s.mesh, s.bvh, of, df, float(tmin), float(tmax), false);
  Hit<T> h;
  if (!ch.hit) return h;
  h.hit = true;
  h.t = T(ch.t);
  const Mesh &m = s.mesh;
  const uint32_t i0 = m.idx[ch.prim * 3], i1 = m.idx[ch.prim * 3 + 1], i2 = m.idx[ch.prim * 3 + 2];
  const V3<T> v0 = v3<T>(&m.pos[size_t(i0) * 3]), v1 = v3<T>(&m.pos[size_t(i1) * 3]), v2 = v3<T>(&m.pos[size_t(i2) * 3]);
  h.p = o + d * h.t;
  const V3<T> nn = normalize(cross(v1 - v0, v2 - v0));
  h.n = nn;
  h.sn = nn;
  h.bary = {T(ch.u), T(ch.v)};
  const T w0 = T(1) - T(ch.u) - T(ch.v), w1 = T(ch.u), w2 = T(ch.v);
  if (!m.nrm.empty()) {
    h.sn = normalize(v3<T>(&m.nrm[size_t(i0) * 3]) * w0 + v3<T>(&m.nrm[size_t(i1) * 3]) * w1 + v3<T>(&m.nrm[size_t(i2) * 3]) * w2);
    if (dot(h.sn, nn) < T(0)) h.sn = h.sn * T(-1);
  }
  if (!m.uv.empty()) {
    const float *a = &m.uv[size_t(i0) * 2], *b = &m.uv[size_t(i1) * 2], *c = &m.uv[size_t(i2) * 2];
    h.uv = {T(a[0]) * w0 + T(b[0]) * w1 + T(c[0]) * w2, T(a[1]) * w0 + T(b[1]) * w1 + T(c[1]) * w2};
  }
  if (!m.col.empty() && m.col[size_t(i0) * 3] >= 0.f) {
    h.color = v3<T>(&m.col[size_t(i0) * 3]) * w0 + v3<T>(&m.col[size_t(i1) * 3]) * w1 + v3<T>(&m.col[size_t(i2) * 3]) * w2;
    h.has_color = true;
  }
  h.prim = ch.prim;
  h.mat = m.mat.empty() ? 0u : m.mat[ch.prim];
  return h;
}
static bool occluded(const Scene &s, const V3<float> &o, const V3<float> &d, float tmin, float tmax) {
  const ClosestHit ch = s.brute ? trace_brute(s.mesh, o, d, tmin, tmax, true) : trace_bvh(s.mesh, s.bvh, o, d, tmin, tmax, true);
  return ch.hit;
}

// A point of the mesh as the material resolve sees it (SurfaceQuery,
// pathtracer.cpp:95-108).
template <class T> struct Query {
  uint32_t prim{0}, mat{0};
  T b1{0}, b2{0};
  V3<T> p, sn, color;
  V2<T> uv;
  bool has_color{false};
};
// The query of a hit: `hsn` is the hit's shading normal after the flat /
// side fix of pathtracer.cpp:906-910.
template <class T> static Query<T> query_of(const Hit<T> &hr, const V3<T> &hsn) {
  Query<T> q;
  q.prim = hr.prim;
  q.mat = hr.mat;
  q.b1 = hr.bary.x;
  q.b2 = hr.bary.y;
  q.p = hr.p;
  q.sn = hsn;
  q.uv = hr.uv;
  q.color = hr.color;
  q.has_color = hr.has_color;
  return q;
}
// fetch_attribute (pathtracer.cpp:111-123)
template <class T> static V4<T> fetch_attribute(const Scene &s, int slot, const Query<T> &q) {
  if (slot == PHX_ATTRIBUTE_WORLD_POSITION) return {q.p.x, q.p.y, q.p.z, T(1)};
  if (slot == PHX_ATTRIBUTE_WORLD_NORMAL) return {q.sn.x, q.sn.y, q.sn.z, T(1)};
  if (slot < 0 || slot >= PHX_ATTRIBUTE_MESH_COUNT || s.attr_off[slot] < 0) return {T(0), T(0), T(0), T(1)};
  const size_t base = size_t(s.attr_off[slot]);
  auto rec = [&](uint32_t vtx) {
    const float *a = &s.attrs[(base + vtx) * 4];
    return V4<T>{T(a[0]), T(a[1]), T(a[2]), T(a[3])};
  };
  const V4<T> a0 = rec(s.mesh.idx[q.prim * 3]), a1 = rec(s.mesh.idx[q.prim * 3 + 1]), a2 = rec(s.mesh.idx[q.prim * 3 + 2]);
  const T w0 = T(1) - q.b1 - q.b2;
  return {a0.x * w0 + a1.x * q.b1 + a2.x * q.b2, a0.y * w0 + a1.y * q.b1 + a2.y * q.b2,
          a0.z * w0 + a1.z * q.b1 + a2.z * q.b2, a0.w * w0 + a1.w * q.b1 + a2.w * q.b2};
}
// SamplerTable::wrap_index (sampler.h:62-71)
static int wrap_index(int i, int n, uint32_t mode) {
  if (mode == PHX_WRAP_REPEAT) return ((i % n) + n) % n;
  if (mode == PHX_WRAP_MIRROR) {
    const int j = ((i % (2 * n)) + 2 * n) % (2 * n);
    return j < n ? j : 2 * n - 1 - j;
  }
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}
// mul_mat4 (math_vec4.h:37-43) then the offset
template <class T> static V4<T> mat4_apply(const float *m, const float *o, const V4<T> &v) {
  V4<T> r{T(m[0]) * v.x + T(m[1]) * v.y + T(m[2]) * v.z + T(m[3]) * v.w,
          T(m[4]) * v.x + T(m[5]) * v.y + T(m[6]) * v.z + T(m[7]) * v.w,
          T(m[8]) * v.x + T(m[9]) * v.y + T(m[10]) * v.z + T(m[11]) * v.w,
          T(m[12]) * v.x + T(m[13]) * v.y + T(m[14]) * v.z + T(m[15]) * v.w};
  r.x += T(o[0]);
  r.y += T(o[1]);
  r.z += T(o[2]);
  r.w += T(o[3]);
  return r;
}
// SamplerTable::sample (sampler.h:73-138)
template <class T> static V4<T> sampler_sample(const Scene &s, int id, const V4<T> &attr) {
  if (id < 0 || size_t(id) >= s.samplers.size()) return {T(0), T(0), T(0), T(1)};
  const phx_sampler_info &si = s.samplers[size_t(id)];
  V4<T> c = mat4_apply<T>(si.in_transform, si.in_offset, attr);
  V4<T> result = c;
  if (si.type != PHX_SAMPLER_TRANSFORM) {
    const int dims = si.type == PHX_SAMPLER_IMAGE1D ? 1 : (si.type == PHX_SAMPLER_IMAGE2D ? 2 : 3);
    const T coord[3] = {c.x, c.y, c.z};
    int i0[3] = {0, 0, 0}, i1[3] = {0, 0, 0};
    T frac[3] = {T(0), T(0), T(0)};
    for (int d = 0; d < dims; ++d) {
      const int n = int(si.size[d]);
      T x = coord[d];
      if (!(x > T(-1e6) && x < T(1e6))) x = T(0);
      if (si.linear) {
        const T f = x * T(n) - T(0.5);
        const T fl = std::floor(f);
        frac[d] = f - fl;
        i0[d] = wrap_index(int(fl), n, si.wrap[d]);
        i1[d] = wrap_index(int(fl) + 1, n, si.wrap[d]);
      } else {
        i0[d] = wrap_index(int(std::floor(x * T(n))), n, si.wrap[d]);
        i1[d] = i0[d];
      }
    }
    const uint32_t sx = si.size[0], sxy = si.size[0] * si.size[1], base = si.offset;
    auto texel = [&](uint32_t i) {
      const float *t = &s.sampler_texels[size_t(i) * 4];
      return V4<T>{T(t[0]), T(t[1]), T(t[2]), T(t[3])};
    };
    V4<T> acc{};
    const int nz = dims > 2 ? 2 : 1, ny = dims > 1 ? 2 : 1;
    for (int kz = 0; kz < nz; ++kz) {
      const T wz = dims > 2 ? (kz ? frac[2] : T(1) - frac[2]) : T(1);
      const uint32_t z = uint32_t(kz ? i1[2] : i0[2]);
      for (int ky = 0; ky < ny; ++ky) {
        const T wy = dims > 1 ? (ky ? frac[1] : T(1) - frac[1]) : T(1);
        const uint32_t y = uint32_t(ky ? i1[1] : i0[1]);
        const V4<T> a = texel(base + z * sxy + y * sx + uint32_t(i0[0]));
        const V4<T> b = texel(base + z * sxy + y * sx + uint32_t(i1[0]));
        const T wt = wy * wz;
        acc.x = acc.x + (a.x * (T(1) - frac[0]) + b.x * frac[0]) * wt;
        acc.y = acc.y + (a.y * (T(1) - frac[0]) + b.y * frac[0]) * wt;
        acc.z = acc.z + (a.z * (T(1) - frac[0]) + b.z * frac[0]) * wt;
        acc.w = acc.w + (a.w * (T(1) - frac[0]) + b.w * frac[0]) * wt;
      }
    }
    result = acc;
  }
  return mat4_apply<T>(si.out_transform, si.out_offset, result);
}
// eval_source (pathtracer.cpp:125-133)
template <class T> static V4<T> eval_source(const Scene &s, const phx_param_source &src, const Query<T> &q) {
  if (src.kind == PHX_SOURCE_ATTRIBUTE) return fetch_attribute<T>(s, src.index, q);
  if (src.index < 0 || size_t(src.index) >= s.samplers.size()) return {T(0), T(0), T(0), T(1)};
  return sampler_sample<T>(s, src.index, fetch_attribute<T>(s, s.samplers[size_t(src.index)].in_attribute, q));
}
// apply_normal_map (pathtracer.cpp:135-176)
template <class T>
static V3<T> apply_normal_map(const Scene &s, const phx_param_source &src, const Query<T> &q, const V3<T> &sn) {
  const V4<T> c = eval_source<T>(s, src, q);
  const V3<T> nm{T(2) * c.x - T(1), T(2) * c.y - T(1), T(2) * c.z - T(1)};
  int slot = src.index;
  if (src.kind != PHX_SOURCE_ATTRIBUTE) {
    if (src.index < 0 || size_t(src.index) >= s.samplers.size()) return sn;
    slot = s.samplers[size_t(src.index)].in_attribute;
  }
  Query<T> corner = q;
  corner.b1 = T(0);
  corner.b2 = T(0);
  const V4<T> a0 = fetch_attribute<T>(s, slot, corner);
  corner.b1 = T(1);
  const V4<T> a1 = fetch_attribute<T>(s, slot, corner);
  corner.b1 = T(0);
  corner.b2 = T(1);
  const V4<T> a2 = fetch_attribute<T>(s, slot, corner);
  const T du1 = a1.x - a0.x, dv1 = a1.y - a0.y, du2 = a2.x - a0.x, dv2 = a2.y - a0.y;
  const T det = du1 * dv2 - du2 * dv1;
  if (!(std::fabs(det) > T(1e-20))) return sn;
  const Mesh &m = s.mesh;
  const V3<T> p0 = v3<T>(&m.pos[size_t(m.idx[q.prim * 3]) * 3]);
  const V3<T> e1 = v3<T>(&m.pos[size_t(m.idx[q.prim * 3 + 1]) * 3]) - p0;
  const V3<T> e2 = v3<T>(&m.pos[size_t(m.idx[q.prim * 3 + 2]) * 3]) - p0;
  V3<T> t = (e1 * dv2 - e2 * dv1) * (T(1) / det);
  V3<T> b = (e2 * du1 - e1 * du2) * (T(1) / det);
  t = t - sn * dot(sn, t);
  b = b - sn * dot(sn, b);
  if (!(dot(t, t) > T(0)) || !(dot(b, b) > T(0))) return sn;
  const V3<T> n = normalize(t) * nm.x + normalize(b) * nm.y + sn * nm.z;
  return dot(n, n) > T(0) ? normalize(n) : sn;
}
// resolve_material (pathtracer.cpp:186-247): textures, then sources, then
// the clamps and the alpha mode.
template <class T> static Mat<T> resolve_material(const Scene &s, const Query<T> &q, T &alpha, const Mat<T> *override_mats) {
  const uint32_t mid = q.mat;
  Mat<T> m = override_mats ? override_mats[mid] : (s.mats.empty() ? Mat<T>{} : mat_from<T>(s.mats[mid]));
  const bool tex = !s.atlas.infos.empty();
  alpha = m.alpha;
  if (m.alpha_tex >= 0 && tex) alpha = s.atlas.sample<T>(m.alpha_tex, q.uv.x, q.uv.y).x;
  if (m.base_color_tex >= 0 && tex) m.base_color = s.atlas.sample<T>(m.base_color_tex, q.uv.x, q.uv.y);
  else if (q.has_color) m.base_color = q.color;
  if (m.roughness_tex >= 0 && tex) m.roughness = s.atlas.sample<T>(m.roughness_tex, q.uv.x, q.uv.y).x;
  if (m.metallic_tex >= 0 && tex) m.metallic = s.atlas.sample<T>(m.metallic_tex, q.uv.x, q.uv.y).x;
  auto c01 = [](T x) { return x < T(0) ? T(0) : (x > T(1) ? T(1) : x); };
  if (m.base_color_src.kind != PHX_SOURCE_CONSTANT) {
    const V4<T> c = eval_source<T>(s, m.base_color_src, q);
    m.base_color = {fmaxT(c.x, T(0)), fmaxT(c.y, T(0)), fmaxT(c.z, T(0))};
    if (m.base_color_alpha) alpha *= c.w;
  }
  if (m.alpha_src.kind != PHX_SOURCE_CONSTANT) alpha *= eval_source<T>(s, m.alpha_src, q).x;
  if (m.metallic_src.kind != PHX_SOURCE_CONSTANT) m.metallic = eval_source<T>(s, m.metallic_src, q).x;
  if (m.roughness_src.kind != PHX_SOURCE_CONSTANT) {
    const T r = c01(eval_source<T>(s, m.roughness_src, q).x);
    m.roughness = m.roughness_src_squared ? r * r : r;
  }
  if (m.transmission_src.kind != PHX_SOURCE_CONSTANT) m.transmission = eval_source<T>(s, m.transmission_src, q).x;
  if (m.emission_src.kind != PHX_SOURCE_CONSTANT) {
    const V4<T> e = eval_source<T>(s, m.emission_src, q);
    m.emission = {fmaxT(e.x, T(0)), fmaxT(e.y, T(0)), fmaxT(e.z, T(0))};
  }
  m.metallic = c01(m.metallic);
  m.roughness = c01(m.roughness);
  m.transmission = c01(m.transmission);
  if (m.alpha_mode == PHX_ALPHA_OPAQUE) alpha = T(1);
  else if (m.alpha_mode == PHX_ALPHA_MASK) alpha = alpha >= m.alpha_cutoff ? T(1) : T(0);
  else alpha = c01(alpha);
  return m;
}
// sample_background_image (pathtracer.cpp:370-389) over the RGBA rows of the
// render parameters' background image.
template <class T> static V4<T> background_at(const phx_render_params &rp, T bs, T bt) {
  const int W = int(rp.background_width), H = int(rp.background_height);
  const T fx = bs * T(W) - T(0.5), fy = bt * T(H) - T(0.5);
  const T flx = std::floor(fx), fly = std::floor(fy);
  const T tx = fx - flx, ty = fy - fly;
  auto cl = [](int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); };
  const int x0 = cl(int(flx), W), x1 = cl(int(flx) + 1, W), y0 = cl(int(fly), H), y1 = cl(int(fly) + 1, H);
  auto px = [&](int x, int y) {
    const float *p = rp.background_image + (size_t(y) * W + x) * 4;
    return V4<T>{T(p[0]), T(p[1]), T(p[2]), T(p[3])};
  };
  const V4<T> p00 = px(x0, y0), p01 = px(x1, y0), p10 = px(x0, y1), p11 = px(x1, y1);
  auto lerp4 = [](const V4<T> &a, const V4<T> &b, T t) {
    return V4<T>{a.x * (T(1) - t) + b.x * t, a.y * (T(1) - t) + b.y * t, a.z * (T(1) - t) + b.z * t,
                 a.w * (T(1) - t) + b.w * t};
  };
  return lerp4(lerp4(p00, p01, tx), lerp4(p10, p11, tx), ty);
}
// The camera-ray background of pixel idx (pathtracer.cpp:862-877): the
// image's colour times its alpha, and the alpha.
template <class T> static V3<T> camera_background(const phx_render_params &rp, uint32_t idx, T &bg_alpha) {
  if (rp.background_image && rp.background_width && rp.background_height) {
    const uint32_t x = idx % rp.width, y = idx / rp.width;
    const T bs = (T(x) + T(0.5)) / T(rp.width), bt = (T(y) + T(0.5)) / T(rp.height);
    const V4<T> bg = background_at<T>(rp, bs, rp.bottom_up ? bt : T(1) - bt);
    bg_alpha = bg.w;
    return V3<T>{bg.x, bg.y, bg.z} * bg.w;
  }
  bg_alpha = T(rp.background_alpha);
  return v3<T>(rp.background_color) * T(rp.background_alpha);
}
template <class T> static bool emits_on_side(const Mat<T> &m, bool front) {  // pathtracer.cpp:250-256
  if (m.flags & PHX_MATERIAL_FLAG_EMIT_BOTH_SIDES) return true;
  const bool back_only = (m.flags & PHX_MATERIAL_FLAG_EMIT_BACK_ONLY) != 0;
  return front ? !back_only : back_only;
}
// pathtracer.cpp:40-47
template <class T> static V3<T> clamp_contribution(const V3<T> &c, int lb, T ds, T is) {
  const T limit = lb > 0 ? is : ds;
  if (!(limit > T(0))) return c;
  const T sum = std::fabs(c.x) + std::fabs(c.y) + std::fabs(c.z);
  return sum > limit ? c * (limit / sum) : c;
}
// pathtracer.cpp:51-58
static bool count_bounce(uint32_t &kinds, uint32_t shift, int limit) {
  uint32_t count = (kinds >> shift) & 0x7fu;
  if (count < 0x7fu) ++count;
  kinds = (kinds & ~(0x7fu << shift)) | (count << shift);
  return limit >= 0 && int(count) > limit;
}
static const uint32_t PATH_EMISSION_ONLY = 0x80000000u;

template <class T> struct Direct { V3<T> wi, L; T pdf{0}, dist{0}; bool mis{false}, valid{false}; };
// pathtracer.cpp:270-367
template <class T> static Direct<T> sample_direct(const Scene &s, const V3<T> &p, Rng &rng, uint32_t strategies,
                                                  const Mat<T> *ov) {
  Direct<T> ds;
  if (strategies == 0) return ds;
  const uint32_t nl = uint32_t(s.nee_lights.size());
  const T pick_pdf = T(1) / T(strategies);
  uint32_t pick = uint32_t(rng.next_f32() * float(strategies));
  if (pick >= strategies) pick = strategies - 1;
  if (pick < nl) {
    const LightSample<T> ls = sample_light<T>(s.nee_lights[pick], p, rng);
    if (!(ls.pdf > T(0))) return ds;
    ds.wi = ls.wi;
    ds.L = ls.Li;
    ds.pdf = ls.pdf * pick_pdf;
    ds.dist = ls.dist;
    ds.mis = false;
    ds.valid = true;
    return ds;
  }
  if (s.emissive_count > 0 && pick == nl) {
    const float xi = rng.next_f32();
    uint32_t lo = 0, hi = s.emissive_count - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) / 2;
      if (s.emissive_cdf[mid] <= xi) lo = mid + 1;
      else hi = mid;
    }
    const uint32_t prim = s.emissive_ids[lo];
    const T pdf_area = T(s.prim_light_pdf[prim]);
    if (!(pdf_area > T(0))) return ds;
    const Mesh &m = s.mesh;
    const V3<T> v0 = v3<T>(&m.pos[size_t(m.idx[prim * 3]) * 3]), v1 = v3<T>(&m.pos[size_t(m.idx[prim * 3 + 1]) * 3]),
                v2 = v3<T>(&m.pos[size_t(m.idx[prim * 3 + 2]) * 3]);
    const T su = std::sqrt(draw<T>(rng));
    const T r2 = draw<T>(rng);
    const T b0 = T(1) - su, b1 = r2 * su, b2 = T(1) - b0 - b1;
    const V3<T> lp = v0 * b0 + v1 * b1 + v2 * b2;
    const V3<T> to = lp - p;
    const T d2 = dot(to, to);
    if (!(d2 > T(0))) return ds;
    const T dist = std::sqrt(d2);
    const V3<T> wi = to * (T(1) / dist);
    const V3<T> ln = normalize(cross(v1 - v0, v2 - v0));
    const T cl = -dot(wi, ln);
    if (std::fabs(cl) < T(1e-6)) return ds;
    // resolve_material at the light point: its texcoords are not interpolated
    // there (SurfaceQuery::uv stays 0), pathtracer.cpp:325-341.
    T alpha = T(1);
    Query<T> q;
    q.prim = prim;
    q.mat = m.mat.empty() ? 0u : m.mat[prim];
    q.b1 = b1;
    q.b2 = b2;
    q.p = lp;
    q.sn = ln;
    const Mat<T> em = resolve_material<T>(s, q, alpha, ov);
    if (!emits_on_side(em, cl > T(0))) return ds;
    ds.wi = wi;
    ds.L = material_emission(em) * alpha;
    ds.pdf = pdf_area * d2 / std::fabs(cl) * pick_pdf;
    ds.dist = dist;
    ds.mis = true;
    ds.valid = true;
    return ds;
  }
  if (s.has_env) {
    T ep = T(0);
    const V3<T> wi = s.env.sample_direction<T>(rng, ep);
    if (!(ep > T(0))) return ds;
    ds.wi = wi;
    ds.L = s.env.evaluate<T>(wi);
    ds.pdf = ep * pick_pdf;
    ds.dist = T(1e30);
    ds.mis = true;
    ds.valid = true;
  }
  return ds;
}

struct Params {
  phx_render_params p;
  phx_camera cam;
};

// Per-pixel accumulators of PathTracer::render (pathtracer.cpp:505-562).
template <class T> struct PixelOut {
  V3<T> accum{};
  T alpha{0};
  uint32_t ns{0};
  V3<T> albedo{}, normal{};
  T depth{T(1e30)};
};

// One pixel, all its samples: pathtracer.cpp:688-1127.  The reference runs the
// bounce loop over all pixels in lockstep; each pixel's state only depends on
// its own streams, so running the pixels one after another gives the same
// values.  `ov` replaces the scene materials (finite-difference engine).
template <class T>
static void render_pixel(const Scene &s, const Params &P, uint32_t idx, PixelOut<T> &out, const Mat<T> *ov) {
  const phx_render_params &rp = P.p;
  const uint32_t w = rp.width, h = rp.height;
  const uint32_t max_depth = rp.max_depth > 0 ? rp.max_depth : 1u;
  const uint32_t strategies = uint32_t(s.nee_lights.size()) + (s.emissive_count > 0 ? 1u : 0u) + (s.has_env ? 1u : 0u);
  const uint32_t max_it = max_depth + 1u + 8u;
  const bool has_vol = !s.vol.empty();
  const T eps = T(s.ray_eps);
  const T cd = T(rp.clamp_direct_sum), ci = T(rp.clamp_indirect_sum);
  const V3<T> ambient = v3<T>(rp.ambient_color) * T(rp.ambient_radiance);
  const uint32_t x = idx % w, y = idx / w;
  out = PixelOut<T>{};
  V3<T> accum_sq{};
  for (uint32_t sidx = 0; sidx < rp.samples_per_pixel; ++sidx) {
    const uint32_t sample = sidx + rp.sample_offset;
    // adaptive sampling (pathtracer.cpp:676-723): from sample 16 on, a pixel
    // with 16 or more samples stops once its largest channel variance over
    // its squared luminance is below 0.01 (its sums no longer change, so it
    // stays stopped)
    if (rp.adaptive_sampling && sidx >= 16u && out.ns >= 16u) {
      const T inv = T(1) / T(out.ns);
      const V3<T> mean = out.accum * inv, sq = accum_sq * inv;
      const V3<T> var = sq - V3<T>{mean.x * mean.x, mean.y * mean.y, mean.z * mean.z};
      const T max_var = fmaxT(var.x, fmaxT(var.y, var.z));
      const T lum = fmaxT(T(0.2126f) * mean.x + T(0.7152f) * mean.y + T(0.0722f) * mean.z, T(1e-4f));
      if (max_var / (lum * lum) < T(0.01f)) break;
    }
    const V3<T> sample_start = out.accum;
    // pt_init_sample: pathtracer.cpp:727-759
    Rng crng = make_rng(idx, sample, 0u);
    T u = (T(x) + draw<T>(crng)) / T(w);
    T v = (T(y) + draw<T>(crng)) / T(h);
    if (!rp.bottom_up) v = T(1) - v;
    u = T(rp.image_region[0]) + (T(rp.image_region[2]) - T(rp.image_region[0])) * u;
    v = T(rp.image_region[1]) + (T(rp.image_region[3]) - T(rp.image_region[1])) * v;
    V3<T> ro, rd;
    camera_ray<T>(P.cam, u, v, crng, ro, rd);
    T rtmin, rtmax;
    clip_range<T>(P.cam, rd, rtmin, rtmax);
    V3<T> tp{T(1), T(1), T(1)};
    bool active = true, primary = true;
    T prev_pdf = T(-1), path_dist = T(0);
    uint32_t depth = 0, kinds = 0;
    out.ns += 1;
    for (uint32_t it = 0; it < max_it && active; ++it) {
      const Hit<T> hr = closest<T>(s, ro, rd, rtmin, rtmax);
      Rng rng = make_rng(idx, sample, 1u + it);
      const bool emission_only = (kinds & PATH_EMISSION_ONLY) != 0u;
      bool nee = false;
      V3<T> nee_c{}, so, sd;
      T stmin = T(1), stmax = T(0);
      // pt_shade: pathtracer.cpp:787-1089
      if (has_vol) {
        V3<T> alb{};
        const T t_end = hr.hit ? hr.t : rtmax;
        const T ts = s.vol.sample_scatter<T>(ro, rd, rtmin, t_end, rng, alb);
        if (ts >= T(0)) {
          const V3<T> p = ro + rd * ts;
          if (primary) {  // :800-809 (the AOVs come from pixel_aov)
            primary = false;
            out.alpha += T(1);
          }
          if (depth >= max_depth || emission_only) {
            active = false;
          } else {
            const V3<T> tpv = tp * alb;
            const T phase = T(0.25) * K<T>::INV_PI;
            const Direct<T> ds = sample_direct<T>(s, p, rng, strategies, ov);
            if (ds.valid) {
              const T wm = ds.mis ? power_heuristic(1, ds.pdf, 1, phase) : T(1);
              nee_c = clamp_contribution(tpv * ds.L * (phase * wm / ds.pdf), int(depth), cd, ci);
              nee = true;
              so = p;
              sd = ds.wi;
              stmin = T(0);
              stmax = ds.dist < T(1e29) ? ds.dist * (T(1) - T(1e-4)) : T(1e30);
            }
            V3<T> tn = tpv;
            bool killed = false;
            if (depth >= 2) {
              const T mt = fmaxT(tpv.x, fmaxT(tpv.y, tpv.z));
              const T surv = fminT(fmaxT(mt, T(0.05)), T(0.95));
              if (draw<T>(rng) >= surv) killed = true;
              else tn = tpv * (T(1) / surv);
            }
            if (killed) {
              active = false;
            } else {
              tp = tn;
              prev_pdf = phase;
              depth += 1;
              if (count_bounce(kinds, 24u, rp.max_volume_bounces)) kinds |= PATH_EMISSION_ONLY;
              ro = p;
              rd = sample_uniform_sphere<T>(rng);
              rtmin = T(0);
              rtmax = T(1e30);
            }
          }
          goto pack;
        }
      }
      if (!hr.hit) {
        V3<T> L{};
        if (primary) {
          if (s.has_env && s.env_visible) {
            L = s.env.evaluate<T>(rd);
            out.alpha += T(1);
          } else {
            T ba;
            L = camera_background<T>(rp, idx, ba);
            out.alpha += ba;
          }
        } else {
          L = ambient;
          if (s.has_env) {
            T wm = T(1);
            if (prev_pdf > T(0)) wm = power_heuristic(1, prev_pdf, 1, s.env.pdf<T>(rd) / T(strategies));
            L = L + s.env.evaluate<T>(rd) * wm;
          }
        }
        V3<T> fm = tp * L;
        if (!primary || (s.has_env && s.env_visible)) fm = clamp_contribution(fm, int(depth) - 1, cd, ci);
        out.accum = out.accum + fm;
        active = false;
        goto pack;
      }
      {
        const V3<T> p = hr.p;
        const V3<T> wo = -rd;
        V3<T> hsn = hr.sn;
        if (!(dot(hsn, hsn) > T(0.25))) hsn = hr.n;
        else if (dot(hsn, hr.n) < T(0)) hsn = hsn * T(-1);
        T alpha = T(1);
        const Query<T> hq = query_of<T>(hr, hsn);
        const Mat<T> mat = resolve_material<T>(s, hq, alpha, ov);
        const bool hidden = (mat.flags & PHX_MATERIAL_FLAG_HIDDEN_TO_CAMERA) != 0 && primary;
        if (hidden || (alpha < T(1) && draw<T>(rng) >= alpha)) {
          ro = p;
          rtmin = eps;
          rtmax = T(1e30);
          path_dist += hr.t;
          if (primary) {
            T tn, tf;
            clip_range<T>(P.cam, rd, tn, tf);
            if (tf < T(1e30)) {
              rtmax = tf - path_dist;
              if (rtmax < eps) {
                rtmin = T(1);
                rtmax = T(0);
              }
            }
          }
          prev_pdf = T(-1);
          goto pack;
        }
        V3<T> n = hr.n, sn = hsn;
        const bool front = dot(n, wo) >= T(0);
        if (!(mat.transmission > T(0)) && !front) {
          n = n * T(-1);
          sn = sn * T(-1);
        }
        if (mat.normal_tex >= 0 && !s.atlas.infos.empty()) {
          V3<T> nm = s.atlas.sample<T>(mat.normal_tex, hr.uv.x, hr.uv.y);
          nm = nm * T(2) - V3<T>{T(1), T(1), T(1)};
          V3<T> ta, bi;
          onb_from_normal(sn, ta, bi);
          sn = normalize(ta * nm.x + bi * nm.y + sn * nm.z);
        }
        if (mat.normal_src.kind != PHX_SOURCE_CONSTANT) sn = apply_normal_map<T>(s, mat.normal_src, hq, sn);
        if (primary) {  // :974-983 (the AOVs come from pixel_aov)
          primary = false;
          out.alpha += T(1);
        }
        if (material_may_emit(mat) && emits_on_side(mat, front)) {
          T wm = T(1);
          if (prev_pdf > T(0) && s.emissive_count > 0) {
            const T pa = T(s.prim_light_pdf[hr.prim]);
            const T cl = std::fabs(dot(hr.n, wo));
            if (pa > T(0) && cl > T(0)) wm = power_heuristic(1, prev_pdf, 1, pa * hr.t * hr.t / cl / T(strategies));
          }
          out.accum = out.accum + clamp_contribution(tp * material_emission(mat) * wm, int(depth) - 1, cd, ci);
        }
        if ((mat.flags & PHX_MATERIAL_FLAG_EMITTER_ONLY) != 0 || depth >= max_depth || emission_only) {
          active = false;
          goto pack;
        }
        {
          const Direct<T> ds = sample_direct<T>(s, p, rng, strategies, ov);
          if (ds.valid) {
            const V3<T> f = disney_bsdf_eval(mat, wo, ds.wi, n, sn);
            if (f.x > T(0) || f.y > T(0) || f.z > T(0)) {
              T wm = T(1);
              if (ds.mis) wm = power_heuristic(1, ds.pdf, 1, disney_bsdf_pdf(mat, wo, ds.wi, n, sn));
              nee_c = clamp_contribution(tp * f * ds.L * (wm / ds.pdf), int(depth), cd, ci);
              nee = true;
              const V3<T> sorg = p + n * eps;
              so = sorg;
              stmin = T(0);
              if (ds.dist < T(1e29)) {
                const V3<T> tl = p + ds.wi * ds.dist - sorg;
                const T sdist = length(tl);
                sd = sdist > T(0) ? tl * (T(1) / sdist) : ds.wi;
                stmax = sdist * (T(1) - T(1e-3));
              } else {
                sd = ds.wi;
                stmax = T(1e30);
              }
            }
          }
        }
        const BsdfSample<T> bs = disney_bsdf_sample(mat, wo, n, sn, rng);
        if (bs.pdf <= T(0) || (bs.f.x <= T(0) && bs.f.y <= T(0) && bs.f.z <= T(0))) {
          active = false;
          goto pack;
        }
        V3<T> tpn = tp * (bs.f * (T(1) / bs.pdf));
        if (depth >= 2) {
          const T mt = fmaxT(tpn.x, fmaxT(tpn.y, tpn.z));
          const T surv = fminT(fmaxT(mt, T(0.05)), T(0.95));
          if (draw<T>(rng) >= surv) {
            active = false;
            goto pack;
          }
          tpn = tpn * (T(1) / surv);
        }
        if (rp.throughput_clamp > 0.f) {
          const T mm = fmaxT(tpn.x, fmaxT(tpn.y, tpn.z));
          if (mm > T(rp.throughput_clamp)) tpn = tpn * (T(rp.throughput_clamp) / mm);
        }
        if (!(tpn.x == tpn.x && tpn.y == tpn.y && tpn.z == tpn.z)) {
          active = false;
          goto pack;
        }
        tp = tpn;
        prev_pdf = bs.is_specular ? T(-1) : bs.pdf;
        depth += 1;
        const V3<T> wi = normalize(bs.wi);
        const T side = dot(wi, hr.n) < T(0) ? T(-1) : T(1);
        const bool through = (dot(wi, hr.n) < T(0)) != (dot(wo, hr.n) < T(0));
        const bool exceeded = through ? count_bounce(kinds, 16u, rp.max_transmission_bounces)
                                      : (bs.is_diffuse ? count_bounce(kinds, 0u, rp.max_diffuse_bounces)
                                                       : count_bounce(kinds, 8u, rp.max_glossy_bounces));
        if (exceeded) kinds |= PATH_EMISSION_ONLY;
        ro = p + hr.n * (side * eps);
        rd = wi;
        rtmin = T(0);
        rtmax = T(1e30);
      }
    pack:
      // trace_occluded + pt_nee_accum_pack: pathtracer.cpp:1092-1121
      if (nee) {
        const V3<float> sof(so), sdf(sd);
        if (!occluded(s, sof, sdf, float(stmin), float(stmax))) {
          V3<T> c = nee_c;
          if (has_vol) {
            Rng trng = make_rng(idx, sample, 0x10000u + it);
            c = c * s.vol.transmittance<T>(so, sd, stmin, stmax, trng);
          }
          if (c.x == c.x && c.y == c.y && c.z == c.z) out.accum = out.accum + c;
        }
      }
    }
    // pt_update_variance (pathtracer.cpp:1129-1139)
    if (rp.adaptive_sampling) {
      const V3<T> sv = out.accum - sample_start;
      accum_sq = accum_sq + V3<T>{sv.x * sv.x, sv.y * sv.y, sv.z * sv.z};
    }
  }
}

// ---- detached finite-difference engine ----------------------------------------
// The estimator the path-replay adjoint differentiates: every sampling
// decision is the float integrator's (render_pixel<float>, the scene's
// materials), while the weights that depend on the trainable parameters are
// evaluated in double with `eval` materials and texels: the BSDF f * cos at
// next-event estimation and at the continuing sample (f / pdf with the float
// pdf), times the parameter-free factors (emission, MIS weights, light pdfs,
// Russian-roulette survival, volume albedo and transmittance).  Delta and
// transmission samples keep their float weight.  Clamps are not applied.
// Central differences of this estimator in double are the reference the
// adjoint's gradients are checked against.
struct DetachedEval {
  const double *mat_params;  // material_count * 5 (base r, g, b, roughness, metallic) or null
  uint32_t mat_count;
  const double *texels;      // atlas pixels * 3 or null (the scene's)
};
static Mat<double> detached_material(const Scene &s, const DetachedEval &e, uint32_t mid, const V2<float> &uv,
                                     bool has_color, const V3<float> &color) {
  Mat<double> m = s.mats.empty() ? Mat<double>{} : mat_from<double>(s.mats[mid]);
  if (e.mat_params && mid < e.mat_count) {
    const double *q = e.mat_params + size_t(mid) * 5;
    m.base_color = {q[0], q[1], q[2]};
    m.roughness = q[3];
    m.metallic = q[4];
  }
  const bool tex = !s.atlas.infos.empty();
  const V2<double> uvd{double(uv.x), double(uv.y)};
  if (m.base_color_tex >= 0 && tex)
    m.base_color = e.texels ? s.atlas.sample_from<double>(e.texels, m.base_color_tex, uvd.x, uvd.y)
                            : s.atlas.sample<double>(m.base_color_tex, uvd.x, uvd.y);
  else if (has_color)
    m.base_color = V3<double>(color);
  if (m.roughness_tex >= 0 && tex) m.roughness = s.atlas.sample<double>(m.roughness_tex, uvd.x, uvd.y).x;
  if (m.metallic_tex >= 0 && tex) m.metallic = s.atlas.sample<double>(m.metallic_tex, uvd.x, uvd.y).x;
  auto c01 = [](double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); };
  m.metallic = c01(m.metallic);
  m.roughness = c01(m.roughness);
  m.transmission = c01(m.transmission);
  return m;
}

static V3<double> detached_pixel(const Scene &s, const Params &P, uint32_t idx, const DetachedEval &E) {
  const phx_render_params &rp = P.p;
  const uint32_t w = rp.width, h = rp.height;
  const uint32_t max_depth = rp.max_depth > 0 ? rp.max_depth : 1u;
  const uint32_t strategies = uint32_t(s.nee_lights.size()) + (s.emissive_count > 0 ? 1u : 0u) + (s.has_env ? 1u : 0u);
  const uint32_t max_it = max_depth + 1u + 8u;
  const bool has_vol = !s.vol.empty();
  const float eps = s.ray_eps;
  const V3<float> ambient = v3<float>(rp.ambient_color) * rp.ambient_radiance;
  const uint32_t x = idx % w, y = idx / w;
  V3<double> acc{};
  auto wv = [](const V3<float> &v) { return V3<double>(v); };
  for (uint32_t sidx = 0; sidx < rp.samples_per_pixel; ++sidx) {
    const uint32_t sample = sidx + rp.sample_offset;
    Rng crng = make_rng(idx, sample, 0u);
    float u = (float(x) + draw<float>(crng)) / float(w);
    float v = (float(y) + draw<float>(crng)) / float(h);
    if (!rp.bottom_up) v = 1.f - v;
    u = rp.image_region[0] + (rp.image_region[2] - rp.image_region[0]) * u;
    v = rp.image_region[1] + (rp.image_region[3] - rp.image_region[1]) * v;
    V3<float> ro, rd;
    camera_ray<float>(P.cam, u, v, crng, ro, rd);
    float rtmin, rtmax;
    clip_range<float>(P.cam, rd, rtmin, rtmax);
    V3<float> tp{1.f, 1.f, 1.f};
    V3<double> tw{1.0, 1.0, 1.0};  // the parameter-dependent throughput
    bool active = true, primary = true;
    float prev_pdf = -1.f, path_dist = 0.f;
    uint32_t depth = 0, kinds = 0;
    for (uint32_t it = 0; it < max_it && active; ++it) {
      const Hit<float> hr = closest<float>(s, ro, rd, rtmin, rtmax);
      Rng rng = make_rng(idx, sample, 1u + it);
      const bool emission_only = (kinds & PATH_EMISSION_ONLY) != 0u;
      bool nee = false;
      V3<double> nee_w{};
      V3<float> so, sd;
      float stmin = 1.f, stmax = 0.f;
      if (has_vol) {
        V3<float> alb{};
        const float t_end = hr.hit ? hr.t : rtmax;
        const float ts = s.vol.sample_scatter<float>(ro, rd, rtmin, t_end, rng, alb);
        if (ts >= 0.f) {
          const V3<float> p = ro + rd * ts;
          if (primary) primary = false;
          if (depth >= max_depth || emission_only) {
            active = false;
          } else {
            const V3<float> tpv = tp * alb;
            const V3<double> twv = tw * wv(alb);
            const float phase = 0.25f * K<float>::INV_PI;
            const Direct<float> ds = sample_direct<float>(s, p, rng, strategies, nullptr);
            if (ds.valid) {
              const float wm = ds.mis ? power_heuristic(1, ds.pdf, 1, phase) : 1.f;
              nee_w = twv * wv(ds.L) * (double(phase) * double(wm) / double(ds.pdf));
              nee = true;
              so = p;
              sd = ds.wi;
              stmin = 0.f;
              stmax = ds.dist < 1e29f ? ds.dist * (1.f - 1e-4f) : 1e30f;
            }
            V3<float> tn = tpv;
            V3<double> twn = twv;
            bool killed = false;
            if (depth >= 2) {
              const float mt = fmaxT(tpv.x, fmaxT(tpv.y, tpv.z));
              const float surv = fminT(fmaxT(mt, 0.05f), 0.95f);
              if (draw<float>(rng) >= surv) killed = true;
              else {
                tn = tpv * (1.f / surv);
                twn = twv * (1.0 / double(surv));
              }
            }
            if (killed) {
              active = false;
            } else {
              tp = tn;
              tw = twn;
              prev_pdf = phase;
              depth += 1;
              if (count_bounce(kinds, 24u, rp.max_volume_bounces)) kinds |= PATH_EMISSION_ONLY;
              ro = p;
              rd = sample_uniform_sphere<float>(rng);
              rtmin = 0.f;
              rtmax = 1e30f;
            }
          }
          goto pack;
        }
      }
      if (!hr.hit) {
        V3<float> L{};
        if (primary) {
          float ba;
          L = (s.has_env && s.env_visible) ? s.env.evaluate<float>(rd) : camera_background<float>(rp, idx, ba);
        } else {
          L = ambient;
          if (s.has_env) {
            float wm = 1.f;
            if (prev_pdf > 0.f) wm = power_heuristic(1, prev_pdf, 1, s.env.pdf<float>(rd) / float(strategies));
            L = L + s.env.evaluate<float>(rd) * wm;
          }
        }
        acc = acc + tw * wv(L);
        active = false;
        goto pack;
      }
      {
        const V3<float> p = hr.p;
        const V3<float> wo = -rd;
        V3<float> hsn = hr.sn;
        if (!(dot(hsn, hsn) > 0.25f)) hsn = hr.n;
        else if (dot(hsn, hr.n) < 0.f) hsn = hsn * -1.f;
        float alpha = 1.f;
        const Query<float> hq = query_of<float>(hr, hsn);
        const Mat<float> mat = resolve_material<float>(s, hq, alpha, nullptr);
        const bool hidden = (mat.flags & PHX_MATERIAL_FLAG_HIDDEN_TO_CAMERA) != 0 && primary;
        if (hidden || (alpha < 1.f && draw<float>(rng) >= alpha)) {
          ro = p;
          rtmin = eps;
          rtmax = 1e30f;
          path_dist += hr.t;
          if (primary) {
            float tn, tf;
            clip_range<float>(P.cam, rd, tn, tf);
            if (tf < 1e30f) {
              rtmax = tf - path_dist;
              if (rtmax < eps) {
                rtmin = 1.f;
                rtmax = 0.f;
              }
            }
          }
          prev_pdf = -1.f;
          goto pack;
        }
        V3<float> n = hr.n, sn = hsn;
        const bool front = dot(n, wo) >= 0.f;
        if (!(mat.transmission > 0.f) && !front) {
          n = n * -1.f;
          sn = sn * -1.f;
        }
        if (mat.normal_tex >= 0 && !s.atlas.infos.empty()) {
          V3<float> nm = s.atlas.sample<float>(mat.normal_tex, hr.uv.x, hr.uv.y);
          nm = nm * 2.f - V3<float>{1.f, 1.f, 1.f};
          V3<float> ta, bi;
          onb_from_normal(sn, ta, bi);
          sn = normalize(ta * nm.x + bi * nm.y + sn * nm.z);
        }
        if (mat.normal_src.kind != PHX_SOURCE_CONSTANT) sn = apply_normal_map<float>(s, mat.normal_src, hq, sn);
        const Mat<double> mw = detached_material(s, E, hr.mat, hr.uv, hr.has_color, hr.color);
        if (primary) primary = false;
        if (material_may_emit(mat) && emits_on_side(mat, front)) {
          float wm = 1.f;
          if (prev_pdf > 0.f && s.emissive_count > 0) {
            const float pa = s.prim_light_pdf[hr.prim];
            const float cl = std::fabs(dot(hr.n, wo));
            if (pa > 0.f && cl > 0.f) wm = power_heuristic(1, prev_pdf, 1, pa * hr.t * hr.t / cl / float(strategies));
          }
          acc = acc + tw * wv(material_emission(mat) * wm);
        }
        if ((mat.flags & PHX_MATERIAL_FLAG_EMITTER_ONLY) != 0 || depth >= max_depth || emission_only) {
          active = false;
          goto pack;
        }
        {
          const Direct<float> ds = sample_direct<float>(s, p, rng, strategies, nullptr);
          if (ds.valid) {
            const V3<float> f = disney_bsdf_eval(mat, wo, ds.wi, n, sn);
            if (f.x > 0.f || f.y > 0.f || f.z > 0.f) {
              float wm = 1.f;
              if (ds.mis) wm = power_heuristic(1, ds.pdf, 1, disney_bsdf_pdf(mat, wo, ds.wi, n, sn));
              const V3<double> fw = disney_bsdf_eval(mw, wv(wo), wv(ds.wi), wv(n), wv(sn));
              nee_w = tw * fw * wv(ds.L) * (double(wm) / double(ds.pdf));
              nee = true;
              const V3<float> sorg = p + n * eps;
              so = sorg;
              stmin = 0.f;
              if (ds.dist < 1e29f) {
                const V3<float> tl = p + ds.wi * ds.dist - sorg;
                const float sdist = length(tl);
                sd = sdist > 0.f ? tl * (1.f / sdist) : ds.wi;
                stmax = sdist * (1.f - 1e-3f);
              } else {
                sd = ds.wi;
                stmax = 1e30f;
              }
            }
          }
        }
        const BsdfSample<float> bs = disney_bsdf_sample(mat, wo, n, sn, rng);
        if (bs.pdf <= 0.f || (bs.f.x <= 0.f && bs.f.y <= 0.f && bs.f.z <= 0.f)) {
          active = false;
          goto pack;
        }
        V3<float> tpn = tp * (bs.f * (1.f / bs.pdf));
        const V3<double> fw =
            bs.is_specular ? wv(bs.f) : disney_bsdf_eval(mw, wv(wo), wv(bs.wi), wv(n), wv(sn));
        V3<double> twn = tw * fw * (1.0 / double(bs.pdf));
        if (depth >= 2) {
          const float mt = fmaxT(tpn.x, fmaxT(tpn.y, tpn.z));
          const float surv = fminT(fmaxT(mt, 0.05f), 0.95f);
          if (draw<float>(rng) >= surv) {
            active = false;
            goto pack;
          }
          tpn = tpn * (1.f / surv);
          twn = twn * (1.0 / double(surv));
        }
        if (!(tpn.x == tpn.x && tpn.y == tpn.y && tpn.z == tpn.z)) {
          active = false;
          goto pack;
        }
        tp = tpn;
        tw = twn;
        prev_pdf = bs.is_specular ? -1.f : bs.pdf;
        depth += 1;
        const V3<float> wi = normalize(bs.wi);
        const float side = dot(wi, hr.n) < 0.f ? -1.f : 1.f;
        const bool through = (dot(wi, hr.n) < 0.f) != (dot(wo, hr.n) < 0.f);
        const bool exceeded = through ? count_bounce(kinds, 16u, rp.max_transmission_bounces)
                                      : (bs.is_diffuse ? count_bounce(kinds, 0u, rp.max_diffuse_bounces)
                                                       : count_bounce(kinds, 8u, rp.max_glossy_bounces));
        if (exceeded) kinds |= PATH_EMISSION_ONLY;
        ro = p + hr.n * (side * eps);
        rd = wi;
        rtmin = 0.f;
        rtmax = 1e30f;
      }
    pack:
      if (nee && !occluded(s, so, sd, stmin, stmax)) {
        double tr = 1.0;
        if (has_vol) {
          Rng trng = make_rng(idx, sample, 0x10000u + it);
          tr = double(s.vol.transmittance<float>(so, sd, stmin, stmax, trng));
        }
        const V3<double> c = nee_w * tr;
        if (c.x == c.x && c.y == c.y && c.z == c.z) acc = acc + c;
      }
    }
  }
  return rp.samples_per_pixel ? acc * (1.0 / double(rp.samples_per_pixel)) : acc;
}

// AOVs are written by a separate, simpler pass: the first camera-ray event of
// the first sample (pathtracer.cpp:800-809, 974-983).  They do not feed the
// color, so the color pass above leaves them out.
template <class T> static void pixel_aov(const Scene &s, const Params &P, uint32_t idx, PixelOut<T> &out) {
  const phx_render_params &rp = P.p;
  const uint32_t w = rp.width, h = rp.height, x = idx % w, y = idx / w;
  const uint32_t max_depth = rp.max_depth > 0 ? rp.max_depth : 1u;
  const uint32_t max_it = max_depth + 1u + 8u;
  const T eps = T(s.ray_eps);
  for (uint32_t sidx = 0; sidx < rp.samples_per_pixel; ++sidx) {
    const uint32_t sample = sidx + rp.sample_offset;
    Rng crng = make_rng(idx, sample, 0u);
    T u = (T(x) + draw<T>(crng)) / T(w);
    T v = (T(y) + draw<T>(crng)) / T(h);
    if (!rp.bottom_up) v = T(1) - v;
    u = T(rp.image_region[0]) + (T(rp.image_region[2]) - T(rp.image_region[0])) * u;
    v = T(rp.image_region[1]) + (T(rp.image_region[3]) - T(rp.image_region[1])) * v;
    V3<T> ro, rd;
    camera_ray<T>(P.cam, u, v, crng, ro, rd);
    T rtmin, rtmax;
    clip_range<T>(P.cam, rd, rtmin, rtmax);
    T path_dist = T(0);
    for (uint32_t it = 0; it < max_it; ++it) {
      const Hit<T> hr = closest<T>(s, ro, rd, rtmin, rtmax);
      Rng rng = make_rng(idx, sample, 1u + it);
      if (!s.vol.empty()) {
        V3<T> alb{};
        const T ts = s.vol.sample_scatter<T>(ro, rd, rtmin, hr.hit ? hr.t : rtmax, rng, alb);
        if (ts >= T(0)) {
          out.depth = path_dist + ts;
          out.normal = -rd;
          out.albedo = alb;
          return;
        }
      }
      if (!hr.hit) break;
      V3<T> hsn = hr.sn;
      if (!(dot(hsn, hsn) > T(0.25))) hsn = hr.n;
      else if (dot(hsn, hr.n) < T(0)) hsn = hsn * T(-1);
      T alpha = T(1);
      const Query<T> hq = query_of<T>(hr, hsn);
      const Mat<T> mat = resolve_material<T>(s, hq, alpha, nullptr);
      const bool hidden = (mat.flags & PHX_MATERIAL_FLAG_HIDDEN_TO_CAMERA) != 0;
      if (hidden || (alpha < T(1) && draw<T>(rng) >= alpha)) {
        ro = hr.p;
        rtmin = eps;
        rtmax = T(1e30);
        path_dist += hr.t;
        T tn, tf;
        clip_range<T>(P.cam, rd, tn, tf);
        if (tf < T(1e30)) {
          rtmax = tf - path_dist;
          if (rtmax < eps) break;
        }
        continue;
      }
      V3<T> sn = hsn;
      const V3<T> wo = -rd;
      if (!(mat.transmission > T(0)) && !(dot(hr.n, wo) >= T(0))) sn = sn * T(-1);
      if (mat.normal_tex >= 0 && !s.atlas.infos.empty()) {
        V3<T> nm = s.atlas.sample<T>(mat.normal_tex, hr.uv.x, hr.uv.y);
        nm = nm * T(2) - V3<T>{T(1), T(1), T(1)};
        V3<T> ta, bi;
        onb_from_normal(sn, ta, bi);
        sn = normalize(ta * nm.x + bi * nm.y + sn * nm.z);
      }
      if (mat.normal_src.kind != PHX_SOURCE_CONSTANT) sn = apply_normal_map<T>(s, mat.normal_src, hq, sn);
      out.depth = path_dist + hr.t;
      out.normal = normalize(sn);
      out.albedo = mat.base_color;
      return;
    }
  }
}

}  // namespace orc

using namespace orc;

// ---- extern "C" API used by the tests -------------------------------------
extern "C" {

void *orc_scene_create(const phx_scene_desc *d, int brute) {
  auto *s = new Scene();
  scene_load(*s, *d);
  s->brute = brute != 0;
  return s;
}
void orc_scene_destroy(void *s) { delete static_cast<Scene *>(s); }
float orc_scene_ray_epsilon(void *s) { return static_cast<Scene *>(s)->ray_eps; }
uint32_t orc_scene_bvh_nodes(void *s) { return uint32_t(static_cast<Scene *>(s)->bvh.nodes.size()); }
uint32_t orc_scene_emissive(void *sp, float *cdf, float *prim_pdf, uint32_t *ids) {
  Scene *s = static_cast<Scene *>(sp);
  if (cdf) std::copy(s->emissive_cdf.begin(), s->emissive_cdf.end(), cdf);
  if (prim_pdf) std::copy(s->prim_light_pdf.begin(), s->prim_light_pdf.end(), prim_pdf);
  if (ids) std::copy(s->emissive_ids.begin(), s->emissive_ids.begin() + s->emissive_count, ids);
  return s->emissive_count;
}

// Closest hit of n rays: t, prim (0xffffffff: none), b1, b2.  mode 0: the
// reference BVH, 1: brute force.
void orc_trace_closest(void *sp, const float *org, const float *dir, const float *tmin, const float *tmax, uint32_t n,
                       float *t_out, uint32_t *prim_out, float *u_out, float *v_out, int mode) {
  Scene *s = static_cast<Scene *>(sp);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < int64_t(n); ++i) {
    const V3<float> o = v3<float>(&org[i * 3]), d = v3<float>(&dir[i * 3]);
    const ClosestHit h = mode ? trace_brute(s->mesh, o, d, tmin[i], tmax[i], false)
                              : trace_bvh(s->mesh, s->bvh, o, d, tmin[i], tmax[i], false);
    t_out[i] = h.hit ? h.t : 0.f;
    prim_out[i] = h.hit ? h.prim : 0xffffffffu;
    u_out[i] = h.hit ? h.u : 0.f;
    v_out[i] = h.hit ? h.v : 0.f;
  }
}
// Full HitResult of n rays (mesh_intersector.h:126-178) in phx_hit layout.
void orc_trace_closest_full(void *sp, const float *org, const float *dir, const float *tmin, const float *tmax, uint32_t n,
                            phx_hit *out) {
  Scene *s = static_cast<Scene *>(sp);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < int64_t(n); ++i) {
    const Hit<float> h = closest<float>(*s, v3<float>(&org[i * 3]), v3<float>(&dir[i * 3]), tmin[i], tmax[i]);
    phx_hit r{};
    r.hit = h.hit;
    if (h.hit) {
      r.t = h.t;
      r.position[0] = h.p.x; r.position[1] = h.p.y; r.position[2] = h.p.z;
      r.normal[0] = h.n.x; r.normal[1] = h.n.y; r.normal[2] = h.n.z;
      r.shading_normal[0] = h.sn.x; r.shading_normal[1] = h.sn.y; r.shading_normal[2] = h.sn.z;
      r.uv[0] = h.uv.x; r.uv[1] = h.uv.y;
      r.bary[0] = h.bary.x; r.bary[1] = h.bary.y;
      r.interpolated_color[0] = h.color.x; r.interpolated_color[1] = h.color.y; r.interpolated_color[2] = h.color.z;
      r.prim_id = h.prim;
      r.material_id = h.mat;
      r.has_interpolated_color = h.has_color;
    }
    out[i] = r;
  }
}
void orc_trace_occluded(void *sp, const float *org, const float *dir, const float *tmin, const float *tmax, uint32_t n,
                        uint32_t *occ, int mode) {
  Scene *s = static_cast<Scene *>(sp);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < int64_t(n); ++i) {
    const V3<float> o = v3<float>(&org[i * 3]), d = v3<float>(&dir[i * 3]);
    const ClosestHit h = mode ? trace_brute(s->mesh, o, d, tmin[i], tmax[i], true)
                              : trace_bvh(s->mesh, s->bvh, o, d, tmin[i], tmax[i], true);
    occ[i] = h.hit ? 1u : 0u;
  }
}

// Disney BSDF on n (material, wo, wi, n, sn) tuples: f*cos (3) and pdf.
void orc_bsdf_eval(const phx_material *mats, const float *wo, const float *wi, const float *ng, const float *sn, uint32_t n,
                   float *f_out, float *pdf_out, int use_double) {
  for (uint32_t i = 0; i < n; ++i) {
    if (use_double) {
      const Mat<double> m = mat_from<double>(mats[i]);
      const V3<double> a = v3<double>(&wo[i * 3]), b = v3<double>(&wi[i * 3]), c = v3<double>(&ng[i * 3]), d = v3<double>(&sn[i * 3]);
      const V3<double> f = disney_bsdf_eval(m, a, b, c, d);
      f_out[i * 3] = float(f.x); f_out[i * 3 + 1] = float(f.y); f_out[i * 3 + 2] = float(f.z);
      pdf_out[i] = float(disney_bsdf_pdf(m, a, b, c, d));
    } else {
      const Mat<float> m = mat_from<float>(mats[i]);
      const V3<float> a = v3<float>(&wo[i * 3]), b = v3<float>(&wi[i * 3]), c = v3<float>(&ng[i * 3]), d = v3<float>(&sn[i * 3]);
      const V3<float> f = disney_bsdf_eval(m, a, b, c, d);
      f_out[i * 3] = f.x; f_out[i * 3 + 1] = f.y; f_out[i * 3 + 2] = f.z;
      pdf_out[i] = disney_bsdf_pdf(m, a, b, c, d);
    }
  }
}
// Central differences in double of disney_bsdf_eval with respect to base
// colour r, g, b, roughness and metallic (step h): df[i*15 + k*3 + c].
void orc_bsdf_eval_grad_fd(const phx_material *mats, const float *wo, const float *wi, const float *ng, const float *sn,
                           uint32_t n, double h, double *df) {
  for (uint32_t i = 0; i < n; ++i) {
    const V3<double> a = v3<double>(&wo[i * 3]), b = v3<double>(&wi[i * 3]), c = v3<double>(&ng[i * 3]),
                     d = v3<double>(&sn[i * 3]);
    for (int k = 0; k < 5; ++k) {
      V3<double> f[2];
      for (int sgn = 0; sgn < 2; ++sgn) {
        Mat<double> m = mat_from<double>(mats[i]);
        const double dh = sgn ? -h : h;
        if (k == 0) m.base_color.x += dh;
        if (k == 1) m.base_color.y += dh;
        if (k == 2) m.base_color.z += dh;
        if (k == 3) m.roughness += dh;
        if (k == 4) m.metallic += dh;
        f[sgn] = disney_bsdf_eval(m, a, b, c, d);
      }
      df[i * 15 + k * 3] = (f[0].x - f[1].x) / (2 * h);
      df[i * 15 + k * 3 + 1] = (f[0].y - f[1].y) / (2 * h);
      df[i * 15 + k * 3 + 2] = (f[0].z - f[1].z) / (2 * h);
    }
  }
}
// disney_bsdf_sample with a per-tuple seed: wi (3), f (3), pdf, flags
// (bit 0 specular, bit 1 diffuse).
void orc_bsdf_sample(const phx_material *mats, const float *wo, const float *ng, const float *sn, const uint32_t *seeds,
                     uint32_t n, float *wi_out, float *f_out, float *pdf_out, uint32_t *flags_out) {
  for (uint32_t i = 0; i < n; ++i) {
    const Mat<float> m = mat_from<float>(mats[i]);
    Rng rng(seeds[i]);
    const BsdfSample<float> s =
        disney_bsdf_sample(m, v3<float>(&wo[i * 3]), v3<float>(&ng[i * 3]), v3<float>(&sn[i * 3]), rng);
    wi_out[i * 3] = s.wi.x; wi_out[i * 3 + 1] = s.wi.y; wi_out[i * 3 + 2] = s.wi.z;
    f_out[i * 3] = s.f.x; f_out[i * 3 + 1] = s.f.y; f_out[i * 3 + 2] = s.f.z;
    pdf_out[i] = s.pdf;
    flags_out[i] = (s.is_specular ? 1u : 0u) | (s.is_diffuse ? 2u : 0u);
  }
}

// disney_lobe_weights (disney_bsdf.h:148-175) at NoV: wd, ws, wt, wc.
void orc_lobe_weights(const phx_material *m, double nov, int use_double, double *out) {
  if (use_double) {
    double w[4];
    disney_lobe_weights(mat_from<double>(*m), nov, w[0], w[1], w[2], w[3]);
    std::copy(w, w + 4, out);
  } else {
    float w[4];
    disney_lobe_weights(mat_from<float>(*m), float(nov), w[0], w[1], w[2], w[3]);
    for (int k = 0; k < 4; ++k) out[k] = w[k];
  }
}
// sample_light (light.h:72-148) from point p with Rng(seed): wi, Li, pdf, dist.
void orc_sample_light(const phx_light *l, const float *p, uint32_t seed, float *out) {
  Rng r(seed);
  const LightSample<float> ls = sample_light<float>(*l, v3<float>(p), r);
  const float v[8] = {ls.wi.x, ls.wi.y, ls.wi.z, ls.Li.x, ls.Li.y, ls.Li.z, ls.pdf, ls.dist};
  std::copy(v, v + 8, out);
}
// TextureAtlas::sample (texture.h:25-64) of texture `id` at (u, v).
void orc_texture_sample(const float *pixels, uint32_t pixel_count, const phx_texture_info *infos, uint32_t count, int id,
                        double u, double v, int use_double, double *out) {
  Atlas a;
  a.pixels.assign(pixels, pixels + size_t(pixel_count) * 3);
  a.infos.assign(infos, infos + count);
  if (use_double) {
    const V3<double> c = a.sample<double>(id, u, v);
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
  } else {
    const V3<float> c = a.sample<float>(id, float(u), float(v));
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
  }
}
// SamplerTable::sample (sampler.h:73-138) of sampler `id` for attribute value `attr`.
void orc_sampler_sample(const phx_sampler_info *infos, uint32_t count, const float *texels, uint32_t texel_count, int id,
                        const double *attr, int use_double, double *out) {
  Scene s;
  s.samplers.assign(infos, infos + count);
  if (texels) s.sampler_texels.assign(texels, texels + size_t(texel_count) * 4);
  if (use_double) {
    const V4<double> r = sampler_sample<double>(s, id, V4<double>{attr[0], attr[1], attr[2], attr[3]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
  } else {
    const V4<float> r = sampler_sample<float>(s, id, V4<float>{float(attr[0]), float(attr[1]), float(attr[2]), float(attr[3])});
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
  }
}
// resolve_material and the sourced normal map (pathtracer.cpp:111-247, 966-972)
// at a point of primitive `prim` with barycentrics (b1, b2), position p,
// shading normal sn and texture coordinates uv: the resolved material, its
// alpha and the shading normal after the normal maps (the normal_tex map
// first, then normal_src).
void orc_surface_resolve(void *sp, uint32_t prim, double b1, double b2, const double *p, const double *sn,
                         const double *uv, phx_material *mat_out, double *alpha_out, double *normal_out) {
  const Scene &s = *static_cast<Scene *>(sp);
  Query<double> q;
  q.prim = prim;
  q.mat = s.mesh.mat.empty() ? 0u : s.mesh.mat[prim];
  q.b1 = b1;
  q.b2 = b2;
  q.p = {p[0], p[1], p[2]};
  q.sn = {sn[0], sn[1], sn[2]};
  q.uv = {uv[0], uv[1]};
  double alpha = 1.0;
  const Mat<double> m = resolve_material<double>(s, q, alpha, nullptr);
  V3<double> n = q.sn;
  if (m.normal_tex >= 0 && !s.atlas.infos.empty()) {
    V3<double> nm = s.atlas.sample<double>(m.normal_tex, q.uv.x, q.uv.y);
    nm = nm * 2.0 - V3<double>{1.0, 1.0, 1.0};
    V3<double> ta, bi;
    onb_from_normal(n, ta, bi);
    n = normalize(ta * nm.x + bi * nm.y + n * nm.z);
  }
  if (m.normal_src.kind != PHX_SOURCE_CONSTANT) n = apply_normal_map<double>(s, m.normal_src, q, n);
  phx_material o = s.mats[q.mat];
  for (int k = 0; k < 3; ++k) {
    o.base_color[k] = float((&m.base_color.x)[k]);
    o.emission[k] = float((&m.emission.x)[k]);
  }
  o.metallic = float(m.metallic);
  o.roughness = float(m.roughness);
  o.transmission = float(m.transmission);
  *mat_out = o;
  *alpha_out = alpha;
  normal_out[0] = n.x; normal_out[1] = n.y; normal_out[2] = n.z;
}

void orc_rng(uint32_t pixel, uint32_t sample, uint32_t dim, uint32_t n, uint32_t *u32_out, float *f32_out) {
  Rng r = make_rng(pixel, sample, dim);
  Rng r2 = r;
  for (uint32_t i = 0; i < n; ++i) {
    if (u32_out) u32_out[i] = r.next_u32();
    if (f32_out) f32_out[i] = r2.next_f32();
  }
}
uint32_t orc_hash(uint32_t x) { return hash_u32(x); }
float orc_power_heuristic(float a, float b) { return power_heuristic(1, a, 1, b); }
void orc_onb(const float *n, float *t, float *b) {
  V3<float> tt, bb;
  onb_from_normal(v3<float>(n), tt, bb);
  t[0] = tt.x; t[1] = tt.y; t[2] = tt.z;
  b[0] = bb.x; b[1] = bb.y; b[2] = bb.z;
}
// Environment map: sample (dir, pdf) with seeds, pdf(dir), evaluate(dir).
void orc_env_sample(void *sp, const uint32_t *seeds, uint32_t n, float *dir, float *pdf) {
  Scene *s = static_cast<Scene *>(sp);
  for (uint32_t i = 0; i < n; ++i) {
    Rng r(seeds[i]);
    float p = 0.f;
    const V3<float> d = s->env.sample_direction<float>(r, p);
    dir[i * 3] = d.x; dir[i * 3 + 1] = d.y; dir[i * 3 + 2] = d.z;
    pdf[i] = p;
  }
}
void orc_env_eval(void *sp, const float *dir, uint32_t n, float *pdf, float *rgb) {
  Scene *s = static_cast<Scene *>(sp);
  for (uint32_t i = 0; i < n; ++i) {
    const V3<float> d = v3<float>(&dir[i * 3]);
    if (pdf) pdf[i] = s->env.pdf<float>(d);
    if (rgb) {
      const V3<float> c = s->env.evaluate<float>(d);
      rgb[i * 3] = c.x; rgb[i * 3 + 1] = c.y; rgb[i * 3 + 2] = c.z;
    }
  }
}
void orc_env_cdf(void *sp, float *marginal, float *conditional) {
  Scene *s = static_cast<Scene *>(sp);
  std::copy(s->env.marginal.begin(), s->env.marginal.end(), marginal);
  std::copy(s->env.conditional.begin(), s->env.conditional.end(), conditional);
}
void orc_camera_perspective(phx_camera *out, const float *from, const float *at, const float *up, float vfov, float aspect,
                            float aperture, float focus) {
  *out = make_perspective(v3<float>(from), v3<float>(at), v3<float>(up), vfov, aspect, aperture, focus);
}
uint32_t orc_volume_majorants(phx_volume *v, const float *values, const float *tf, uint32_t tf_size, float *out,
                              uint32_t cap) {
  std::vector<float> m;
  const uint32_t n = build_majorants(*v, values, tf, tf_size, m);
  if (out) std::copy(m.begin(), m.begin() + std::min(n, cap), out);
  return n;
}
void orc_volume_transmittance(void *sp, const float *org, const float *dir, float tmin, float tmax, const uint32_t *seeds,
                              uint32_t n, float *tr) {
  Scene *s = static_cast<Scene *>(sp);
  for (uint32_t i = 0; i < n; ++i) {
    Rng r(seeds[i]);
    tr[i] = s->vol.transmittance<float>(v3<float>(org), v3<float>(dir), tmin, tmax, r);
  }
}
void orc_volume_scatter(void *sp, const float *org, const float *dir, float tmin, float tmax, const uint32_t *seeds,
                        uint32_t n, float *t_out) {
  Scene *s = static_cast<Scene *>(sp);
  for (uint32_t i = 0; i < n; ++i) {
    Rng r(seeds[i]);
    V3<float> alb{};
    t_out[i] = s->vol.sample_scatter<float>(v3<float>(org), v3<float>(dir), tmin, tmax, r, alb);
  }
}

// The detached estimator (detached_pixel) of every pixel in double: the image
// when `image` is not NULL, and the returned sum over pixels and channels of
// delta[p * 3 + c] * image[p * 3 + c] when `delta` is not NULL.
double orc_render_detached(void *sp, const phx_camera *cam, const phx_render_params *rp, const double *mat_params,
                           uint32_t mat_count, const double *texels, const float *delta, double *image, int threads) {
  Scene *s = static_cast<Scene *>(sp);
  Params P{*rp, *cam};
  const DetachedEval E{mat_params, mat_count, texels};
  const uint32_t npix = rp->width * rp->height;
  if (threads > 0) omp_set_num_threads(threads);
  double total = 0.0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
  for (int64_t i = 0; i < int64_t(npix); ++i) {
    const uint32_t idx = uint32_t(i);
    const V3<double> c = detached_pixel(*s, P, idx, E);
    if (image) {
      image[idx * 3] = c.x;
      image[idx * 3 + 1] = c.y;
      image[idx * 3 + 2] = c.z;
    }
    if (delta) total += double(delta[idx * 3]) * c.x + double(delta[idx * 3 + 1]) * c.y + double(delta[idx * 3 + 2]) * c.z;
  }
  return total;
}

// PathTracer::render (pathtracer.cpp:564-1159) on the CPU with `threads`
// OpenMP threads (0: all).  Outputs are W*H rows as rendered (row y = idx / W):
// color rgb, alpha, depth, normal, albedo, the samples each pixel took (any
// may be NULL).  `pixel_begin`,
// `pixel_end` restrict the work to a pixel range (bounded CPU baselines).
// Returns the wall time in seconds.
double orc_render(void *sp, const phx_camera *cam, const phx_render_params *rp, float *color, float *alpha, float *depth,
                  float *normal, float *albedo, uint32_t *nsamp, int threads, int use_double, uint32_t pixel_begin,
                  uint32_t pixel_end) {
  Scene *s = static_cast<Scene *>(sp);
  Params P{*rp, *cam};
  const uint32_t npix = rp->width * rp->height;
  if (pixel_end == 0 || pixel_end > npix) pixel_end = npix;
  if (threads > 0) omp_set_num_threads(threads);
  const double t0 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = int64_t(pixel_begin); i < int64_t(pixel_end); ++i) {
    const uint32_t idx = uint32_t(i);
    float c[3], a, d, nn[3], al[3];
    uint32_t ns;
    if (use_double) {
      PixelOut<double> o;
      render_pixel<double>(*s, P, idx, o, nullptr);
      pixel_aov<double>(*s, P, idx, o);
      const double inv = o.ns ? 1.0 / double(o.ns) : 0.0;
      c[0] = float(o.accum.x * inv); c[1] = float(o.accum.y * inv); c[2] = float(o.accum.z * inv);
      a = float(o.alpha * inv);
      d = float(o.depth);
      nn[0] = float(o.normal.x); nn[1] = float(o.normal.y); nn[2] = float(o.normal.z);
      al[0] = float(o.albedo.x); al[1] = float(o.albedo.y); al[2] = float(o.albedo.z);
      ns = o.ns;
    } else {
      PixelOut<float> o;
      render_pixel<float>(*s, P, idx, o, nullptr);
      pixel_aov<float>(*s, P, idx, o);
      const float inv = o.ns ? 1.f / float(o.ns) : 0.f;
      c[0] = o.accum.x * inv; c[1] = o.accum.y * inv; c[2] = o.accum.z * inv;
      a = o.alpha * inv;
      d = o.depth;
      nn[0] = o.normal.x; nn[1] = o.normal.y; nn[2] = o.normal.z;
      al[0] = o.albedo.x; al[1] = o.albedo.y; al[2] = o.albedo.z;
      ns = o.ns;
    }
    if (color) { color[idx * 3] = c[0]; color[idx * 3 + 1] = c[1]; color[idx * 3 + 2] = c[2]; }
    if (alpha) alpha[idx] = a;
    if (depth) depth[idx] = d;
    if (normal) { normal[idx * 3] = nn[0]; normal[idx * 3 + 1] = nn[1]; normal[idx * 3 + 2] = nn[2]; }
    if (albedo) { albedo[idx * 3] = al[0]; albedo[idx * 3 + 1] = al[1]; albedo[idx * 3 + 2] = al[2]; }
    if (nsamp) nsamp[idx] = ns;
  }
  return omp_get_wtime() - t0;
}

// Display transforms (tone_mapping.h:7-28, color.h:9-12) in photon_render's
// postprocess order (photon_render.cpp:29-35) with its to_u8 (:24-27):
// exposure factor pow(2, e), tone map, clamp01, transfer curve, clamp01.
void orc_display(const float *rgb, uint32_t channels, uint64_t n, const phx_display *t, float *out, uint8_t *out8) {
  const float scale = std::pow(2.f, t->exposure);  // exposure_adjust (tone_mapping.h:17-20)
  auto c01 = [](float x) { return x < 0.f ? 0.f : (x > 1.f ? 1.f : x); };
  for (uint64_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) {
      float c = rgb[i * channels + k] * scale;
      if (t->tonemap == PHX_TONEMAP_ACES) {  // tone_mapping.h:7-12
        const float a = 2.51f, b = 0.03f, cc = 2.43f, d = 0.59f, e = 0.14f;
        const float num = c * (c * a + b);
        const float den = c * (c * cc + d) + e;
        c = num / den;
      } else if (t->tonemap == PHX_TONEMAP_REINHARD) {  // :14-16
        c = c / (1.f + c);
      }
      if (t->clamp) c = c01(c);
      if (t->curve == PHX_CURVE_SRGB)  // gamma_correct, :22-28
        c = c <= 0.0031308f ? 12.92f * c : 1.055f * std::pow(c, 1.f / 2.4f) - 0.055f;
      else if (t->curve == PHX_CURVE_GAMMA22)  // linear_to_srgb, color.h:9-12
        c = std::pow(c, 1.f / 2.2f);
      if (t->clamp) c = c01(c);
      if (out) out[i * 3 + k] = c;
      if (out8) out8[i * 3 + k] = static_cast<uint8_t>(c01(c) * 255.999f);
    }
  }
}

}  // extern "C"
