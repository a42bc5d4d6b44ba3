// Device view of a scene and the light, environment, texture, camera and
// volume routines the shading kernels use.
//
// Semantics: light.h:45-148 (sample_light), environment_map.h:40-153
// (nearest-texel evaluate, marginal/conditional CDF sampling, pdf),
// texture.h:25-64 (bilinear repeat lookup, v flipped), camera.h:143-182,
// tf_volume.h:43-283 (macro-cell delta / ratio tracking), pathtracer.cpp:186-247
// (material resolve with atlas textures) and pathtracer.cpp:270-367
// (one uniformly picked light strategy).
#pragma once

#include "../../include/phx.h"
#include "dev_bsdf.h"
#include "dev_math.h"

namespace phx {

// 8-wide BVH node of 80 bytes (five 16-byte words), child boxes quantised to
// 8 bits in the node's frame: on axis a, child k spans
//   [org[a] + qlo[a][k] * 2^ex[a], org[a] + qhi[a][k] * 2^ex[a]]
// (the builder rounds the quantised bounds outward against the exact value of
// that expression).  16-bit planes in 128-byte nodes measured 7 % slower: the
// 8-bit node is read with five loads instead of eight, and the traversal fits
// six waves per SIMD instead of five.  Children
// [0, ninner) are the inner nodes child_base + k; children [ninner, nchild)
// are leaves whose triangles follow one another from tri_base in slot order,
// slot k holding ((leaf_cnt >> 2k) & 3) + 1 triangles.
constexpr int kNodeWidth = 8;
struct alignas(16) BvhNode {
  float org[3];
  int8_t ex[3];
  uint8_t ninner;
  uint32_t child_base;
  uint32_t tri_base;
  uint16_t leaf_cnt;
  uint8_t nchild;
  uint8_t pad0;
  uint32_t pad1;
  uint8_t qlo[3][kNodeWidth];  // bytes 32..55: 8 bytes per axis
  uint8_t qhi[3][kNodeWidth];  // bytes 56..79
};
static_assert(sizeof(BvhNode) == 80, "BvhNode is five 16-byte words");
// Stack entry of a leaf: ~((first triangle << 3) | (count - 1)); inner nodes
// are pushed as their index (>= 0).
PHX_HD int leaf_code(uint32_t first, uint32_t count) { return int(~((first << 3) | (count - 1u))); }

// Buckets of the environment CDF search guides (a power of two).  Each search
// step is a dependent load of ~1 us, and an environment sample makes two
// searches: with 8192 buckets most searches end in 0-1 steps (the per-row
// conditional guides take 4 * 8193 B per row, 16.8 MB for a 1024x512 map).
// On the 737K-triangle frame: 128 buckets 47.5 ms, 2048 46.7 ms, 8192 46.3 ms.
constexpr uint32_t kEnvGuide = 8192;
// Buckets of the emissive-triangle CDF search guide (a power of two).
constexpr uint32_t kEmGuide = 1024;

// Triangles in leaf order: v0, e1 = v1 - v0, e2 = v2 - v0 (the Moller-Trumbore
// edge vectors, computed with the same float subtraction the reference does)
// and the primitive id in v0's w.
struct alignas(16) TriRec {
  float4 v0;
  float4 e1;
  float4 e2;
};

struct DevScene {
  const BvhNode *nodes;
  const TriRec *tris;
  // In leaf order beside tris: the triangle's vertex indices and material id,
  // so a hit's attributes load one dependent step after the hit, not two.
  const uint4 *tri_attr;
  const uint32_t *indices;
  const float *positions;
  const float *normals;
  const float *texcoords;
  const float *colors;
  const uint32_t *material_ids;
  const phx_material *materials;
  const phx_light *lights;
  const uint32_t *em_ids;
  const float *em_cdf;
  const uint32_t *em_guide;  // kEmGuide + 1 search bounds of em_cdf
  const float4 *em_recs;     // per em_cdf entry: the emitter record of em_record()
  const float *prim_light_pdf;
  const float *env_pixels;
  const float *env_marginal;
  const float *env_cond;
  const uint32_t *env_guide_m;  // kEnvGuide + 1 search bounds of the marginal CDF
  const uint32_t *env_guide_c;  // per row: kEnvGuide + 1 bounds of its conditional CDF
  const float *tex_pixels;
  const phx_texture_info *tex_infos;
  const phx_volume *volumes;
  const float *vol_values;
  const float4 *vol_tf;
  const float *majorants;
  uint32_t material_count;
  uint32_t light_count;
  uint32_t em_count;
  uint32_t env_w, env_h;
  uint32_t tex_count;
  uint32_t vol_count;
  uint32_t vol_tf_count;  // float4 entries of vol_tf, all volumes together
  uint32_t strategy_count;
  uint32_t triangle_count;
  int32_t has_env, env_visible;
  float ray_eps;
  float w2t[9];
  float t2w[9];
};

// ---- lights (light.h:56-148) ------------------------------------------------
struct LightSample {
  vec3 wi, Li;
  float pdf, dist;
};
PHX_DEV vec3 sphere_light_point(const phx_light &l, vec3 p, Rng &rng) {
  const vec3 pos = ld3(l.position);
  if (l.radius <= 0.f) return pos;
  vec3 t, b;
  onb(normalize(pos - p), t, b);
  const vec2 d = concentric_disk(rng);
  return pos + (t * d.x + b * d.y) * l.radius;
}
PHX_DEV LightSample sample_light(const phx_light &l, vec3 p, Rng &rng) {
  const vec3 color = ld3(l.color);
  const LightSample none{{0.f, 1.f, 0.f}, {0.f, 0.f, 0.f}, 0.f, 0.f};
  if (l.type == PHX_LIGHT_POINT || l.type == PHX_LIGHT_SPOT) {
    const vec3 ct = ld3(l.position) - p;
    const float cd2 = dot(ct, ct);
    const vec3 to = sphere_light_point(l, p, rng) - p;
    const float dist = length(to);
    const vec3 wi = dist > 0.f ? to * (1.f / dist) : vec3{0.f, 1.f, 0.f};
    const float inv = cd2 > 0.f ? 1.f / cd2 : 0.f;
    if (l.type == PHX_LIGHT_POINT) return {wi, color * (l.intensity * inv), 1.f, dist};
    const vec3 sd = normalize(ld3(l.direction));
    const float ca = dot(-normalize(ct), sd);
    const float ci = cosf(fmaxf(l.spot_angle - l.spot_falloff, 0.f));
    const float co = cosf(l.spot_angle);
    float fall;
    if (ca >= ci) {
      fall = 1.f;
    } else if (ca <= co) {
      fall = 0.f;
    } else {
      float t = (ca - co) / (ci - co);
      t = t < 0.f ? 0.f : (t > 1.f ? 1.f : t);
      fall = t * t * (3.f - 2.f * t);
    }
    return {wi, color * (l.intensity * inv * fall), 1.f, dist};
  }
  if (l.type == PHX_LIGHT_DIRECTIONAL) {
    vec3 wi = -normalize(ld3(l.direction));
    if (l.radius > 0.f) {
      const float cmax = cosf(l.radius);
      const float ct = 1.f - rng.next() * (1.f - cmax);
      const float st = sqrtf(fmaxf(0.f, 1.f - ct * ct));
      const float phi = kTwoPi * rng.next();
      vec3 t, b;
      onb(wi, t, b);
      wi = normalize(t * (st * cosf(phi)) + b * (st * sinf(phi)) + wi * ct);
    }
    return {wi, color * l.intensity, 1.f, 1e30f};
  }
  if (l.type == PHX_LIGHT_AREA) {
    const float u = rng.next(), v = rng.next();
    const vec3 e1 = ld3(l.edge1), e2 = ld3(l.edge2);
    const vec3 sp = ld3(l.position) + e1 * u + e2 * v;
    const vec3 ln = normalize(cross(e1, e2));
    const vec3 to = sp - p;
    const float dist = length(to);
    if (dist < 1e-6f) return none;
    const vec3 wi = to * (1.f / dist);
    const float cl = -dot(wi, ln);
    if (cl <= 0.f) return none;
    const float area = l.area > 0.f ? l.area : length(cross(e1, e2));
    return {wi, color * l.intensity, (1.f / area) * (dist * dist) / cl, dist * 0.999f};
  }
  return none;
}

// ---- environment map ---------------------------------------------------------
PHX_DEV vec3 mat3mul(const float *m, vec3 d) {
  return {m[0] * d.x + m[1] * d.y + m[2] * d.z, m[3] * d.x + m[4] * d.y + m[5] * d.z,
          m[6] * d.x + m[7] * d.y + m[8] * d.z};
}
PHX_DEV vec2 dir_to_uv(vec3 d) {
  const vec3 nd = normalize(d);
  const float phi = atan2f(nd.z, nd.x);
  const float y = fminf(fmaxf(nd.y, -1.f), 1.f);
  const float theta = acosf(y);
  return {(phi + kPi) * kInvTwoPi, theta * kInvPi};
}
PHX_DEV vec3 uv_to_dir(vec2 uv) {
  const float phi = uv.x * kTwoPi - kPi;
  const float theta = uv.y * kPi;
  const float st = __sinf(theta);
  return {st * __cosf(phi), __cosf(theta), st * __sinf(phi)};
}
PHX_DEV void env_texel(const DevScene &s, vec3 dir, uint32_t &x, uint32_t &y, vec2 &uv) {
  uv = dir_to_uv(mat3mul(s.w2t, dir));
  const float xf = uv.x * float(s.env_w), yf = uv.y * float(s.env_h);
  x = uint32_t(fminf(float(s.env_w - 1), fmaxf(0.f, xf)));
  y = uint32_t(fminf(float(s.env_h - 1), fmaxf(0.f, yf)));
}
PHX_DEV vec3 env_eval(const DevScene &s, vec3 dir) {
  uint32_t x, y;
  vec2 uv;
  env_texel(s, dir, x, y, uv);
  return ld3(s.env_pixels + (size_t(y) * s.env_w + x) * 3);
}
PHX_DEV float env_pdf(const DevScene &s, vec3 dir) {
  uint32_t x, y;
  vec2 uv;
  env_texel(s, dir, x, y, uv);
  const float rp = s.env_marginal[y + 1] - s.env_marginal[y];
  const float *c = s.env_cond + size_t(y) * (s.env_w + 1);
  const float cp = c[x + 1] - c[x];
  const float st = sinf(uv.y * kPi);
  return (rp * cp) * float(s.env_w) * float(s.env_h) / (kTwoPi * kPi * fmaxf(st, 1e-6f));
}
// Largest i in [0, n) with cdf[i] <= u (cdf[0] = 0), the search of
// environment_map.cpp's sampling.  guide[b] is that index for u = b / kEnvGuide,
// so for u in [b, b + 1) / kEnvGuide the answer lies in [guide[b],
// guide[b + 1]]: the search starts there and returns the same index in a few
// steps instead of log2(n).  kEnvGuide is a power of two: u * kEnvGuide and
// b / kEnvGuide are exact in float.
PHX_DEV uint32_t cdf_search(const float *cdf, uint32_t n, const uint32_t *guide, float u) {
  const uint32_t b = min(uint32_t(u * float(kEnvGuide)), kEnvGuide - 1u);
  uint32_t lo = guide[b], hi = min(guide[b + 1] + 1u, n);
  while (lo + 1 < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (cdf[mid] <= u) lo = mid;
    else hi = mid;
  }
  return lo;
}
// Marginal then conditional CDF search, each over a table with a leading 0
// (environment_map.h:56-115).  Also returns the sampled texel (col, row).
PHX_DEV vec3 env_sample(const DevScene &s, Rng &rng, float &pdf, uint32_t &col, uint32_t &row) {
  const float u1 = rng.next(), u2 = rng.next();
  row = cdf_search(s.env_marginal, s.env_h, s.env_guide_m, u2);
  const float *c = s.env_cond + size_t(row) * (s.env_w + 1);
  col = cdf_search(c, s.env_w, s.env_guide_c + size_t(row) * (kEnvGuide + 1), u1);
  const float rp = s.env_marginal[row + 1] - s.env_marginal[row];
  const float cp = c[col + 1] - c[col];
  const float u = (float(col) + rng.next()) / float(s.env_w);
  const float v = (float(row) + rng.next()) / float(s.env_h);
  const vec3 dir = mat3mul(s.t2w, uv_to_dir(vec2{u, v}));
  const float st = __sinf(v * kPi);
  pdf = (rp * cp) * float(s.env_w) * float(s.env_h) / (kTwoPi * kPi * fmaxf(st, 1e-6f));
  return dir;
}

// ---- textures (texture.h:25-64) ---------------------------------------------
PHX_DEV vec3 tex_sample(const DevScene &s, int id, float u, float v) {
  if (id < 0 || uint32_t(id) >= s.tex_count) return {1.f, 1.f, 1.f};
  const phx_texture_info ti = s.tex_infos[id];
  if (!ti.width || !ti.height) return {1.f, 1.f, 1.f};
  u = u - floorf(u);
  v = 1.f - (v - floorf(v));
  const float fx = u * float(ti.width) - 0.5f, fy = v * float(ti.height) - 0.5f;
  const int ix = int(floorf(fx)), iy = int(floorf(fy));
  const float tx = fx - floorf(fx), ty = fy - floorf(fy);
  const int W = int(ti.width), H = int(ti.height);
  const uint32_t x0 = uint32_t(((ix % W) + W) % W), x1 = uint32_t((((ix + 1) % W) + W) % W);
  const uint32_t y0 = uint32_t(((iy % H) + H) % H), y1 = uint32_t((((iy + 1) % H) + H) % H);
  const float *px = s.tex_pixels + size_t(ti.offset) * 3;
  const vec3 c00 = ld3(px + (size_t(y0) * ti.width + x0) * 3), c10 = ld3(px + (size_t(y0) * ti.width + x1) * 3);
  const vec3 c01 = ld3(px + (size_t(y1) * ti.width + x0) * 3), c11 = ld3(px + (size_t(y1) * ti.width + x1) * 3);
  const vec3 top = c00 * (1.f - tx) + c10 * tx;
  const vec3 bot = c01 * (1.f - tx) + c11 * tx;
  return top * (1.f - ty) + bot * ty;
}

// The four atlas texels tex_sample blends at (u, v) and their weights.
struct TexTaps {
  uint32_t idx[4];  // atlas pixel indices: (x0,y0), (x1,y0), (x0,y1), (x1,y1)
  float w[4];
  bool valid;
};
PHX_DEV TexTaps tex_taps(const DevScene &s, int id, float u, float v) {
  TexTaps t{};
  t.valid = false;
  if (id < 0 || uint32_t(id) >= s.tex_count) return t;
  const phx_texture_info ti = s.tex_infos[id];
  if (!ti.width || !ti.height) return t;
  u = u - floorf(u);
  v = 1.f - (v - floorf(v));
  const float fx = u * float(ti.width) - 0.5f, fy = v * float(ti.height) - 0.5f;
  const int ix = int(floorf(fx)), iy = int(floorf(fy));
  const float tx = fx - floorf(fx), ty = fy - floorf(fy);
  const int W = int(ti.width), H = int(ti.height);
  const uint32_t x0 = uint32_t(((ix % W) + W) % W), x1 = uint32_t((((ix + 1) % W) + W) % W);
  const uint32_t y0 = uint32_t(((iy % H) + H) % H), y1 = uint32_t((((iy + 1) % H) + H) % H);
  t.idx[0] = ti.offset + y0 * ti.width + x0;
  t.idx[1] = ti.offset + y0 * ti.width + x1;
  t.idx[2] = ti.offset + y1 * ti.width + x0;
  t.idx[3] = ti.offset + y1 * ti.width + x1;
  t.w[0] = (1.f - tx) * (1.f - ty);
  t.w[1] = tx * (1.f - ty);
  t.w[2] = (1.f - tx) * ty;
  t.w[3] = tx * ty;
  t.valid = true;
  return t;
}

// Material at a hit with atlas textures applied (pathtracer.cpp:186-247).
// TEX = false leaves the texture lookups out of the code (for scenes without
// textures, where they never run).
template <bool TEX = true>
PHX_DEV phx_material resolve_material(const DevScene &s, uint32_t mid, vec2 uv, bool has_color, vec3 color,
                                      float &alpha) {
  phx_material m = s.materials[mid];
  const bool tex = TEX && s.tex_count > 0;
  alpha = m.alpha;
  if (m.alpha_tex >= 0 && tex) alpha = tex_sample(s, m.alpha_tex, uv.x, uv.y).x;
  if (m.base_color_tex >= 0 && tex) {
    const vec3 c = tex_sample(s, m.base_color_tex, uv.x, uv.y);
    m.base_color[0] = c.x;
    m.base_color[1] = c.y;
    m.base_color[2] = c.z;
  } else if (has_color) {
    m.base_color[0] = color.x;
    m.base_color[1] = color.y;
    m.base_color[2] = color.z;
  }
  if (m.roughness_tex >= 0 && tex) m.roughness = tex_sample(s, m.roughness_tex, uv.x, uv.y).x;
  if (m.metallic_tex >= 0 && tex) m.metallic = tex_sample(s, m.metallic_tex, uv.x, uv.y).x;
  m.metallic = clamp01(m.metallic);
  m.roughness = clamp01(m.roughness);
  m.transmission = clamp01(m.transmission);
  if (m.alpha_mode == PHX_ALPHA_OPAQUE) alpha = 1.f;
  else if (m.alpha_mode == PHX_ALPHA_MASK) alpha = alpha >= m.alpha_cutoff ? 1.f : 0.f;
  else alpha = clamp01(alpha);
  return m;
}
PHX_DEV vec3 material_emission(const phx_material &m) { return ld3(m.emission) * m.emission_strength; }
PHX_DEV bool material_may_emit(const phx_material &m) {
  if (m.emission_strength <= 0.f) return false;
  return m.emission_tex >= 0 || m.emission[0] > 0.f || m.emission[1] > 0.f || m.emission[2] > 0.f;
}
PHX_DEV bool emits_on_side(const phx_material &m, bool front) {
  if (m.flags & PHX_MATERIAL_FLAG_EMIT_BOTH_SIDES) return true;
  const bool back_only = (m.flags & PHX_MATERIAL_FLAG_EMIT_BACK_ONLY) != 0;
  return front ? !back_only : back_only;
}

// The emitter record of emissive-CDF entry i (four float4): the triangle's
// vertices with its area pdf (pathtracer.cpp:411-503) and the emitter's
// radiance, emission times alpha of its material resolved at uv (0, 0), with
// bit 0 / bit 1 of r3.y set when it emits toward its front / back side.
PHX_DEV void em_record(const DevScene &s, uint32_t i, float4 *out) {
  const uint32_t prim = s.em_ids[i];
  const float pa = s.prim_light_pdf[prim];
  const vec3 v0 = ld3(s.positions + size_t(s.indices[prim * 3]) * 3);
  const vec3 v1 = ld3(s.positions + size_t(s.indices[prim * 3 + 1]) * 3);
  const vec3 v2 = ld3(s.positions + size_t(s.indices[prim * 3 + 2]) * 3);
  float alpha = 1.f;
  const uint32_t mid = s.material_ids ? s.material_ids[prim] : 0u;
  const phx_material em = resolve_material(s, mid, vec2{0.f, 0.f}, false, vec3{0.f, 0.f, 0.f}, alpha);
  const vec3 L = material_emission(em) * alpha;
  const uint32_t sides = (emits_on_side(em, true) ? 1u : 0u) | (emits_on_side(em, false) ? 2u : 0u);
  out[0] = make_float4(v0.x, v0.y, v0.z, pa);
  out[1] = make_float4(v1.x, v1.y, v1.z, L.x);
  out[2] = make_float4(v2.x, v2.y, v2.z, L.y);
  out[3] = make_float4(L.z, __uint_as_float(sides), 0.f, 0.f);
}

// ---- light strategy pick (pathtracer.cpp:270-367) ---------------------------
struct Direct {
  vec3 wi, L;
  float pdf, dist;
  bool mis, valid;
};
// LIGHTS = false leaves the analytic lights out of the code (for scenes
// without them, where that branch never runs).
template <bool LIGHTS = true>
PHX_DEV Direct sample_direct(const DevScene &s, vec3 p, Rng &rng) {
  Direct ds{{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, 0.f, 0.f, false, false};
  const uint32_t nstrat = s.strategy_count;
  if (nstrat == 0) return ds;
  const float pick_pdf = 1.f / float(nstrat);
  uint32_t pick = uint32_t(rng.next() * float(nstrat));
  if (pick >= nstrat) pick = nstrat - 1;
  if (LIGHTS && pick < s.light_count) {
    const LightSample ls = sample_light(s.lights[pick], p, rng);
    if (!(ls.pdf > 0.f)) return ds;
    ds.wi = ls.wi;
    ds.L = ls.Li;
    ds.pdf = ls.pdf * pick_pdf;
    ds.dist = ls.dist;
    ds.valid = true;
    return ds;
  }
  if (s.em_count > 0 && pick == s.light_count) {
    // The search of pathtracer.cpp:287-295 (first entry above xi, at most
    // em_count - 1), started from the guide's bounds for xi's bucket; then one
    // record holds the triangle, its pdf and the emitter's radiance.
    const float xi = rng.next();
    const uint32_t bucket = min(uint32_t(xi * float(kEmGuide)), kEmGuide - 1u);
    uint32_t lo = s.em_guide[bucket], hi = s.em_guide[bucket + 1];
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s.em_cdf[mid] <= xi) lo = mid + 1;
      else hi = mid;
    }
    const float4 *rec = s.em_recs + size_t(lo) * 4;
    const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3];
    const float pa = r0.w;
    if (!(pa > 0.f)) return ds;
    const vec3 v0{r0.x, r0.y, r0.z}, v1{r1.x, r1.y, r1.z}, v2{r2.x, r2.y, r2.z};
    const float su = sqrtf(rng.next());
    const float r2u = rng.next();
    const float b0 = 1.f - su, b1 = r2u * su, b2 = 1.f - b0 - b1;
    const vec3 lp = v0 * b0 + v1 * b1 + v2 * b2;
    const vec3 to = lp - p;
    const float d2 = dot(to, to);
    if (!(d2 > 0.f)) return ds;
    const float dist = sqrtf(d2);
    const vec3 wi = to * (1.f / dist);
    const vec3 ln = normalize(cross(v1 - v0, v2 - v0));
    const float cl = -dot(wi, ln);
    if (fabsf(cl) < 1e-6f) return ds;
    if (!(__float_as_uint(r3.y) & (cl > 0.f ? 1u : 2u))) return ds;
    ds.wi = wi;
    ds.L = vec3{r1.w, r2.w, r3.x};
    ds.pdf = pa * d2 / fabsf(cl) * pick_pdf;
    ds.dist = dist;
    ds.mis = true;
    ds.valid = true;
    return ds;
  }
  if (s.has_env) {
    float ep = 0.f;
    uint32_t col, row;
    const vec3 wi = env_sample(s, rng, ep, col, row);
    // The radiance is the texel the direction maps back to, as env_eval does;
    // that is the sampled texel but at texel edges, so it is loaded beside the
    // direction's arithmetic and only reloaded when the two differ.
    const vec3 Ls = ld3(s.env_pixels + (size_t(row) * s.env_w + col) * 3);
    if (!(ep > 0.f)) return ds;
    uint32_t x, y;
    vec2 uv;
    env_texel(s, wi, x, y, uv);
    ds.wi = wi;
    ds.L = (x == col && y == row) ? Ls : ld3(s.env_pixels + (size_t(y) * s.env_w + x) * 3);
    ds.pdf = ep * pick_pdf;
    ds.dist = 1e30f;
    ds.mis = true;
    ds.valid = true;
  }
  return ds;
}

// ---- camera (camera.h:143-182) ----------------------------------------------
PHX_DEV void clip_range(const phx_camera &c, vec3 dir, float &tn, float &tf) {
  tn = 0.f;
  tf = 1e30f;
  if (c.clip_near <= 0.f && c.clip_far >= 1e30f) return;
  const float ca = dot(dir, ld3(c.w) * -1.f);
  if (ca <= 1e-6f) {
    tn = 1.f;
    tf = 0.f;
    return;
  }
  if (c.clip_near > 0.f) tn = c.clip_near / ca;
  if (c.clip_far < 1e30f) tf = c.clip_far / ca;
}
PHX_DEV void camera_ray(const phx_camera &c, float s, float t, Rng &rng, vec3 &org, vec3 &dir) {
  const vec3 ll = ld3(c.lower_left), hor = ld3(c.horizontal), ver = ld3(c.vertical), o = ld3(c.origin);
  if (c.is_ortho) {
    org = ll + hor * s + ver * t;
    dir = -ld3(c.w);
    return;
  }
  if (c.lens_radius <= 0.f) {
    org = o;
    dir = normalize(ll + hor * s + ver * t - o);
    return;
  }
  const vec2 d = concentric_disk(rng);
  const vec2 rd{d.x * c.lens_radius, d.y * c.lens_radius};
  const vec3 off = ld3(c.u) * rd.x + ld3(c.v) * rd.y;
  org = o + off;
  dir = normalize(ll + hor * s + ver * t - o - off);
}

// ---- volumes (tf_volume.h:43-283) --------------------------------------------
PHX_DEV vec3 to_local_point(const phx_volume &v, vec3 p) {
  const float *m = v.world_to_local;
  return {m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
          m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
}
PHX_DEV vec3 to_local_vector(const phx_volume &v, vec3 d) {
  const float *m = v.world_to_local;
  return {m[0] * d.x + m[1] * d.y + m[2] * d.z, m[4] * d.x + m[5] * d.y + m[6] * d.z,
          m[8] * d.x + m[9] * d.y + m[10] * d.z};
}
PHX_DEV bool vol_clip(const phx_volume &v, vec3 org, vec3 dir, float &t0, float &t1) {
  const float o[3] = {org.x, org.y, org.z}, d[3] = {dir.x, dir.y, dir.z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float lo = v.origin[a];
    const float hi = lo + v.spacing[a] * float(v.dims[a] > 1 ? v.dims[a] - 1 : 1);
    if (d[a] == 0.f) {
      if (o[a] < lo || o[a] > hi) return false;
      continue;
    }
    const float inv = 1.f / d[a];
    const float ta = (lo - o[a]) * inv, tb = (hi - o[a]) * inv;
    const float tn = ta < tb ? ta : tb, tf = ta < tb ? tb : ta;
    t0 = tn > t0 ? tn : t0;
    t1 = tf < t1 ? tf : t1;
    if (t1 <= t0) return false;
  }
  return true;
}
PHX_DEV float vol_field(const DevScene &s, const phx_volume &v, vec3 p) {
  const float g[3] = {(p.x - v.origin[0]) / v.spacing[0], (p.y - v.origin[1]) / v.spacing[1],
                      (p.z - v.origin[2]) / v.spacing[2]};
  int i0[3], i1[3];
  float fr[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int n = int(v.dims[a]);
    float x = g[a];
    x = x < 0.f ? 0.f : (x > float(n - 1) ? float(n - 1) : x);
    if (v.linear) {
      int i = int(x);
      if (i > n - 2) i = n - 2;
      if (i < 0) i = 0;
      i0[a] = i;
      i1[a] = i + 1 < n ? i + 1 : i;
      fr[a] = x - float(i);
    } else {
      int i = int(x + 0.5f);
      i = i > n - 1 ? n - 1 : i;
      i0[a] = i1[a] = i;
      fr[a] = 0.f;
    }
  }
  const size_t sx = v.dims[0], sxy = size_t(v.dims[0]) * v.dims[1];
  const float *base = s.vol_values + v.data_offset;
  auto at = [&](int x, int y, int z) { return base[size_t(z) * sxy + size_t(y) * sx + size_t(x)]; };
  if (!v.linear) return at(i0[0], i0[1], i0[2]);
  const float c00 = at(i0[0], i0[1], i0[2]) * (1.f - fr[0]) + at(i1[0], i0[1], i0[2]) * fr[0];
  const float c10 = at(i0[0], i1[1], i0[2]) * (1.f - fr[0]) + at(i1[0], i1[1], i0[2]) * fr[0];
  const float c01 = at(i0[0], i0[1], i1[2]) * (1.f - fr[0]) + at(i1[0], i0[1], i1[2]) * fr[0];
  const float c11 = at(i0[0], i1[1], i1[2]) * (1.f - fr[0]) + at(i1[0], i1[1], i1[2]) * fr[0];
  const float c0 = c00 * (1.f - fr[1]) + c10 * fr[1];
  const float c1 = c01 * (1.f - fr[1]) + c11 * fr[1];
  return c0 * (1.f - fr[2]) + c1 * fr[2];
}
PHX_DEV float4 vol_classify(const DevScene &s, const phx_volume &v, vec3 p) {
  const float value = vol_field(s, v, p);
  const float range = v.value_hi - v.value_lo;
  float x = range != 0.f ? (value - v.value_lo) / range : 0.f;
  x = x < 0.f ? 0.f : (x > 1.f ? 1.f : x);
  const float f = x * float(v.tf_size - 1);
  uint32_t i = uint32_t(f);
  if (i > v.tf_size - 2) i = v.tf_size >= 2 ? v.tf_size - 2 : 0;
  const float fr = v.tf_size >= 2 ? f - float(i) : 0.f;
  const float4 a = s.vol_tf[v.tf_offset + i];
  const float4 b = s.vol_tf[v.tf_offset + (v.tf_size >= 2 ? i + 1 : i)];
  float4 c;
  c.x = a.x * (1.f - fr) + b.x * fr;
  c.y = a.y * (1.f - fr) + b.y * fr;
  c.z = a.z * (1.f - fr) + b.z * fr;
  c.w = (a.w * (1.f - fr) + b.w * fr) * v.density_scale;
  return c;
}
PHX_DEV float vol_track(const DevScene &s, const phx_volume &v, vec3 org, vec3 dir, float t0, float t1, Rng &rng,
                        bool scatter, vec3 &albedo, float &tr) {
  const float o[3] = {org.x, org.y, org.z}, d[3] = {dir.x, dir.y, dir.z};
  int cell[3], step[3];
  float t_next[3], t_delta[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int m = int(v.macro_dims[a]);
    const float extent = v.spacing[a] * float(v.dims[a] > 1 ? v.dims[a] - 1 : 1);
    const float cpu = float(m) / extent;
    const float g = (o[a] + d[a] * t0 - v.origin[a]) * cpu;
    int c = int(g);
    c = c < 0 ? 0 : (c > m - 1 ? m - 1 : c);
    cell[a] = c;
    const float speed = d[a] * cpu;
    if (speed > 0.f) {
      step[a] = 1;
      t_delta[a] = 1.f / speed;
      t_next[a] = t0 + (float(c + 1) - g) * t_delta[a];
    } else if (speed < 0.f) {
      step[a] = -1;
      t_delta[a] = -1.f / speed;
      t_next[a] = t0 + (g - float(c)) * t_delta[a];
    } else {
      step[a] = 0;
      t_delta[a] = 1e30f;
      t_next[a] = 1e30f;
    }
  }
  float t = t0;
  int lookups = 0;
  for (int visited = 0; visited < 4096; ++visited) {
    const int axis = t_next[0] < t_next[1] ? (t_next[0] < t_next[2] ? 0 : 2) : (t_next[1] < t_next[2] ? 1 : 2);
    const float tna = axis == 0 ? t_next[0] : (axis == 1 ? t_next[1] : t_next[2]);
    const float t_exit = tna < t1 ? tna : t1;
    const float bound =
        s.majorants[v.macro_offset + (uint32_t(cell[2]) * v.macro_dims[1] + uint32_t(cell[1])) * v.macro_dims[0] +
                    uint32_t(cell[0])];
    if (bound > 0.f) {
      const float inv_bound = 1.f / bound;
      while (lookups < 8192) {
        t -= logf(1.f - rng.next()) * inv_bound;
        if (t >= t_exit) break;
        ++lookups;
        const float4 c = vol_classify(s, v, org + dir * t);
        if (scatter) {
          if (rng.next() * bound < c.w) {
            albedo = {c.x, c.y, c.z};
            return t;
          }
        } else {
          tr *= 1.f - c.w * inv_bound;
          if (tr < 0.05f) {
            if (!(tr > 0.f) || rng.next() * 0.05f >= tr) {
              tr = 0.f;
              return -1.f;
            }
            tr = 0.05f;
          }
        }
      }
    }
    t = t_exit;
    if (t_exit >= t1 || lookups >= 8192) break;
    // the step along `axis`, with selects: an index computed at run time into
    // these arrays (or into v.macro_dims) would put them in scratch memory
    const bool a0 = axis == 0, a1 = axis == 1;
    const int ca = (a0 ? cell[0] : (a1 ? cell[1] : cell[2])) + (a0 ? step[0] : (a1 ? step[1] : step[2]));
    const int ma = int(a0 ? v.macro_dims[0] : (a1 ? v.macro_dims[1] : v.macro_dims[2]));
    if (ca < 0 || ca >= ma) break;
    if (a0) {
      cell[0] = ca;
      t_next[0] += t_delta[0];
    } else if (a1) {
      cell[1] = ca;
      t_next[1] += t_delta[1];
    } else {
      cell[2] = ca;
      t_next[2] += t_delta[2];
    }
  }
  return -1.f;
}
PHX_DEV float vol_sample_scatter(const DevScene &s, vec3 ro, vec3 rd, float tmin, float tmax, Rng &rng, vec3 &albedo) {
  float t_event = -1.f, t_limit = tmax, unused = 1.f;
  for (uint32_t i = 0; i < s.vol_count; ++i) {
    const phx_volume v = s.volumes[i];
    if (v.majorant <= 0.f || v.tf_size == 0) continue;
    const vec3 org = to_local_point(v, ro), dir = to_local_vector(v, rd);
    float t0 = tmin, t1 = t_limit;
    if (!vol_clip(v, org, dir, t0, t1)) continue;
    vec3 ea{0.f, 0.f, 0.f};
    const float t = vol_track(s, v, org, dir, t0, t1, rng, true, ea, unused);
    if (t >= 0.f) {
      t_event = t;
      t_limit = t;
      albedo = ea;
    }
  }
  return t_event;
}
PHX_DEV float vol_transmittance(const DevScene &s, vec3 ro, vec3 rd, float tmin, float tmax, Rng &rng) {
  float tr = 1.f;
  vec3 unused{0.f, 0.f, 0.f};
  for (uint32_t i = 0; i < s.vol_count; ++i) {
    const phx_volume v = s.volumes[i];
    if (v.majorant <= 0.f || v.tf_size == 0) continue;
    const vec3 org = to_local_point(v, ro), dir = to_local_vector(v, rd);
    float t0 = tmin, t1 = tmax;
    if (!vol_clip(v, org, dir, t0, t1)) continue;
    vol_track(s, v, org, dir, t0, t1, rng, false, unused, tr);
    if (tr <= 0.f) return 0.f;
  }
  return tr;
}

}  // namespace phx
