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
// The unflipped geometric normal of a triangle record (mesh_intersector.h:139).
PHX_DEV vec3 tri_normal(const TriRec &r) {
  return normalize(cross(vec3{r.e1.x, r.e1.y, r.e1.z}, vec3{r.e2.x, r.e2.y, r.e2.z}));
}

struct DevScene {
  const BvhNode *nodes;
  const TriRec *tris;
  // In leaf order beside tris: the triangle's vertex indices and material id,
  // so a hit's attributes load one dependent step after the hit, not two.
  const uint4 *tri_attr;
  // In leaf order beside tris: the geometric normal normalize(cross(e1, e2))
  // and the primitive id bits, so shading reads 16 bytes instead of the 48-byte
  // record and skips the normalisation (null: computed from tris).
  const float4 *tri_ng;
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
  const float4 *attributes;  // per-vertex RGBA records of the attribute slots
  const phx_sampler_info *samplers;
  const float4 *sampler_texels;
  int32_t attr_offset[PHX_ATTRIBUTE_MESH_COUNT];  // first record of each slot, -1: absent
  uint32_t sampler_count;
  uint32_t has_sources;  // a material takes a parameter from an attribute or a sampler
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

// A point of the scene mesh as shading sees it (SurfaceQuery,
// pathtracer.cpp:95-108): primitive and material, barycentrics (b1, b2) of
// the second and third vertex, position, geometric normal, the shading normal
// (interpolated, flat where degenerate, on the geometric normal's side,
// :906-910), texture coordinates and vertex colour.
struct SurfaceHit {
  vec3 p, ng, sn;
  vec2 uv;
  vec3 color;
  bool has_color;
  uint32_t prim, mid;
  float b1, b2;
};

// ---- attributes and samplers (sampler.h:25-139, pathtracer.cpp:111-132) ------
PHX_DEV float4 make_rgba(vec3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
// Attribute `slot` at the point: the world position or shading normal, or the
// per-vertex records of a mesh slot blended by the barycentrics
// (a0 * (1 - b1 - b2) + a1 * b1 + a2 * b2); (0, 0, 0, 1) for an absent slot.
PHX_DEV float4 fetch_attribute(const DevScene &s, int slot, const SurfaceHit &q) {
  if (slot == PHX_ATTRIBUTE_WORLD_POSITION) return make_rgba(q.p, 1.f);
  if (slot == PHX_ATTRIBUTE_WORLD_NORMAL) return make_rgba(q.sn, 1.f);
  // the slot's first record by selects: an index computed at run time into the
  // scene's array would put the scene in scratch memory
  const int off = slot == 0 ? s.attr_offset[0]
                : slot == 1 ? s.attr_offset[1]
                : slot == 2 ? s.attr_offset[2]
                : slot == 3 ? s.attr_offset[3]
                : slot == 4 ? s.attr_offset[4] : -1;
  if (off < 0) return make_float4(0.f, 0.f, 0.f, 1.f);
  const float4 *rec = s.attributes + off;
  const float4 a0 = rec[s.indices[q.prim * 3]], a1 = rec[s.indices[q.prim * 3 + 1]], a2 = rec[s.indices[q.prim * 3 + 2]];
  const float w0 = 1.f - q.b1 - q.b2;
  return make_float4(a0.x * w0 + a1.x * q.b1 + a2.x * q.b2, a0.y * w0 + a1.y * q.b1 + a2.y * q.b2,
                     a0.z * w0 + a1.z * q.b1 + a2.z * q.b2, a0.w * w0 + a1.w * q.b1 + a2.w * q.b2);
}
// Texel index of a coordinate outside [0, n) under a wrap mode (sampler.h:62-71).
PHX_DEV int wrap_texel(int i, int n, uint32_t mode) {
  if (mode == PHX_WRAP_REPEAT) return ((i % n) + n) % n;
  if (mode == PHX_WRAP_MIRROR) {
    const int j = ((i % (2 * n)) + 2 * n) % (2 * n);
    return j < n ? j : 2 * n - 1 - j;
  }
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}
// Row-major 4x4 matrix times a column vector, plus an offset.
PHX_DEV float4 affine4(const float *m, const float *o, float4 v) {
  return make_float4(m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3] * v.w + o[0],
                     m[4] * v.x + m[5] * v.y + m[6] * v.z + m[7] * v.w + o[1],
                     m[8] * v.x + m[9] * v.y + m[10] * v.z + m[11] * v.w + o[2],
                     m[12] * v.x + m[13] * v.y + m[14] * v.z + m[15] * v.w + o[3]);
}
// SamplerTable::sample (sampler.h:73-138): the input transform, a nearest or
// multilinear lookup in a 1-3D image (texel rows from coordinate 0), the output
// transform; (0, 0, 0, 1) for an unknown sampler.
PHX_DEV float4 sampler_lookup(const DevScene &s, int id, float4 attr) {
  if (id < 0 || uint32_t(id) >= s.sampler_count) return make_float4(0.f, 0.f, 0.f, 1.f);
  const phx_sampler_info &si = s.samplers[id];
  float4 r = affine4(si.in_transform, si.in_offset, attr);
  if (si.type != PHX_SAMPLER_TRANSFORM) {
    const int dims = si.type == PHX_SAMPLER_IMAGE1D ? 1 : (si.type == PHX_SAMPLER_IMAGE2D ? 2 : 3);
    const float coord[3] = {r.x, r.y, r.z};
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    float fr[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (d >= dims) continue;
      const int n = int(si.size[d]);
      float x = coord[d];
      if (!(x > -1e6f && x < 1e6f)) x = 0.f;  // keeps the texel finite for degenerate coordinates
      if (si.linear) {
        const float f = x * float(n) - 0.5f;
        const float fl = floorf(f);
        fr[d] = f - fl;
        lo[d] = wrap_texel(int(fl), n, si.wrap[d]);
        hi[d] = wrap_texel(int(fl) + 1, n, si.wrap[d]);
      } else {
        lo[d] = hi[d] = wrap_texel(int(floorf(x * float(n))), n, si.wrap[d]);
      }
    }
    const uint32_t sx = si.size[0], sxy = si.size[0] * si.size[1];
    const float4 *tx = s.sampler_texels + si.offset;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    const int nz = dims > 2 ? 2 : 1, ny = dims > 1 ? 2 : 1;
    for (int kz = 0; kz < nz; ++kz) {
      const float wz = dims > 2 ? (kz ? fr[2] : 1.f - fr[2]) : 1.f;
      const uint32_t z = uint32_t(kz ? hi[2] : lo[2]);
      for (int ky = 0; ky < ny; ++ky) {
        const float wy = dims > 1 ? (ky ? fr[1] : 1.f - fr[1]) : 1.f;
        const uint32_t row = z * sxy + uint32_t(ky ? hi[1] : lo[1]) * sx;
        const float4 a = tx[row + uint32_t(lo[0])], b = tx[row + uint32_t(hi[0])];
        const float w = wy * wz, u0 = 1.f - fr[0], u1 = fr[0];
        acc.x = acc.x + (a.x * u0 + b.x * u1) * w;
        acc.y = acc.y + (a.y * u0 + b.y * u1) * w;
        acc.z = acc.z + (a.z * u0 + b.z * u1) * w;
        acc.w = acc.w + (a.w * u0 + b.w * u1) * w;
      }
    }
    r = acc;
  }
  return affine4(si.out_transform, si.out_offset, r);
}
// A non-constant parameter source at the point (pathtracer.cpp:125-133).
PHX_DEV float4 eval_source(const DevScene &s, phx_param_source src, const SurfaceHit &q) {
  if (src.kind == PHX_SOURCE_ATTRIBUTE) return fetch_attribute(s, src.index, q);
  if (src.index < 0 || uint32_t(src.index) >= s.sampler_count) return make_float4(0.f, 0.f, 0.f, 1.f);
  return sampler_lookup(s, src.index, fetch_attribute(s, s.samplers[src.index].in_attribute, q));
}
// Shading normal from a sourced tangent-space normal map (pathtracer.cpp:135-176):
// the tangent frame is dP/du, dP/dv of the triangle in the coordinates the map
// is looked up with, made orthogonal to `sn`; `sn` itself where that frame
// is degenerate.
PHX_DEV vec3 apply_normal_map(const DevScene &s, phx_param_source src, const SurfaceHit &q, vec3 sn) {
  const float4 c = eval_source(s, src, q);
  const vec3 nm{2.f * c.x - 1.f, 2.f * c.y - 1.f, 2.f * c.z - 1.f};
  int slot = src.index;
  if (src.kind != PHX_SOURCE_ATTRIBUTE) {
    if (src.index < 0 || uint32_t(src.index) >= s.sampler_count) return sn;
    slot = s.samplers[src.index].in_attribute;
  }
  SurfaceHit corner = q;
  corner.b1 = 0.f;
  corner.b2 = 0.f;
  const float4 a0 = fetch_attribute(s, slot, corner);
  corner.b1 = 1.f;
  const float4 a1 = fetch_attribute(s, slot, corner);
  corner.b1 = 0.f;
  corner.b2 = 1.f;
  const float4 a2 = fetch_attribute(s, slot, corner);
  const float du1 = a1.x - a0.x, dv1 = a1.y - a0.y, du2 = a2.x - a0.x, dv2 = a2.y - a0.y;
  const float det = du1 * dv2 - du2 * dv1;
  if (!(fabsf(det) > 1e-20f)) return sn;
  const vec3 p0 = ld3(s.positions + size_t(s.indices[q.prim * 3]) * 3);
  const vec3 e1 = ld3(s.positions + size_t(s.indices[q.prim * 3 + 1]) * 3) - p0;
  const vec3 e2 = ld3(s.positions + size_t(s.indices[q.prim * 3 + 2]) * 3) - p0;
  vec3 t = (e1 * dv2 - e2 * dv1) * (1.f / det);
  vec3 b = (e2 * du1 - e1 * du2) * (1.f / det);
  t = t - sn * dot(sn, t);
  b = b - sn * dot(sn, b);
  if (!(dot(t, t) > 0.f) || !(dot(b, b) > 0.f)) return sn;
  const vec3 n = normalize(t) * nm.x + normalize(b) * nm.y + sn * nm.z;
  return dot(n, n) > 0.f ? normalize(n) : sn;
}

// Material at a surface point with atlas textures and parameter sources
// applied (resolve_material, pathtracer.cpp:186-247); `alpha` is the
// probability that a ray interacts with the surface.  TEX = false / SRC =
// false leave the texture / source lookups out of the code (scenes without
// textures / sources, where they never run).
template <bool TEX = true, bool SRC = true>
PHX_DEV phx_material resolve_material(const DevScene &s, const SurfaceHit &q, float &alpha) {
  phx_material m = s.materials[q.mid];
  const bool tex = TEX && s.tex_count > 0;
  alpha = m.alpha;
  if (m.alpha_tex >= 0 && tex) alpha = tex_sample(s, m.alpha_tex, q.uv.x, q.uv.y).x;
  if (m.base_color_tex >= 0 && tex) {
    const vec3 c = tex_sample(s, m.base_color_tex, q.uv.x, q.uv.y);
    m.base_color[0] = c.x;
    m.base_color[1] = c.y;
    m.base_color[2] = c.z;
  } else if (q.has_color) {
    m.base_color[0] = q.color.x;
    m.base_color[1] = q.color.y;
    m.base_color[2] = q.color.z;
  }
  if (m.roughness_tex >= 0 && tex) m.roughness = tex_sample(s, m.roughness_tex, q.uv.x, q.uv.y).x;
  if (m.metallic_tex >= 0 && tex) m.metallic = tex_sample(s, m.metallic_tex, q.uv.x, q.uv.y).x;
  if (SRC && s.has_sources) {
    if (m.base_color_src.kind != PHX_SOURCE_CONSTANT) {
      const float4 c = eval_source(s, m.base_color_src, q);
      m.base_color[0] = fmaxf(c.x, 0.f);
      m.base_color[1] = fmaxf(c.y, 0.f);
      m.base_color[2] = fmaxf(c.z, 0.f);
      if (m.base_color_alpha) alpha *= c.w;
    }
    if (m.alpha_src.kind != PHX_SOURCE_CONSTANT) alpha *= eval_source(s, m.alpha_src, q).x;
    if (m.metallic_src.kind != PHX_SOURCE_CONSTANT) m.metallic = eval_source(s, m.metallic_src, q).x;
    if (m.roughness_src.kind != PHX_SOURCE_CONSTANT) {
      const float r = clamp01(eval_source(s, m.roughness_src, q).x);
      m.roughness = m.roughness_src_squared ? r * r : r;
    }
    if (m.transmission_src.kind != PHX_SOURCE_CONSTANT) m.transmission = eval_source(s, m.transmission_src, q).x;
    if (m.emission_src.kind != PHX_SOURCE_CONSTANT) {
      const float4 e = eval_source(s, m.emission_src, q);
      m.emission[0] = fmaxf(e.x, 0.f);
      m.emission[1] = fmaxf(e.y, 0.f);
      m.emission[2] = fmaxf(e.z, 0.f);
    }
  }
  m.metallic = clamp01(m.metallic);
  m.roughness = clamp01(m.roughness);
  m.transmission = clamp01(m.transmission);
  if (m.alpha_mode == PHX_ALPHA_OPAQUE) alpha = 1.f;
  else if (m.alpha_mode == PHX_ALPHA_MASK) alpha = alpha >= m.alpha_cutoff ? 1.f : 0.f;
  else alpha = clamp01(alpha);
  return m;
}
PHX_DEV vec3 material_emission(const phx_material &m) { return ld3(m.emission) * m.emission_strength; }
// material.h:97-103: also true when the emission comes from a source.
PHX_DEV bool material_may_emit(const phx_material &m) {
  if (m.emission_strength <= 0.f) return false;
  return m.emission_src.kind != PHX_SOURCE_CONSTANT || m.emission_tex >= 0 || m.emission[0] > 0.f ||
         m.emission[1] > 0.f || m.emission[2] > 0.f;
}
PHX_DEV bool emits_on_side(const phx_material &m, bool front) {
  if (m.flags & PHX_MATERIAL_FLAG_EMIT_BOTH_SIDES) return true;
  const bool back_only = (m.flags & PHX_MATERIAL_FLAG_EMIT_BACK_ONLY) != 0;
  return front ? !back_only : back_only;
}
// The emitter's radiance depends on the sampled point: its emission or alpha
// comes from a parameter source.
PHX_DEV bool emission_varies(const phx_material &m) {
  return m.emission_src.kind != PHX_SOURCE_CONSTANT || m.alpha_src.kind != PHX_SOURCE_CONSTANT ||
         (m.base_color_src.kind != PHX_SOURCE_CONSTANT && m.base_color_alpha);
}

// Bit of the emitter record's side mask: the radiance is resolved per sample.
constexpr uint32_t kEmitVaries = 4u;
// The emitter record of emissive-CDF entry i (four float4): the triangle's
// vertices with its area pdf (pathtracer.cpp:411-503); the emitter's radiance,
// emission times alpha of its material resolved at uv (0, 0) as
// sample_direct does (:325-345), with bit 0 / bit 1 of r3.y set when it emits
// toward its front / back side; the primitive and material in r3.z, r3.w.
// When the radiance varies over the triangle (kEmitVaries in r3.y),
// sample_direct resolves the material at the sampled point instead.
PHX_DEV void em_record(const DevScene &s, uint32_t i, float4 *out) {
  const uint32_t prim = s.em_ids[i];
  const float pa = s.prim_light_pdf[prim];
  const vec3 v0 = ld3(s.positions + size_t(s.indices[prim * 3]) * 3);
  const vec3 v1 = ld3(s.positions + size_t(s.indices[prim * 3 + 1]) * 3);
  const vec3 v2 = ld3(s.positions + size_t(s.indices[prim * 3 + 2]) * 3);
  float alpha = 1.f;
  SurfaceHit q{};
  q.prim = prim;
  q.mid = s.material_ids ? s.material_ids[prim] : 0u;
  q.p = v0;
  q.ng = q.sn = normalize(cross(v1 - v0, v2 - v0));
  const phx_material em = resolve_material(s, q, alpha);
  const vec3 L = material_emission(em) * alpha;
  const uint32_t sides = (emits_on_side(em, true) ? 1u : 0u) | (emits_on_side(em, false) ? 2u : 0u) |
                         (emission_varies(em) ? kEmitVaries : 0u);
  out[0] = make_float4(v0.x, v0.y, v0.z, pa);
  out[1] = make_float4(v1.x, v1.y, v1.z, L.x);
  out[2] = make_float4(v2.x, v2.y, v2.z, L.y);
  out[3] = make_float4(L.z, __uint_as_float(sides), __uint_as_float(prim), __uint_as_float(q.mid));
}

// ---- light strategy pick (pathtracer.cpp:270-367) ---------------------------
struct Direct {
  vec3 wi, L;
  float pdf, dist;
  bool mis, valid;
};
// LIGHTS = false leaves the analytic lights, SRC = false the per-sample
// resolve of emitters with sourced radiance out of the code (for scenes
// without them, where those branches never run).
template <bool LIGHTS = true, bool SRC = true>
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
    const uint32_t sides = __float_as_uint(r3.y);
    vec3 L{r1.w, r2.w, r3.x};
    if (SRC && (sides & kEmitVaries)) {
      // the emitter's material at the sampled point (pathtracer.cpp:325-345)
      SurfaceHit q{};
      q.prim = __float_as_uint(r3.z);
      q.mid = __float_as_uint(r3.w);
      q.b1 = b1;
      q.b2 = b2;
      q.p = lp;
      q.ng = q.sn = ln;
      float alpha = 1.f;
      const phx_material em = resolve_material(s, q, alpha);
      if (!emits_on_side(em, cl > 0.f)) return ds;
      L = material_emission(em) * alpha;
    } else if (!(sides & (cl > 0.f ? 1u : 2u))) {
      return ds;
    }
    ds.wi = wi;
    ds.L = L;
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
// A ray's walk through a volume's macro-cell grid, one cell at a time (a 3-D
// DDA).  Per axis: `at` is the cell, `dir` the step (+1, -1 or 0), `face` the
// ray parameter at which the walk leaves the cell through that axis's next
// face and `stride` the parameter between two such faces.  The arithmetic is
// the reference's (tf_volume.h:150-186), so the cells and their boundaries
// fall where its do.  (Three named axes rather than arrays: arrays indexed
// by the loop variable kept the walk and the volume record in scratch.)
struct WalkAxis {
  int at, dir;
  float face, stride;
  PHX_DEV void start(uint32_t cells, uint32_t dims, float spacing, float origin, float o, float d, float t0) {
    const float extent = spacing * float(dims > 1 ? dims - 1 : 1);
    const float per_unit = float(int(cells)) / extent;  // cells per unit length
    const float g = (o + d * t0 - origin) * per_unit;
    const int c = int(g);
    at = c < 0 ? 0 : (c > int(cells) - 1 ? int(cells) - 1 : c);
    const float rate = d * per_unit;  // cells per unit of t
    if (rate > 0.f) {
      dir = 1;
      stride = 1.f / rate;
      face = t0 + (float(at + 1) - g) * stride;
    } else if (rate < 0.f) {
      dir = -1;
      stride = -1.f / rate;
      face = t0 + (g - float(at)) * stride;
    } else {
      dir = 0;
      stride = 1e30f;
      face = 1e30f;
    }
  }
  // Moves to the neighbouring cell when `go` (by value selects: branches per
  // axis were merged by the compiler into one update through a selected
  // address, which put the walk in scratch memory).
  PHX_DEV void step_if(bool go) {
    at = go ? at + dir : at;
    face = go ? face + stride : face;
  }
};
struct MacroWalk {
  WalkAxis x, y, z;
  PHX_DEV void start(const phx_volume &v, vec3 o, vec3 d, float t0) {
    x.start(v.macro_dims[0], v.dims[0], v.spacing[0], v.origin[0], o.x, d.x, t0);
    y.start(v.macro_dims[1], v.dims[1], v.spacing[1], v.origin[1], o.y, d.y, t0);
    z.start(v.macro_dims[2], v.dims[2], v.spacing[2], v.origin[2], o.z, d.z, t0);
  }
  // the axis whose face comes first (ties go to the later axis, as the
  // reference's comparisons do)
  PHX_DEV int next_axis() const { return x.face < y.face ? (x.face < z.face ? 0 : 2) : (y.face < z.face ? 1 : 2); }
  // that face's parameter (the smallest of the three; no NaN arises here)
  PHX_DEV float next_face() const { return fminf(fminf(x.face, y.face), z.face); }
  PHX_DEV uint32_t cell_index(const uint32_t m[3]) const {
    return (uint32_t(z.at) * m[1] + uint32_t(y.at)) * m[0] + uint32_t(x.at);
  }
  // Crosses into the neighbour along `ax` of a grid of m[3] cells; false
  // (and no move) when that leaves the grid.
  PHX_DEV bool cross(const uint32_t m[3], int ax) {
    const bool ox = ax == 0, oy = ax == 1, oz = !ox && !oy;
    const int next = ox ? x.at + x.dir : (oy ? y.at + y.dir : z.at + z.dir);
    const int cells = int(ox ? m[0] : (oy ? m[1] : m[2]));
    if (next < 0 || next >= cells) return false;
    x.step_if(ox);
    y.step_if(oy);
    z.step_if(oz);
    return true;
  }
};

// Bounds of the walk: cells per ray and density lookups per ray
// (tf_volume.h:188-190).
constexpr int kMaxWalkCells = 4096, kMaxLookups = 8192;

// Tracking along [t0, t1] of one volume in its local frame, cell by cell
// with each cell's majorant (tf_volume.h:192-283).  SCATTER: delta tracking,
// returns the first real collision's distance (its albedo in `albedo`) or
// -1.  Otherwise ratio tracking: scales `tr` by the transmittance, with
// Russian roulette below 0.05, and returns -1 (tr = 0 when the roulette ends
// the ray).  Each cell entered with a majorant draws a fresh free-flight
// distance; a distance past the cell's far face is dropped at the face.
template <bool SCATTER>
PHX_DEV float vol_track(const DevScene &s, const phx_volume &v, vec3 org, vec3 dir, float t0, float t1, Rng &rng,
                        vec3 &albedo, float &tr) {
  MacroWalk w;
  w.start(v, org, dir, t0);
  const uint32_t cells_of[3] = {v.macro_dims[0], v.macro_dims[1], v.macro_dims[2]};
  float t = t0;
  int lookups = 0;
  for (int cells = 0; cells < kMaxWalkCells; ++cells) {
    const int ax = w.next_axis();
    const float face = w.next_face();
    const float seg_end = face < t1 ? face : t1;
    const float sigma_max = s.majorants[v.macro_offset + w.cell_index(cells_of)];
    if (sigma_max > 0.f) {
      const float rcp_max = 1.f / sigma_max;
      while (lookups < kMaxLookups) {
        t -= logf(1.f - rng.next()) * rcp_max;
        if (t >= seg_end) break;
        ++lookups;
        const float4 c = vol_classify(s, v, org + dir * t);
        if (SCATTER) {
          if (rng.next() * sigma_max < c.w) {  // a real collision
            albedo = {c.x, c.y, c.z};
            return t;
          }
        } else {
          tr *= 1.f - c.w * rcp_max;
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
    t = seg_end;
    if (seg_end >= t1 || lookups >= kMaxLookups || !w.cross(cells_of, ax)) break;
  }
  return -1.f;
}
// The nearest real collision in front of tmax over all volumes (each one
// tracked only up to the nearest collision found so far), or -1.
PHX_DEV float vol_sample_scatter(const DevScene &s, vec3 ro, vec3 rd, float tmin, float tmax, Rng &rng, vec3 &albedo) {
  float nearest = -1.f, limit = tmax, tr_unused = 1.f;
  for (uint32_t i = 0; i < s.vol_count; ++i) {
    const phx_volume v = s.volumes[i];
    if (v.majorant <= 0.f || v.tf_size == 0) continue;
    const vec3 org = to_local_point(v, ro), dir = to_local_vector(v, rd);
    float t0 = tmin, t1 = limit;
    if (!vol_clip(v, org, dir, t0, t1)) continue;
    vec3 a{0.f, 0.f, 0.f};
    const float t = vol_track<true>(s, v, org, dir, t0, t1, rng, a, tr_unused);
    if (t >= 0.f) {
      nearest = t;
      limit = t;
      albedo = a;
    }
  }
  return nearest;
}
// Transmittance along [tmin, tmax] through every volume.
PHX_DEV float vol_transmittance(const DevScene &s, vec3 ro, vec3 rd, float tmin, float tmax, Rng &rng) {
  float tr = 1.f;
  vec3 albedo_unused{0.f, 0.f, 0.f};
  for (uint32_t i = 0; i < s.vol_count; ++i) {
    const phx_volume v = s.volumes[i];
    if (v.majorant <= 0.f || v.tf_size == 0) continue;
    const vec3 org = to_local_point(v, ro), dir = to_local_vector(v, rd);
    float t0 = tmin, t1 = tmax;
    if (!vol_clip(v, org, dir, t0, t1)) continue;
    vol_track<false>(s, v, org, dir, t0, t1, rng, albedo_unused, tr);
    if (tr <= 0.f) return 0.f;
  }
  return tr;
}

}  // namespace phx
