// ORACLE - test infrastructure only (see orc_core.h).  Lights, environment
// map, textures, camera, triangle test, the reference's binary BVH and the
// transfer-function volumes, restated from the reference's headers.
#pragma once

#include "orc_core.h"

namespace orc {

// ---- lights (light.h:45-148) -------------------------------------------------
template <class T> struct LightSample { V3<T> wi, Li; T pdf{0}, dist{0}; };
template <class T> inline T clampT(T x, T a, T b) { return x < a ? a : (x > b ? b : x); }
template <class T> inline T smoothstep(T e0, T e1, T x) {
  const T t = clampT((x - e0) / (e1 - e0), T(0), T(1));
  return t * t * (T(3) - T(2) * t);
}
template <class T> inline V3<T> sample_sphere_light_point(const phx_light &l, const V3<T> &p, Rng &rng) {  // :58-67
  const V3<T> pos = v3<T>(l.position);
  if (T(l.radius) <= T(0)) return pos;
  const V3<T> axis = normalize(pos - p);
  V3<T> t, b;
  onb_from_normal(axis, t, b);
  const V2<T> d = sample_concentric_disk<T>(rng);
  return pos + (t * d.x + b * d.y) * T(l.radius);
}
template <class T> inline LightSample<T> sample_light(const phx_light &l, const V3<T> &p, Rng &rng) {
  const T INF = T(1e30);
  const V3<T> color = v3<T>(l.color);
  const T intensity = T(l.intensity);
  if (l.type == PHX_LIGHT_POINT) {  // :76-84
    const V3<T> ct = v3<T>(l.position) - p;
    const T cd2 = dot(ct, ct);
    const V3<T> to = sample_sphere_light_point<T>(l, p, rng) - p;
    const T dist = length(to);
    const V3<T> wi = dist > T(0) ? (to * (T(1) / dist)) : V3<T>{T(0), T(1), T(0)};
    const T inv = cd2 > T(0) ? (T(1) / cd2) : T(0);
    return {wi, color * (intensity * inv), T(1), dist};
  }
  if (l.type == PHX_LIGHT_DIRECTIONAL) {  // :86-99
    V3<T> wi = -normalize(v3<T>(l.direction));
    if (T(l.radius) > T(0)) {
      const T cos_max = std::cos(T(l.radius));
      const T ct = T(1) - draw<T>(rng) * (T(1) - cos_max);
      const T st = std::sqrt(fmaxT(T(0), T(1) - ct * ct));
      const T phi = K<T>::TWO_PI * draw<T>(rng);
      V3<T> t, b;
      onb_from_normal(wi, t, b);
      wi = normalize(t * (st * std::cos(phi)) + b * (st * std::sin(phi)) + wi * ct);
    }
    return {wi, color * intensity, T(1), INF};
  }
  if (l.type == PHX_LIGHT_SPOT) {  // :101-125
    const V3<T> ct = v3<T>(l.position) - p;
    const T cd2 = dot(ct, ct);
    const V3<T> to = sample_sphere_light_point<T>(l, p, rng) - p;
    const T dist = length(to);
    const V3<T> wi = dist > T(0) ? (to * (T(1) / dist)) : V3<T>{T(0), T(1), T(0)};
    const V3<T> sd = normalize(v3<T>(l.direction));
    const V3<T> from = -normalize(ct);
    const T ca = dot(from, sd);
    const T ci = std::cos(fmaxT(T(l.spot_angle) - T(l.spot_falloff), T(0)));
    const T co = std::cos(T(l.spot_angle));
    T fall = T(0);
    if (ca >= ci) fall = T(1);
    else if (ca <= co) fall = T(0);
    else fall = smoothstep(co, ci, ca);
    const T inv = cd2 > T(0) ? (T(1) / cd2) : T(0);
    return {wi, color * (intensity * inv * fall), T(1), dist};
  }
  if (l.type == PHX_LIGHT_AREA) {  // :127-145
    const T u = draw<T>(rng), v = draw<T>(rng);
    const V3<T> e1 = v3<T>(l.edge1), e2 = v3<T>(l.edge2);
    const V3<T> sp = v3<T>(l.position) + e1 * u + e2 * v;
    const V3<T> ln = normalize(cross(e1, e2));
    const V3<T> to = sp - p;
    const T dist = length(to);
    if (dist < T(1e-6)) return {V3<T>{T(0), T(1), T(0)}, V3<T>{}, T(0), T(0)};
    const V3<T> wi = to * (T(1) / dist);
    const T cl = -dot(wi, ln);
    if (cl <= T(0)) return {V3<T>{T(0), T(1), T(0)}, V3<T>{}, T(0), T(0)};
    const T area = T(l.area) > T(0) ? T(l.area) : length(cross(e1, e2));
    const T pdf = (T(1) / area) * (dist * dist) / cl;
    return {wi, color * intensity, pdf, dist * T(0.999)};
  }
  return {V3<T>{T(0), T(1), T(0)}, V3<T>{}, T(0), T(0)};
}

// ---- environment map (environment_map.h:13-156, environment_map.cpp:8-65) ----
struct EnvMap {
  uint32_t width{0}, height{0};
  std::vector<float> pixels;       // H*W*3, row 0 = v = 0 (+Y of the map)
  std::vector<float> marginal;     // H+1
  std::vector<float> conditional;  // H*(W+1)
  float w2t[9]{1, 0, 0, 0, 1, 0, 0, 0, 1};
  float t2w[9]{1, 0, 0, 0, 1, 0, 0, 0, 1};

  // environment_map.cpp:8-65, in float as the reference does it.
  void build_cdf() {
    const float PI = 3.14159265358979323846f;
    marginal.assign(height + 1, 0.f);
    conditional.assign(size_t(height) * (width + 1), 0.f);
    std::vector<float> row_w(height, 0.f);
    marginal[0] = 0.f;
    for (uint32_t y = 0; y < height; ++y) {
      float *cond = &conditional[size_t(y) * (width + 1)];
      cond[0] = 0.f;
      const float theta = (float(y) + 0.5f) * (PI / float(height));
      const float st = std::sin(theta);
      float sum = 0.f;
      for (uint32_t x = 0; x < width; ++x) {
        const float *p = &pixels[(size_t(y) * width + x) * 3];
        sum += 0.2126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2];
        cond[x + 1] = sum;
      }
      if (sum > 0.f) {
        const float inv = 1.f / sum;
        for (uint32_t x = 0; x <= width; ++x) cond[x] *= inv;
        cond[width] = 1.f;
      } else {
        for (uint32_t x = 0; x <= width; ++x) cond[x] = float(x) / float(width);
      }
      row_w[y] = sum * st;
    }
    float total = 0.f;
    for (uint32_t y = 0; y < height; ++y) {
      total += row_w[y];
      marginal[y + 1] = total;
    }
    if (total > 0.f) {
      const float inv = 1.f / total;
      for (uint32_t y = 0; y <= height; ++y) marginal[y] *= inv;
      marginal[height] = 1.f;
    } else {
      for (uint32_t y = 0; y <= height; ++y) marginal[y] = float(y) / float(height);
    }
  }
  template <class T> V3<T> to_tex(const V3<T> &d) const {
    const float *m = w2t;
    return {T(m[0]) * d.x + T(m[1]) * d.y + T(m[2]) * d.z, T(m[3]) * d.x + T(m[4]) * d.y + T(m[5]) * d.z,
            T(m[6]) * d.x + T(m[7]) * d.y + T(m[8]) * d.z};
  }
  template <class T> V3<T> to_world(const V3<T> &d) const {
    const float *m = t2w;
    return {T(m[0]) * d.x + T(m[1]) * d.y + T(m[2]) * d.z, T(m[3]) * d.x + T(m[4]) * d.y + T(m[5]) * d.z,
            T(m[6]) * d.x + T(m[7]) * d.y + T(m[8]) * d.z};
  }
  // environment_map.h:136-145
  template <class T> static V2<T> dir_to_uv(const V3<T> &d) {
    const V3<T> nd = normalize(d);
    const T phi = std::atan2(nd.z, nd.x);
    const T y = clampT(nd.y, T(-1), T(1));
    const T theta = std::acos(y);
    return {(phi + K<T>::PI) * K<T>::INV_TWO_PI, theta * K<T>::INV_PI};
  }
  // environment_map.h:147-153
  template <class T> static V3<T> uv_to_dir(const V2<T> &uv) {
    const T phi = uv.x * K<T>::TWO_PI - K<T>::PI;
    const T theta = uv.y * K<T>::PI;
    const T st = std::sin(theta);
    return {st * std::cos(phi), std::cos(theta), st * std::sin(phi)};
  }
  // environment_map.h:40-54: nearest texel.
  template <class T> void texel(const V3<T> &dir, uint32_t &x, uint32_t &y, V2<T> &uv) const {
    uv = dir_to_uv(to_tex(dir));
    const T xf = uv.x * T(width), yf = uv.y * T(height);
    x = uint32_t(fminT(T(width - 1), fmaxT(T(0), xf)));
    y = uint32_t(fminT(T(height - 1), fmaxT(T(0), yf)));
  }
  template <class T> V3<T> evaluate(const V3<T> &dir) const {
    if (!width || !height) return {};
    uint32_t x, y;
    V2<T> uv;
    texel(dir, x, y, uv);
    return v3<T>(&pixels[(size_t(y) * width + x) * 3]);
  }
  // environment_map.h:56-115
  template <class T> V3<T> sample_direction(Rng &rng, T &pdf) const {
    pdf = T(0);
    if (!width || !height) return {T(0), T(1), T(0)};
    const float u1 = rng.next_f32(), u2 = rng.next_f32();
    uint32_t lo = 0, hi = height;
    while (lo + 1 < hi) {
      const uint32_t mid = (lo + hi) / 2;
      if (marginal[mid] <= u2) lo = mid;
      else hi = mid;
    }
    const uint32_t row = lo;
    const float *cond = &conditional[size_t(row) * (width + 1)];
    lo = 0;
    hi = width;
    while (lo + 1 < hi) {
      const uint32_t mid = (lo + hi) / 2;
      if (cond[mid] <= u1) lo = mid;
      else hi = mid;
    }
    const uint32_t col = lo;
    const T row_pdf = T(marginal[row + 1] - marginal[row]);
    const T col_pdf = T(cond[col + 1] - cond[col]);
    const T u = (T(col) + draw<T>(rng)) / T(width);
    const T v = (T(row) + draw<T>(rng)) / T(height);
    const V3<T> dir = to_world(uv_to_dir(V2<T>{u, v}));
    const T st = std::sin(v * K<T>::PI);
    const T denom = K<T>::TWO_PI * K<T>::PI * fmaxT(st, T(1e-6));
    pdf = (row_pdf * col_pdf) * T(width) * T(height) / denom;
    return dir;
  }
  // environment_map.h:117-136
  template <class T> T pdf(const V3<T> &dir) const {
    if (!width || !height) return T(0);
    uint32_t x, y;
    V2<T> uv;
    texel(dir, x, y, uv);
    const T row_pdf = T(marginal[y + 1] - marginal[y]);
    const T col_pdf = T(conditional[size_t(y) * (width + 1) + x + 1] - conditional[size_t(y) * (width + 1) + x]);
    const T st = std::sin(uv.y * K<T>::PI);
    const T denom = K<T>::TWO_PI * K<T>::PI * fmaxT(st, T(1e-6));
    return (row_pdf * col_pdf) * T(width) * T(height) / denom;
  }
};

// ---- texture atlas (texture.h:25-64) ----------------------------------------
struct Atlas {
  std::vector<float> pixels;
  std::vector<phx_texture_info> infos;
  template <class T> V3<T> sample(int id, T u, T v) const { return sample_from<T>(pixels.data(), id, u, v); }
  // The same lookup over another pixel array of the atlas layout (float or double).
  template <class T, class P> V3<T> sample_from(const P *pix, int id, T u, T v) const {
    if (id < 0 || uint32_t(id) >= infos.size()) return {T(1), T(1), T(1)};
    const phx_texture_info ti = infos[id];
    if (!ti.width || !ti.height) return {T(1), T(1), T(1)};
    u = u - std::floor(u);
    v = T(1) - (v - std::floor(v));
    const T fx = u * T(ti.width) - T(0.5), fy = v * T(ti.height) - T(0.5);
    const int ix = int(std::floor(fx)), iy = int(std::floor(fy));
    const T tx = fx - std::floor(fx), ty = fy - std::floor(fy);
    auto wrap = [](int i, uint32_t sz) { const int s = int(sz); return uint32_t(((i % s) + s) % s); };
    const uint32_t x0 = wrap(ix, ti.width), x1 = wrap(ix + 1, ti.width);
    const uint32_t y0 = wrap(iy, ti.height), y1 = wrap(iy + 1, ti.height);
    auto px = [&](uint32_t x, uint32_t y) {
      const P *q = &pix[(size_t(ti.offset) + y * ti.width + x) * 3];
      return V3<T>{T(q[0]), T(q[1]), T(q[2])};
    };
    const V3<T> top = px(x0, y0) * (T(1) - tx) + px(x1, y0) * tx;
    const V3<T> bot = px(x0, y1) * (T(1) - tx) + px(x1, y1) * tx;
    return top * (T(1) - ty) + bot * ty;
  }
};

// ---- camera (camera.h:36-183) ------------------------------------------------
inline phx_camera make_perspective(V3<float> from, V3<float> at, V3<float> vup, float vfov, float aspect,
                                   float aperture, float focus_dist) {
  const float theta = vfov * 3.1415926535f / 180.f;
  const float h = std::tan(theta * 0.5f);
  const float vh = 2.f * h, vw = aspect * vh;
  const V3<float> w = normalize(from - at);
  const V3<float> u = normalize(cross(vup, w));
  const V3<float> v = cross(w, u);
  const V3<float> hor = u * (vw * focus_dist), ver = v * (vh * focus_dist);
  const V3<float> ll = from - hor * 0.5f - ver * 0.5f - w * focus_dist;
  phx_camera c{};
  auto put = [](float *d, const V3<float> &s) { d[0] = s.x; d[1] = s.y; d[2] = s.z; };
  put(c.origin, from); put(c.lower_left, ll); put(c.horizontal, hor); put(c.vertical, ver);
  put(c.u, u); put(c.v, v); put(c.w, w);
  c.lens_radius = aperture * 0.5f;
  c.focus_dist = focus_dist;
  c.is_ortho = 0;
  c.clip_near = 0.f;
  c.clip_far = 1e30f;
  return c;
}
// camera.h:143-159
template <class T> inline void clip_range(const phx_camera &c, const V3<T> &dir, T &tn, T &tf) {
  tn = T(0);
  tf = T(1e30);
  if (c.clip_near <= 0.f && c.clip_far >= 1e30f) return;
  const T ca = dot(dir, v3<T>(c.w) * T(-1));
  if (ca <= T(1e-6)) { tn = T(1); tf = T(0); return; }
  if (c.clip_near > 0.f) tn = T(c.clip_near) / ca;
  if (c.clip_far < 1e30f) tf = T(c.clip_far) / ca;
}
// camera.h:167-182
template <class T> inline void camera_ray(const phx_camera &c, T s, T t, Rng &rng, V3<T> &org, V3<T> &dir) {
  const V3<T> ll = v3<T>(c.lower_left), hor = v3<T>(c.horizontal), ver = v3<T>(c.vertical), o = v3<T>(c.origin);
  if (c.is_ortho) {
    org = ll + hor * s + ver * t;
    dir = -v3<T>(c.w);
    return;
  }
  if (T(c.lens_radius) <= T(0)) {
    org = o;
    dir = normalize(ll + hor * s + ver * t - o);
    return;
  }
  const V2<T> d = sample_concentric_disk<T>(rng);
  const V2<T> rd{d.x * T(c.lens_radius), d.y * T(c.lens_radius)};
  const V3<T> off = v3<T>(c.u) * rd.x + v3<T>(c.v) * rd.y;
  const V3<T> target = ll + hor * s + ver * t;
  org = o + off;
  dir = normalize(target - o - off);
}

// ---- triangle test (triangle_intersect.h:18-52) -----------------------------
struct TriHit { float t{0}, u{0}, v{0}; bool hit{false}; };
inline TriHit intersect_triangle(const V3<float> &org, const V3<float> &dir, const V3<float> &v0, const V3<float> &v1,
                                 const V3<float> &v2, float tmin, float tmax) {
  const V3<float> e1 = v1 - v0, e2 = v2 - v0;
  const V3<float> p = cross(dir, e2);
  const float det = dot(e1, p);
  if (std::fabs(det) < 1e-30f) return {};
  const float inv = 1.f / det;
  const V3<float> tv = org - v0;
  const float u = dot(tv, p) * inv;
  if (u < 0.f || u > 1.f) return {};
  const V3<float> q = cross(tv, e1);
  const float v = dot(dir, q) * inv;
  if (v < 0.f || (u + v) > 1.f) return {};
  const float t = dot(e2, q) * inv;
  if (t < tmin || t > tmax) return {};
  return {t, u, v, true};
}

// ---- the reference's binary BVH (bvh.cpp:20-125) and traversal
// (mesh_intersector.h:30-124).  Also the structure of the Kokkos-HIP proxy. --
struct Box { V3<float> lo{1e30f, 1e30f, 1e30f}, hi{-1e30f, -1e30f, -1e30f}; };
inline void extend(Box &b, const V3<float> &p) {  // math_aabb.h:13-20
  b.lo.x = b.lo.x < p.x ? b.lo.x : p.x; b.lo.y = b.lo.y < p.y ? b.lo.y : p.y; b.lo.z = b.lo.z < p.z ? b.lo.z : p.z;
  b.hi.x = b.hi.x > p.x ? b.hi.x : p.x; b.hi.y = b.hi.y > p.y ? b.hi.y : p.y; b.hi.z = b.hi.z > p.z ? b.hi.z : p.z;
}
struct Node { Box b; uint32_t left{0}, right{0}, begin{0}, count{0}, leaf{0}; };
struct Mesh {
  std::vector<float> pos;      // V*3
  std::vector<uint32_t> idx;   // T*3
  std::vector<float> nrm;      // V*3 or empty
  std::vector<float> uv;       // V*2 or empty
  std::vector<float> col;      // V*3 or empty
  std::vector<uint32_t> mat;   // T or empty
  uint32_t ntri() const { return uint32_t(idx.size() / 3); }
  V3<float> P(uint32_t i) const { return v3<float>(&pos[size_t(i) * 3]); }
};
struct Bvh {
  std::vector<Node> nodes;
  std::vector<uint32_t> prim_ids;
  uint32_t root{0};
};
void build_bvh(const Mesh &m, Bvh &out);

struct ClosestHit { float t{0}, u{0}, v{0}; uint32_t prim{0}; bool hit{false}; };
// mesh_intersector.h:30-45
inline float aabb_entry(const Box &b, const V3<float> &o, const V3<float> &inv, float tmin, float tmax) {
  const float tx0 = (b.lo.x - o.x) * inv.x, tx1 = (b.hi.x - o.x) * inv.x;
  const float ty0 = (b.lo.y - o.y) * inv.y, ty1 = (b.hi.y - o.y) * inv.y;
  const float tz0 = (b.lo.z - o.z) * inv.z, tz1 = (b.hi.z - o.z) * inv.z;
  const float tn = std::fmax(std::fmax(std::fmin(tx0, tx1), std::fmin(ty0, ty1)), std::fmax(std::fmin(tz0, tz1), tmin));
  const float tf = std::fmin(std::fmin(std::fmax(tx0, tx1), std::fmax(ty0, ty1)), std::fmin(std::fmax(tz0, tz1), tmax));
  return tn <= tf ? tn : -1.f;
}
inline V3<float> safe_inv(const V3<float> &d) {  // :47-53
  auto inv = [](float x) { return 1.f / (std::fabs(x) > 1e-20f ? x : (x < 0.f ? -1e-20f : 1e-20f)); };
  return {inv(d.x), inv(d.y), inv(d.z)};
}
ClosestHit trace_bvh(const Mesh &m, const Bvh &bvh, const V3<float> &o, const V3<float> &d, float tmin, float tmax,
                     bool any_hit);
// All triangles in index order; a tie at equal t goes to the later triangle,
// which is what the reference's "t <= best.t accepted" rule gives for that order.
ClosestHit trace_brute(const Mesh &m, const V3<float> &o, const V3<float> &d, float tmin, float tmax, bool any_hit);

// ---- transfer-function volumes (tf_volume.h:43-283) ------------------------
struct Volumes {
  std::vector<phx_volume> vols;
  std::vector<float> values;
  std::vector<float> tf;  // RGBA
  std::vector<float> majorants;
  bool empty() const { return vols.empty(); }
  template <class T> static V3<T> to_local_point(const phx_volume &v, const V3<T> &p) {
    const float *m = v.world_to_local;
    return {T(m[0]) * p.x + T(m[1]) * p.y + T(m[2]) * p.z + T(m[3]), T(m[4]) * p.x + T(m[5]) * p.y + T(m[6]) * p.z + T(m[7]),
            T(m[8]) * p.x + T(m[9]) * p.y + T(m[10]) * p.z + T(m[11])};
  }
  template <class T> static V3<T> to_local_vector(const phx_volume &v, const V3<T> &d) {
    const float *m = v.world_to_local;
    return {T(m[0]) * d.x + T(m[1]) * d.y + T(m[2]) * d.z, T(m[4]) * d.x + T(m[5]) * d.y + T(m[6]) * d.z,
            T(m[8]) * d.x + T(m[9]) * d.y + T(m[10]) * d.z};
  }
  template <class T> static bool clip(const phx_volume &v, const V3<T> &org, const V3<T> &dir, T &t0, T &t1) {  // :50-69
    const T o[3] = {org.x, org.y, org.z}, d[3] = {dir.x, dir.y, dir.z};
    for (int a = 0; a < 3; ++a) {
      const T lo = T(v.origin[a]);
      const T hi = lo + T(v.spacing[a]) * T(v.dims[a] > 1 ? v.dims[a] - 1 : 1);
      if (d[a] == T(0)) {
        if (o[a] < lo || o[a] > hi) return false;
        continue;
      }
      const T inv = T(1) / d[a];
      const T ta = (lo - o[a]) * inv, tb = (hi - o[a]) * inv;
      const T tn = ta < tb ? ta : tb, tf = ta < tb ? tb : ta;
      t0 = tn > t0 ? tn : t0;
      t1 = tf < t1 ? tf : t1;
      if (t1 <= t0) return false;
    }
    return true;
  }
  template <class T> T field(const phx_volume &v, const V3<T> &p) const {  // :71-110
    const T g[3] = {(p.x - T(v.origin[0])) / T(v.spacing[0]), (p.y - T(v.origin[1])) / T(v.spacing[1]),
                    (p.z - T(v.origin[2])) / T(v.spacing[2])};
    int i0[3], i1[3];
    T fr[3];
    for (int a = 0; a < 3; ++a) {
      const int n = int(v.dims[a]);
      T x = g[a];
      x = x < T(0) ? T(0) : (x > T(n - 1) ? T(n - 1) : x);
      if (v.linear) {
        int i = int(x);
        if (i > n - 2) i = n - 2;
        if (i < 0) i = 0;
        i0[a] = i;
        i1[a] = i + 1 < n ? i + 1 : i;
        fr[a] = x - T(i);
      } else {
        int i = int(x + T(0.5));
        i = i > n - 1 ? n - 1 : i;
        i0[a] = i1[a] = i;
        fr[a] = T(0);
      }
    }
    const uint32_t sx = v.dims[0], sxy = v.dims[0] * v.dims[1];
    auto at = [&](int x, int y, int z) { return T(values[size_t(v.data_offset) + size_t(z) * sxy + size_t(y) * sx + size_t(x)]); };
    if (!v.linear) return at(i0[0], i0[1], i0[2]);
    const T c00 = at(i0[0], i0[1], i0[2]) * (T(1) - fr[0]) + at(i1[0], i0[1], i0[2]) * fr[0];
    const T c10 = at(i0[0], i1[1], i0[2]) * (T(1) - fr[0]) + at(i1[0], i1[1], i0[2]) * fr[0];
    const T c01 = at(i0[0], i0[1], i1[2]) * (T(1) - fr[0]) + at(i1[0], i0[1], i1[2]) * fr[0];
    const T c11 = at(i0[0], i1[1], i1[2]) * (T(1) - fr[0]) + at(i1[0], i1[1], i1[2]) * fr[0];
    const T c0 = c00 * (T(1) - fr[1]) + c10 * fr[1];
    const T c1 = c01 * (T(1) - fr[1]) + c11 * fr[1];
    return c0 * (T(1) - fr[2]) + c1 * fr[2];
  }
  template <class T> V4<T> classify(const phx_volume &v, const V3<T> &p) const {  // :113-130
    const T value = field(v, p);
    const T range = T(v.value_hi) - T(v.value_lo);
    T x = range != T(0) ? (value - T(v.value_lo)) / range : T(0);
    x = x < T(0) ? T(0) : (x > T(1) ? T(1) : x);
    const T f = x * T(v.tf_size - 1);
    uint32_t i = uint32_t(f);
    if (i > v.tf_size - 2) i = v.tf_size >= 2 ? v.tf_size - 2 : 0;
    const T fr = v.tf_size >= 2 ? f - T(i) : T(0);
    const float *a = &tf[(size_t(v.tf_offset) + i) * 4];
    const float *b = &tf[(size_t(v.tf_offset) + (v.tf_size >= 2 ? i + 1 : i)) * 4];
    V4<T> c{T(a[0]) * (T(1) - fr) + T(b[0]) * fr, T(a[1]) * (T(1) - fr) + T(b[1]) * fr,
            T(a[2]) * (T(1) - fr) + T(b[2]) * fr, T(a[3]) * (T(1) - fr) + T(b[3]) * fr};
    c.w = c.w * T(v.density_scale);
    return c;
  }
  // :150-232
  template <class T> T track(const phx_volume &v, const V3<T> &org, const V3<T> &dir, T t0, T t1, Rng &rng, bool scatter,
                             V3<T> &albedo, T &tr) const {
    const T o[3] = {org.x, org.y, org.z}, d[3] = {dir.x, dir.y, dir.z};
    int cell[3], step[3];
    T t_next[3], t_delta[3];
    for (int a = 0; a < 3; ++a) {
      const int m = int(v.macro_dims[a]);
      const T extent = T(v.spacing[a]) * T(v.dims[a] > 1 ? v.dims[a] - 1 : 1);
      const T cpu = T(m) / extent;
      const T g = (o[a] + d[a] * t0 - T(v.origin[a])) * cpu;
      int c = int(g);
      c = c < 0 ? 0 : (c > m - 1 ? m - 1 : c);
      cell[a] = c;
      const T speed = d[a] * cpu;
      if (speed > T(0)) {
        step[a] = 1; t_delta[a] = T(1) / speed; t_next[a] = t0 + (T(c + 1) - g) * t_delta[a];
      } else if (speed < T(0)) {
        step[a] = -1; t_delta[a] = T(-1) / speed; t_next[a] = t0 + (g - T(c)) * t_delta[a];
      } else {
        step[a] = 0; t_delta[a] = T(1e30); t_next[a] = T(1e30);
      }
    }
    T t = t0;
    int lookups = 0;
    for (int visited = 0; visited < 4096; ++visited) {
      const int axis = t_next[0] < t_next[1] ? (t_next[0] < t_next[2] ? 0 : 2) : (t_next[1] < t_next[2] ? 1 : 2);
      const T t_exit = t_next[axis] < t1 ? t_next[axis] : t1;
      const T bound = T(majorants[size_t(v.macro_offset) +
                                  (size_t(cell[2]) * v.macro_dims[1] + size_t(cell[1])) * v.macro_dims[0] + size_t(cell[0])]);
      if (bound > T(0)) {
        const T inv_bound = T(1) / bound;
        while (lookups < 8192) {
          t -= std::log(T(1) - draw<T>(rng)) * inv_bound;
          if (t >= t_exit) break;
          ++lookups;
          const V4<T> c = classify(v, org + dir * t);
          if (scatter) {
            if (draw<T>(rng) * bound < c.w) {
              albedo = {c.x, c.y, c.z};
              return t;
            }
          } else {
            tr *= T(1) - c.w * inv_bound;
            if (tr < T(0.05)) {
              if (!(tr > T(0)) || draw<T>(rng) * T(0.05) >= tr) {
                tr = T(0);
                return T(-1);
              }
              tr = T(0.05);
            }
          }
        }
      }
      t = t_exit;
      if (t_exit >= t1 || lookups >= 8192) break;
      cell[axis] += step[axis];
      if (cell[axis] < 0 || cell[axis] >= int(v.macro_dims[axis])) break;
      t_next[axis] += t_delta[axis];
    }
    return T(-1);
  }
  // :237-261
  template <class T> T sample_scatter(const V3<T> &ro, const V3<T> &rd, T tmin, T tmax, Rng &rng, V3<T> &albedo) const {
    T t_event = T(-1), t_limit = tmax, unused = T(1);
    for (const phx_volume &v : vols) {
      if (v.majorant <= 0.f || v.tf_size == 0) continue;
      const V3<T> org = to_local_point(v, ro), dir = to_local_vector(v, rd);
      T t0 = tmin, t1 = t_limit;
      if (!clip(v, org, dir, t0, t1)) continue;
      V3<T> ea{};
      const T t = track(v, org, dir, t0, t1, rng, true, ea, unused);
      if (t >= T(0)) {
        t_event = t;
        t_limit = t;
        albedo = ea;
      }
    }
    return t_event;
  }
  // :264-283
  template <class T> T transmittance(const V3<T> &ro, const V3<T> &rd, T tmin, T tmax, Rng &rng) const {
    T tr = T(1);
    V3<T> unused{};
    for (const phx_volume &v : vols) {
      if (v.majorant <= 0.f || v.tf_size == 0) continue;
      const V3<T> org = to_local_point(v, ro), dir = to_local_vector(v, rd);
      T t0 = tmin, t1 = tmax;
      if (!clip(v, org, dir, t0, t1)) continue;
      track(v, org, dir, t0, t1, rng, false, unused, tr);
      if (tr <= T(0)) return T(0);
    }
    return tr;
  }
};

// Macro-cell majorants (RenderBridge.cpp:656-778).
uint32_t build_majorants(phx_volume &v, const float *values, const float *tf_rgba, uint32_t tf_size,
                         std::vector<float> &out);

}  // namespace orc
