// ORACLE - test infrastructure only.  A CPU restatement of the reference's
// path-tracing semantics that the parity tests, the smoke check and the CPU
// baseline leg of bench.py use as the checker.  Nothing in the product links,
// loads or calls it.  Each function cites the reference file:line it follows
// (paths relative to the reference repository root).
//
// The code is templated on the scalar type: T = float restates the reference
// bit-for-bit where its arithmetic is IEEE single (no FMA contraction: build
// with -ffp-contract=off, as the reference's x86-64 Release build has none);
// T = double is the FP64 engine of the finite-difference gradient checks.
//
// Pinning (tests/test_oracle.py, tests/test_oracle_pins.py): RNG, hash,
// power heuristic, ONB, triangle test, reference BVH (against brute force),
// Disney special cases (Lambert, metal GGX, sheen, clearcoat, anisotropy
// limit, the subsurface term at NoL = 1, 1/2 and 1/4), the lobe weights
// (hand-derived known answers, and the lobe each draw picks), the dielectric
// lobe (Fresnel coin against the textbook reflectance, Snell's law, total
// internal reflection, Schlick pass-through), the reference's light and
// texture test cases plus closed forms, samplers, parameter sources and
// normal maps, the background image, env map, camera, majorants, the exact
// draws of delta and ratio tracking, and integrator closed forms.
// Parity unpinned: transcendentals at the ulp level (the reference goes
// through Kokkos 5.0.1 math to the platform libm; here host libm).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/phx.h"

namespace orc {

using u32 = uint32_t;

template <class T> struct V3 {
  T x{0}, y{0}, z{0};
  V3() = default;
  V3(T a, T b, T c) : x(a), y(b), z(c) {}
  template <class S> explicit V3(const V3<S> &o) : x(T(o.x)), y(T(o.y)), z(T(o.z)) {}
};
template <class T> struct V2 { T x{0}, y{0}; };
template <class T> struct V4 { T x{0}, y{0}, z{0}, w{0}; };

// math.h:14-112
template <class T> inline V3<T> operator+(V3<T> a, const V3<T> &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(const V3<T> &a, const V3<T> &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(const V3<T> &a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(const V3<T> &a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, const V3<T> &a) { return a * s; }
template <class T> inline V3<T> operator*(const V3<T> &a, const V3<T> &b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline T dot(const V3<T> &a, const V3<T> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(const V3<T> &a, const V3<T> &b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T length(const V3<T> &v) { return std::sqrt(dot(v, v)); }
// math.h:56-60: multiply by the reciprocal, zero vector stays zero.
template <class T> inline V3<T> normalize(const V3<T> &v) {
  const T len = length(v);
  return len > T(0) ? (v * (T(1) / len)) : V3<T>{};
}
// math.h:102-105
template <class T> inline V3<T> reflect(const V3<T> &I, const V3<T> &N) { return I - (T(2) * dot(I, N)) * N; }
// math.h:107-114
template <class T> inline V3<T> refract(const V3<T> &I, const V3<T> &N, T eta) {
  const T cosi = -dot(I, N);
  const T k = T(1) - eta * eta * (T(1) - cosi * cosi);
  if (k < T(0)) return {};
  return eta * I + (eta * cosi - std::sqrt(k)) * N;
}
// math.h:116-119
template <class T> inline V3<T> lerp(const V3<T> &a, const V3<T> &b, T t) { return a + (b - a) * t; }
template <class T> inline T fmaxT(T a, T b) { return std::fmax(a, b); }
template <class T> inline T fminT(T a, T b) { return std::fmin(a, b); }
template <class T> inline T luminance(const V3<T> &c) { return T(0.2126) * c.x + T(0.7152) * c.y + T(0.0722) * c.z; }
template <class T> inline V3<T> v3(const float *p) { return {T(p[0]), T(p[1]), T(p[2])}; }

template <class T> struct K {
  static constexpr T PI = T(3.14159265358979323846);
  static constexpr T INV_PI = T(1) / PI;
  static constexpr T TWO_PI = T(2) * PI;
  static constexpr T INV_TWO_PI = T(1) / TWO_PI;
};
// sampling.h:10-13 (f32 constants): PI is the float nearest pi, and 1/PI etc.
// are computed in float.  For T = float the K<float> values equal them.

// ---- RNG (rng.h:9-30) and per-sample streams (pathtracer.cpp:20-34) -------
struct Rng {
  u32 state{1u};
  explicit Rng(u32 seed) : state(seed ? seed : 1u) {}
  u32 next_u32() {
    u32 x = state;
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    state = x;
    return x;
  }
  // rng.h:25-29: top 24 bits, exactly representable in float.
  float next_f32() { return float(next_u32() >> 8) * (1.0f / 16777216.0f); }
};
inline u32 hash_u32(u32 x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}
inline Rng make_rng(u32 pixel, u32 sample, u32 dim) {
  return Rng(hash_u32(pixel + hash_u32(sample + hash_u32(dim + 0x9e3779b9U))));
}

// Random draws are produced in float by the reference; the double engine
// consumes the same float draws widened.
template <class T> inline T draw(Rng &r) { return T(r.next_f32()); }

// ---- sampling.h --------------------------------------------------------------
// sampling.h:16-23
template <class T> inline T power_heuristic(int nf, T fPdf, int ng, T gPdf) {
  const T f = T(nf) * fPdf, g = T(ng) * gPdf;
  const T f2 = f * f, g2 = g * g;
  return (f2 + g2) > T(0) ? f2 / (f2 + g2) : T(0);
}
// sampling.h:25-40 (Duff et al. branchless ONB, then renormalized)
template <class T> inline void onb_from_normal(const V3<T> &n, V3<T> &t, V3<T> &b) {
  if (n.z < T(-0.9999999f)) {
    t = {T(0), T(-1), T(0)};
    b = {T(-1), T(0), T(0)};
    return;
  }
  const T a = T(1) / (T(1) + n.z);
  const T bb = -n.x * n.y * a;
  t = {T(1) - n.x * n.x * a, bb, -n.x};
  b = {bb, T(1) - n.y * n.y * a, -n.y};
  t = normalize(t);
  b = normalize(b);
}
// sampling.h:48-51
template <class T> inline T cosine_hemisphere_pdf(T c) { return fmaxT(c, T(0)) * K<T>::INV_PI; }
// sampling.h:53-76 (Shirley-Chiu concentric map)
template <class T> inline V2<T> sample_concentric_disk(Rng &rng) {
  const T u = T(2) * draw<T>(rng) - T(1);
  const T v = T(2) * draw<T>(rng) - T(1);
  if (u == T(0) && v == T(0)) return {T(0), T(0)};
  T r, theta;
  if (std::fabs(u) > std::fabs(v)) {
    r = std::fabs(u);
    theta = (K<T>::PI / T(4)) * (v / u);
  } else {
    r = std::fabs(v);
    theta = (K<T>::PI / T(2)) - (K<T>::PI / T(4)) * (u / v);
  }
  return {r * std::cos(theta), r * std::sin(theta)};
}
template <class T> struct DirPdf { V3<T> dir; T pdf{0}; };
// sampling.h:78-92
template <class T> inline DirPdf<T> sample_cosine_hemisphere(const V3<T> &n, Rng &rng) {
  V3<T> t, b;
  onb_from_normal(n, t, b);
  const V2<T> d = sample_concentric_disk<T>(rng);
  const T z = std::sqrt(fmaxT(T(0), T(1) - d.x * d.x - d.y * d.y));
  V3<T> dir = d.x * t + d.y * b + z * n;
  dir = normalize(dir);
  return {dir, cosine_hemisphere_pdf(dot(dir, n))};
}
// sampling.h:94-100
template <class T> inline V3<T> sample_uniform_sphere(Rng &rng) {
  const T z = T(1) - T(2) * draw<T>(rng);
  const T r = std::sqrt(fmaxT(T(0), T(1) - z * z));
  const T phi = K<T>::TWO_PI * draw<T>(rng);
  return {r * std::cos(phi), r * std::sin(phi), z};
}

// ---- Material (material.h:35-103) --------------------------------------------
template <class T> struct Mat {
  V3<T> base_color{T(0.8), T(0.8), T(0.8)};
  T metallic{0}, roughness{T(0.5)}, ior{T(1.5)}, transmission{0}, specular{T(0.5)};
  T clearcoat{0}, clearcoat_roughness{T(0.03)}, sheen{0}, sheen_tint{T(0.5)}, specular_tint{0}, anisotropic{0};
  V3<T> emission{};
  T emission_strength{0}, subsurface{0};
  V3<T> subsurface_color{T(1), T(1), T(1)};
  T alpha{1};
  int thin{0}, flags{0}, alpha_mode{0};
  T alpha_cutoff{T(0.5)};
  int base_color_tex{-1}, normal_tex{-1}, roughness_tex{-1}, metallic_tex{-1}, emission_tex{-1}, alpha_tex{-1};
  // material.h:62-81
  phx_param_source base_color_src{}, alpha_src{}, metallic_src{}, roughness_src{}, emission_src{}, transmission_src{},
      normal_src{};
  int base_color_alpha{0}, roughness_src_squared{0};
};
template <class T> inline Mat<T> mat_from(const phx_material &m) {
  Mat<T> o;
  o.base_color = v3<T>(m.base_color);
  o.metallic = T(m.metallic);
  o.roughness = T(m.roughness);
  o.ior = T(m.ior);
  o.transmission = T(m.transmission);
  o.specular = T(m.specular);
  o.clearcoat = T(m.clearcoat);
  o.clearcoat_roughness = T(m.clearcoat_roughness);
  o.sheen = T(m.sheen);
  o.sheen_tint = T(m.sheen_tint);
  o.specular_tint = T(m.specular_tint);
  o.anisotropic = T(m.anisotropic);
  o.emission = v3<T>(m.emission);
  o.emission_strength = T(m.emission_strength);
  o.subsurface = T(m.subsurface);
  o.subsurface_color = v3<T>(m.subsurface_color);
  o.alpha = T(m.alpha);
  o.thin = m.thin;
  o.flags = m.flags;
  o.alpha_mode = m.alpha_mode;
  o.alpha_cutoff = T(m.alpha_cutoff);
  o.base_color_tex = m.base_color_tex;
  o.normal_tex = m.normal_tex;
  o.roughness_tex = m.roughness_tex;
  o.metallic_tex = m.metallic_tex;
  o.emission_tex = m.emission_tex;
  o.alpha_tex = m.alpha_tex;
  o.base_color_src = m.base_color_src;
  o.alpha_src = m.alpha_src;
  o.metallic_src = m.metallic_src;
  o.roughness_src = m.roughness_src;
  o.emission_src = m.emission_src;
  o.transmission_src = m.transmission_src;
  o.normal_src = m.normal_src;
  o.base_color_alpha = m.base_color_alpha;
  o.roughness_src_squared = m.roughness_src_squared;
  return o;
}
// material.h:84-87
template <class T> inline V3<T> material_emission(const Mat<T> &m) { return m.emission * m.emission_strength; }
// material.h:97-103
template <class T> inline bool material_may_emit(const Mat<T> &m) {
  if (m.emission_strength <= T(0)) return false;
  return m.emission_src.kind != PHX_SOURCE_CONSTANT || m.emission_tex >= 0 || m.emission.x > T(0) ||
         m.emission.y > T(0) || m.emission.z > T(0);
}

// ---- Disney BSDF (disney_bsdf.h) ---------------------------------------------
template <class T> struct BsdfSample {
  V3<T> wi, f;
  T pdf{0};
  bool is_specular{false}, is_diffuse{false};
};
template <class T> inline T saturate(T x) { return fminT(T(1), fmaxT(T(0), x)); }          // :21-24
template <class T> inline T dielectric_f0(T ior) { const T a = (T(1) - ior) / (T(1) + ior); return a * a; }  // :26-30
template <class T> inline V3<T> fresnel_schlick(T c, V3<T> F0) {                              // :32-38
  const T m = T(1) - saturate(c), m2 = m * m, m5 = m2 * m2 * m;
  return F0 + (V3<T>{T(1), T(1), T(1)} - F0) * m5;
}
template <class T> inline V3<T> fresnel_schlick(T c, V3<T> F0, V3<T> F90) {                   // :40-46
  const T m = T(1) - saturate(c), m2 = m * m, m5 = m2 * m2 * m;
  return F0 + (F90 - F0) * m5;
}
template <class T> inline T fresnel_dielectric(T cos_i, T eta) {                              // :49-59
  cos_i = saturate(cos_i);
  const T sin2_t = eta * eta * (T(1) - cos_i * cos_i);
  if (sin2_t >= T(1)) return T(1);
  const T cos_t = std::sqrt(T(1) - sin2_t);
  const T r_par = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
  const T r_perp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
  return T(0.5) * (r_par * r_par + r_perp * r_perp);
}
template <class T> inline T D_ggx(T NoH, T a) {                                               // :61-66
  const T a2 = a * a;
  const T d = NoH * NoH * (a2 - T(1)) + T(1);
  return a2 / (K<T>::PI * d * d);
}
template <class T> inline T D_ggx_aniso(T NoH, T XoH, T YoH, T ax, T ay) {                   // :68-72
  const T d = (XoH * XoH) / (ax * ax) + (YoH * YoH) / (ay * ay) + NoH * NoH;
  return T(1) / (K<T>::PI * ax * ay * d * d);
}
template <class T> inline void aniso_roughness(T r, T an, T &ax, T &ay) {                     // :74-79
  const T aspect = std::sqrt(T(1) - T(0.9) * an);
  ax = fmaxT(r / aspect, T(0.001));
  ay = fmaxT(r * aspect, T(0.001));
}
template <class T> inline T G1_smith(T NoV, T a) {                                            // :81-88
  const T a2 = a * a;
  const T denom = NoV + std::sqrt(a2 + (T(1) - a2) * NoV * NoV);
  if (denom <= T(0)) return T(0);
  return (T(2) * NoV) / denom;
}
template <class T> inline T G_smith(T NoV, T NoL, T a) { return G1_smith(NoV, a) * G1_smith(NoL, a); }  // :90-93
template <class T> inline T ggx_pdf_reflection(T NoH, T VoH, T a) {                           // :95-100
  const T D = D_ggx(NoH, a);
  const T denom = T(4) * fmaxT(VoH, T(1e-7));
  return D * NoH / denom;
}
template <class T> inline T disney_specular_scale(const Mat<T> &m) { return fmaxT(T(2) * m.specular, T(0)); }  // :104-107
template <class T> inline void disney_specular_f0_f90(const Mat<T> &m, V3<T> &F0, V3<T> &F90) {  // :111-125
  const T scale = disney_specular_scale(m);
  const T opaque = T(1) - saturate(m.transmission);
  const T f0d = fminT(dielectric_f0(m.ior) * scale, T(1)) * opaque;
  V3<T> F0d{f0d, f0d, f0d};
  if (m.specular_tint > T(0)) {
    const T lum = T(0.2126) * m.base_color.x + T(0.7152) * m.base_color.y + T(0.0722) * m.base_color.z;
    const V3<T> c_tint = lum > T(0) ? m.base_color * (T(1) / lum) : V3<T>{T(1), T(1), T(1)};
    F0d = lerp(F0d, c_tint * f0d, m.specular_tint);
  }
  F0 = lerp(F0d, m.base_color, m.metallic);
  const T f90d = fminT(scale, T(1)) * opaque;
  F90 = lerp(V3<T>{f90d, f90d, f90d}, V3<T>{T(1), T(1), T(1)}, m.metallic);
}
template <class T> inline T disney_diffuse_scale(const Mat<T> &m, T NoV) {                   // :129-137
  const T scale = disney_specular_scale(m);
  const T f0d = fminT(dielectric_f0(m.ior) * scale, T(1));
  const T f90d = fminT(scale, T(1));
  const T mm = T(1) - saturate(NoV), m2 = mm * mm;
  return T(1) - (f0d + (f90d - f0d) * m2 * m2 * mm);
}
template <class T> inline bool disney_specular_is_delta(const Mat<T> &m) { return m.roughness < T(0.001); }  // :141-144
template <class T>
inline void disney_lobe_weights(const Mat<T> &m, T NoV, T &wd, T &ws, T &wt, T &wc) {       // :148-175
  const T opaque = T(1) - saturate(m.transmission);
  const T dielectric = T(1) - saturate(m.metallic);
  const T diffuse_energy = dielectric * opaque * disney_diffuse_scale(m, NoV);
  const T base_lum = T(0.2126) * m.base_color.x + T(0.7152) * m.base_color.y + T(0.0722) * m.base_color.z;
  wd = diffuse_energy > T(0) ? diffuse_energy * fmaxT(base_lum, T(0.05)) : T(0);
  if (m.sheen > T(0) && dielectric * opaque > T(0)) wd += T(0.1) * m.sheen;
  V3<T> F0, F90;
  disney_specular_f0_f90(m, F0, F90);
  const V3<T> F = fresnel_schlick(NoV, F0, F90);
  const T spec = T(0.2126) * F.x + T(0.7152) * F.y + T(0.0722) * F.z;
  ws = spec > T(0) ? fmaxT(spec, T(0.05)) : T(0);
  wt = saturate(m.transmission) * dielectric;
  wc = T(0);
  if (m.clearcoat > T(0)) {
    const T mm = T(1) - saturate(NoV), m2 = mm * mm;
    wc = fmaxT(m.clearcoat * T(0.25) * (T(0.04) + T(0.96) * m2 * m2 * mm), T(0.02));
  }
}
template <class T> inline V3<T> eval_diffuse(const Mat<T> &m, T NoL, T NoV) {                // :177-196
  const T w = (T(1) - m.metallic) * (T(1) - m.transmission) * disney_diffuse_scale(m, NoV);
  if (w <= T(0)) return {};
  if (m.subsurface > T(0)) {
    const V3<T> diffuse = m.base_color * (w * K<T>::INV_PI * NoL);
    const T Fd90 = T(0.5);
    const T mm = T(1) - NoL;
    const T Fss = T(1) + (Fd90 - T(1)) * mm * mm * mm * mm * mm;
    const T ss = T(1.25) * (Fss * (T(1) / (NoL + T(0.001)) - T(0.5)) + T(0.5));
    const V3<T> ss_color = lerp(m.base_color, m.subsurface_color, m.subsurface);
    const V3<T> sst = ss_color * (w * K<T>::INV_PI * NoL * ss);
    return lerp(diffuse, sst, m.subsurface);
  }
  return m.base_color * (w * K<T>::INV_PI * NoL);
}
template <class T>
inline V3<T> eval_specular_reflection(const Mat<T> &m, const V3<T> &wo, const V3<T> &wi, const V3<T> &n, T a) {  // :200-234
  const T NoV = saturate(dot(n, wo)), NoL = saturate(dot(n, wi));
  if (NoV <= T(0) || NoL <= T(0)) return {};
  const V3<T> h = normalize(wi + wo);
  const T NoH = saturate(dot(n, h)), VoH = saturate(dot(wo, h));
  V3<T> F0, F90;
  disney_specular_f0_f90(m, F0, F90);
  const V3<T> F = fresnel_schlick(VoH, F0, F90);
  T D, G;
  if (m.anisotropic > T(0)) {
    T ax, ay;
    aniso_roughness(a, m.anisotropic, ax, ay);
    V3<T> t, b;
    onb_from_normal(n, t, b);
    D = D_ggx_aniso(NoH, dot(t, h), dot(b, h), ax, ay);
    G = G_smith(NoV, NoL, ax) * G_smith(NoV, NoL, ay);
    G = std::sqrt(G);
  } else {
    D = D_ggx(NoH, a);
    G = G_smith(NoV, NoL, a);
  }
  const T denom = T(4) * fmaxT(NoV, T(1e-7));
  return F * (D * G / denom);
}
template <class T>
inline V3<T> eval_sheen(const Mat<T> &m, const V3<T> &wo, const V3<T> &wi, const V3<T> &n) {  // :236-256
  if (m.sheen <= T(0) || m.metallic >= T(1)) return {};
  const V3<T> h = normalize(wi + wo);
  const T HoL = std::fabs(dot(h, wi));
  const T lum = T(0.2126) * m.base_color.x + T(0.7152) * m.base_color.y + T(0.0722) * m.base_color.z;
  const V3<T> c_tint = lum > T(0) ? m.base_color * (T(1) / lum) : V3<T>{T(1), T(1), T(1)};
  const V3<T> c_sheen = lerp(V3<T>{T(1), T(1), T(1)}, c_tint, m.sheen_tint);
  const T mm = T(1) - HoL, m2 = mm * mm, m5 = m2 * m2 * mm;
  const T NoL = saturate(dot(n, wi));
  const T w = (T(1) - m.metallic) * (T(1) - m.transmission);
  return c_sheen * (m.sheen * m5 * w * NoL);
}
template <class T>
inline V3<T> eval_clearcoat(const Mat<T> &m, const V3<T> &wo, const V3<T> &wi, const V3<T> &n) {  // :258-283
  if (m.clearcoat <= T(0)) return {};
  const T NoV = saturate(dot(n, wo)), NoL = saturate(dot(n, wi));
  if (NoV <= T(0) || NoL <= T(0)) return {};
  const T a = fmaxT(m.clearcoat_roughness, T(0.001));
  const V3<T> h = normalize(wi + wo);
  const T NoH = saturate(dot(n, h)), VoH = saturate(dot(wo, h));
  const T F0 = dielectric_f0(T(1.5));
  const V3<T> F = fresnel_schlick(VoH, V3<T>{F0, F0, F0});
  const T D = D_ggx(NoH, a), G = G_smith(NoV, NoL, a);
  const T denom = T(4) * fmaxT(NoV, T(1e-7));
  const T w = m.clearcoat * T(0.25);
  return F * (w * D * G / denom);
}
// :288-307 - f * cos(theta_i) of the reflection lobes.
template <class T>
inline V3<T> disney_bsdf_eval(const Mat<T> &m, const V3<T> &wo, const V3<T> &wi, const V3<T> &n, const V3<T> &sn) {
  if (dot(wo, n) <= T(0) || dot(wi, n) <= T(0)) return {};
  const T NoL = saturate(dot(sn, wi)), NoV = saturate(dot(sn, wo));
  if (NoL <= T(0) || NoV <= T(0)) return {};
  const T a = fmaxT(m.roughness, T(0.001));
  V3<T> f = eval_diffuse(m, NoL, NoV);
  if (!disney_specular_is_delta(m)) f = f + eval_specular_reflection(m, wo, wi, sn, a);
  f = f + eval_sheen(m, wo, wi, sn);
  f = f + eval_clearcoat(m, wo, wi, sn);
  return f;
}
// :310-355
template <class T>
inline T disney_bsdf_pdf(const Mat<T> &m, const V3<T> &wo, const V3<T> &wi, const V3<T> &n, const V3<T> &sn) {
  if (dot(wo, n) <= T(0) || dot(wi, n) <= T(0)) return T(0);
  const T NoL = saturate(dot(sn, wi)), NoV = saturate(dot(sn, wo));
  if (NoL <= T(0) || NoV <= T(0)) return T(0);
  const T a = fmaxT(m.roughness, T(0.001));
  T wd, ws, wt, wc;
  disney_lobe_weights(m, NoV, wd, ws, wt, wc);
  const T sum = wd + ws + wt + wc;
  if (sum <= T(0)) return T(0);
  const T pd = wd / sum, ps = ws / sum, pc = wc / sum;
  T pdf = T(0);
  if (pd > T(0)) pdf += pd * (NoL * K<T>::INV_PI);
  if (ps > T(0) && !disney_specular_is_delta(m)) {
    const V3<T> h = normalize(wi + wo);
    pdf += ps * ggx_pdf_reflection(saturate(dot(sn, h)), saturate(dot(wo, h)), a);
  }
  if (pc > T(0)) {
    const V3<T> h = normalize(wi + wo);
    const T acc = fmaxT(m.clearcoat_roughness, T(0.001));
    pdf += pc * ggx_pdf_reflection(saturate(dot(sn, h)), saturate(dot(wo, h)), acc);
  }
  return fmaxT(pdf, T(0));
}
// :357-375
template <class T> inline V3<T> sample_ggx_half_vector(const V3<T> &n, T a, Rng &rng) {
  V3<T> t, b;
  onb_from_normal(n, t, b);
  const T r1 = draw<T>(rng), r2 = draw<T>(rng);
  const T a2 = a * a;
  const T ct = std::sqrt((T(1) - r1) / (T(1) + (a2 - T(1)) * r1));
  const T st = std::sqrt(fmaxT(T(0), T(1) - ct * ct));
  const T phi = T(2) * K<T>::PI * r2;
  return normalize(t * (st * std::cos(phi)) + b * (st * std::sin(phi)) + n * ct);
}
// :381-521
template <class T>
inline BsdfSample<T> disney_bsdf_sample(const Mat<T> &m, const V3<T> &wo, const V3<T> &n, const V3<T> &sn, Rng &rng) {
  BsdfSample<T> out;
  const bool front = dot(wo, n) > T(0);
  if (!front && m.transmission <= T(0)) return out;
  const T rough = fmaxT(m.roughness, T(0.001));
  const T NoV = saturate(std::fabs(dot(sn, wo)));
  T wd, ws, wt, wc, te = T(1);
  if (!front) {
    wd = T(0); ws = T(0); wt = T(1); wc = T(0);
  } else {
    disney_lobe_weights(m, NoV, wd, ws, wt, wc);
    te = wt;
  }
  const T sum = wd + ws + wt + wc;
  if (sum <= T(0)) return out;
  const T pd = wd / sum, ps = ws / sum, pt = wt / sum, pc = wc / sum;
  const T xi = draw<T>(rng);
  if (xi < pd) {
    const DirPdf<T> s = sample_cosine_hemisphere(sn, rng);
    out.is_diffuse = true;
    out.wi = s.dir;
    out.f = disney_bsdf_eval(m, wo, out.wi, n, sn);
    out.pdf = disney_bsdf_pdf(m, wo, out.wi, n, sn);
    return out;
  }
  if (xi < pd + ps) {
    if (disney_specular_is_delta(m)) {
      out.wi = reflect(-wo, sn);
      if (dot(out.wi, n) <= T(0)) out.wi = reflect(-wo, n);
      if (dot(out.wi, n) <= T(0)) return BsdfSample<T>{};
      V3<T> F0, F90;
      disney_specular_f0_f90(m, F0, F90);
      const V3<T> F = fresnel_schlick(saturate(dot(wo, sn)), F0, F90);
      out.f = F * (T(1) / ps);
      out.pdf = T(1);
      out.is_specular = true;
      return out;
    }
    const V3<T> h = sample_ggx_half_vector(sn, rough, rng);
    out.wi = reflect(-wo, h);
    if (dot(out.wi, n) <= T(0)) return BsdfSample<T>{};
    out.f = disney_bsdf_eval(m, wo, out.wi, n, sn);
    out.pdf = disney_bsdf_pdf(m, wo, out.wi, n, sn);
    return out;
  }
  if (xi < pd + ps + pt && m.transmission > T(0)) {
    const V3<T> nn = front ? n : n * T(-1);
    V3<T> snf = front ? sn : sn * T(-1);
    if (dot(snf, wo) <= T(0)) snf = nn;
    V3<T> h = snf;
    if (m.roughness >= T(0.001)) {
      h = sample_ggx_half_vector(snf, rough, rng);
      if (dot(h, wo) <= T(0)) h = snf;
    }
    const bool refractive = (m.flags & PHX_MATERIAL_FLAG_REFRACTIVE) != 0 && m.thin == 0;
    const T ior = fmaxT(m.ior, T(1.0001));
    const T eta = refractive ? (front ? (T(1) / ior) : ior) : T(1);
    const T cos_i = saturate(dot(wo, h));
    T F;
    if (refractive) {
      F = fresnel_dielectric(cos_i, eta);
    } else {
      const T f0 = dielectric_f0(m.ior);
      F = f0 + (T(1) - f0) * std::pow(T(1) - cos_i, T(5));
    }
    const T lobe_weight = te / pt;
    out.pdf = T(1);
    out.is_specular = true;
    if (draw<T>(rng) < F) {
      out.wi = reflect(-wo, h);
      if (dot(out.wi, nn) <= T(0)) out.wi = reflect(-wo, nn);
      if (dot(out.wi, nn) <= T(0)) return BsdfSample<T>{};
      out.f = V3<T>{T(1), T(1), T(1)} * lobe_weight;
      return out;
    }
    if (refractive) {
      out.wi = normalize(refract(-wo, h, eta));
      if (dot(out.wi, nn) >= T(0)) out.wi = normalize(refract(-wo, snf, eta));
      if (dot(out.wi, nn) >= T(0)) return BsdfSample<T>{};
    } else {
      out.wi = normalize(-wo);
    }
    out.f = m.base_color * lobe_weight;
    return out;
  }
  if (pc > T(0)) {
    const T acc = fmaxT(m.clearcoat_roughness, T(0.001));
    const V3<T> h = sample_ggx_half_vector(sn, acc, rng);
    out.wi = reflect(-wo, h);
    if (dot(out.wi, n) <= T(0)) return BsdfSample<T>{};
    out.f = disney_bsdf_eval(m, wo, out.wi, n, sn);
    out.pdf = disney_bsdf_pdf(m, wo, out.wi, n, sn);
    return out;
  }
  return out;
}

}  // namespace orc
