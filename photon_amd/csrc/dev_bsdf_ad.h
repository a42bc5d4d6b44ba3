// Forward-mode derivatives of the Disney BSDF evaluation (f * cos, the
// reflection lobes of bsdf_eval) with respect to the five trainable material
// parameters: base colour r, g, b, GGX roughness and metallic.  Every other
// material parameter and all directions are constants.  The formulas are
// those of dev_bsdf.h (disney_bsdf.h:21-307) carried on dual numbers; the
// path-replay adjoint uses the derivatives only (values come from bsdf_eval).
#pragma once

#include "dev_bsdf.h"

namespace phx {

constexpr int kGradParams = 5;  // base r, g, b, roughness, metallic

struct D5 {
  float v;
  float d[kGradParams];
};
PHX_DEV D5 dconst(float c) {
  D5 r;
  r.v = c;
#pragma unroll
  for (int k = 0; k < kGradParams; ++k) r.d[k] = 0.f;
  return r;
}
PHX_DEV D5 dvar(float c, int k) {
  D5 r = dconst(c);
  r.d[k] = 1.f;
  return r;
}
PHX_DEV D5 operator+(const D5 &a, const D5 &b) {
  D5 r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < kGradParams; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
PHX_DEV D5 operator-(const D5 &a, const D5 &b) {
  D5 r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < kGradParams; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
PHX_DEV D5 operator*(const D5 &a, const D5 &b) {
  D5 r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < kGradParams; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
PHX_DEV D5 operator*(const D5 &a, float b) {
  D5 r;
  r.v = a.v * b;
#pragma unroll
  for (int k = 0; k < kGradParams; ++k) r.d[k] = a.d[k] * b;
  return r;
}
PHX_DEV D5 operator*(float b, const D5 &a) { return a * b; }
PHX_DEV D5 operator+(const D5 &a, float b) {
  D5 r = a;
  r.v = a.v + b;
  return r;
}
PHX_DEV D5 operator+(float b, const D5 &a) { return a + b; }
PHX_DEV D5 operator-(const D5 &a, float b) {
  D5 r = a;
  r.v = a.v - b;
  return r;
}
PHX_DEV D5 operator-(float b, const D5 &a) {
  D5 r;
  r.v = b - a.v;
#pragma unroll
  for (int k = 0; k < kGradParams; ++k) r.d[k] = -a.d[k];
  return r;
}
PHX_DEV D5 operator/(const D5 &a, const D5 &b) {
  D5 r;
  r.v = a.v / b.v;
  const float inv2 = 1.f / (b.v * b.v);
#pragma unroll
  for (int k = 0; k < kGradParams; ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) * inv2;
  return r;
}
PHX_DEV D5 operator/(const D5 &a, float b) {
  D5 r;
  r.v = a.v / b;
#pragma unroll
  for (int k = 0; k < kGradParams; ++k) r.d[k] = a.d[k] / b;
  return r;
}
PHX_DEV D5 operator/(float a, const D5 &b) { return dconst(a) / b; }
PHX_DEV D5 dsqrt(const D5 &a) {
  D5 r;
  r.v = sqrtf(a.v);
  const float h = r.v > 0.f ? 0.5f / r.v : 0.f;
#pragma unroll
  for (int k = 0; k < kGradParams; ++k) r.d[k] = a.d[k] * h;
  return r;
}
// fmaxf(a, c) / fminf(a, c) against a constant
PHX_DEV D5 dmax(const D5 &a, float c) { return a.v >= c ? a : dconst(c); }
PHX_DEV D5 dmin(const D5 &a, float c) { return a.v <= c ? a : dconst(c); }
PHX_DEV D5 dsat(const D5 &a) { return dmin(dmax(a, 0.f), 1.f); }

struct DV3 {
  D5 x, y, z;
};
PHX_DEV DV3 dv3(vec3 v) { return {dconst(v.x), dconst(v.y), dconst(v.z)}; }
PHX_DEV DV3 operator+(const DV3 &a, const DV3 &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
PHX_DEV DV3 operator-(const DV3 &a, const DV3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
PHX_DEV DV3 operator*(const DV3 &a, const D5 &s) { return {a.x * s, a.y * s, a.z * s}; }
PHX_DEV DV3 operator*(const DV3 &a, float s) { return {a.x * s, a.y * s, a.z * s}; }
PHX_DEV DV3 operator*(const DV3 &a, const DV3 &b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
PHX_DEV DV3 dlerp(const DV3 &a, const DV3 &b, const D5 &t) { return a + (b - a) * t; }
PHX_DEV DV3 dlerp(const DV3 &a, const DV3 &b, float t) { return a + (b - a) * t; }
PHX_DEV D5 dlum(const DV3 &c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }

// The trainable parameters of a resolved material as dual numbers.
struct DMat {
  DV3 base;
  D5 rough, metal;
};
PHX_DEV DMat dmat(const phx_material &m) {
  return {{dvar(m.base_color[0], 0), dvar(m.base_color[1], 1), dvar(m.base_color[2], 2)}, dvar(m.roughness, 3),
          dvar(m.metallic, 4)};
}

PHX_DEV D5 D_ggx_d(float NoH, const D5 &a) {
  const D5 a2 = a * a;
  const D5 d = NoH * NoH * (a2 - 1.f) + 1.f;
  return a2 / ((d * d) * kPi);
}
PHX_DEV D5 D_ggx_aniso_d(float NoH, float XoH, float YoH, const D5 &ax, const D5 &ay) {
  const D5 d = (XoH * XoH) / (ax * ax) + (YoH * YoH) / (ay * ay) + NoH * NoH;
  return 1.f / (ax * ay * (d * d) * kPi);
}
PHX_DEV D5 G1_d(float NoV, const D5 &a) {
  const D5 a2 = a * a;
  const D5 den = NoV + dsqrt(a2 + (1.f - a2) * (NoV * NoV));
  return den.v <= 0.f ? dconst(0.f) : (2.f * NoV) / den;
}
PHX_DEV D5 G2_d(float NoV, float NoL, const D5 &a) { return G1_d(NoV, a) * G1_d(NoL, a); }

PHX_DEV void spec_f0_f90_d(const phx_material &m, const DMat &t, DV3 &F0, DV3 &F90) {
  const float scale = spec_scale(m);
  const float opaque = 1.f - saturate(m.transmission);
  const float f0d = fminf(dielectric_f0(m.ior) * scale, 1.f) * opaque;
  DV3 F0d = dv3(vec3{f0d, f0d, f0d});
  if (m.specular_tint > 0.f) {
    const D5 lum = dlum(t.base);
    const DV3 tint = lum.v > 0.f ? t.base * (1.f / lum) : dv3(vec3{1.f, 1.f, 1.f});
    F0d = dlerp(F0d, tint * f0d, m.specular_tint);
  }
  F0 = dlerp(F0d, t.base, t.metal);
  const float f90d = fminf(scale, 1.f) * opaque;
  F90 = dlerp(dv3(vec3{f90d, f90d, f90d}), dv3(vec3{1.f, 1.f, 1.f}), t.metal);
}

PHX_DEV DV3 eval_diffuse_d(const phx_material &m, const DMat &t, float NoL, float NoV) {
  const D5 w = (1.f - t.metal) * (1.f - m.transmission) * diffuse_scale(m, NoV);
  if (w.v <= 0.f) return dv3(vec3{0.f, 0.f, 0.f});
  if (m.subsurface > 0.f) {
    const DV3 dif = t.base * (w * kInvPi * NoL);
    const float mm = 1.f - NoL;
    const float Fss = 1.f + (0.5f - 1.f) * mm * mm * mm * mm * mm;
    const float ss = 1.25f * (Fss * (1.f / (NoL + 0.001f) - 0.5f) + 0.5f);
    const DV3 ssc = dlerp(t.base, dv3(ld3(m.subsurface_color)), m.subsurface);
    return dlerp(dif, ssc * (w * kInvPi * NoL * ss), m.subsurface);
  }
  return t.base * (w * kInvPi * NoL);
}
PHX_DEV DV3 eval_spec_d(const phx_material &m, const DMat &t, vec3 wo, vec3 wi, vec3 n, const D5 &a) {
  const float NoV = saturate(dot(n, wo)), NoL = saturate(dot(n, wi));
  if (NoV <= 0.f || NoL <= 0.f) return dv3(vec3{0.f, 0.f, 0.f});
  const vec3 h = normalize(wi + wo);
  const float NoH = saturate(dot(n, h)), VoH = saturate(dot(wo, h));
  DV3 F0, F90;
  spec_f0_f90_d(m, t, F0, F90);
  const float m5 = schlick_m5(VoH);
  const DV3 F = F0 + (F90 - F0) * m5;
  D5 D, G;
  if (m.anisotropic > 0.f) {
    const float aspect = sqrtf(1.f - 0.9f * m.anisotropic);
    const D5 ax = dmax(a / aspect, 0.001f), ay = dmax(a * aspect, 0.001f);
    vec3 tt, b;
    onb(n, tt, b);
    D = D_ggx_aniso_d(NoH, dot(tt, h), dot(b, h), ax, ay);
    G = dsqrt(G2_d(NoV, NoL, ax) * G2_d(NoV, NoL, ay));
  } else {
    D = D_ggx_d(NoH, a);
    G = G2_d(NoV, NoL, a);
  }
  return F * (D * G / (4.f * fmaxf(NoV, 1e-7f)));
}
PHX_DEV DV3 eval_sheen_d(const phx_material &m, const DMat &t, vec3 wo, vec3 wi, vec3 n) {
  if (m.sheen <= 0.f || t.metal.v >= 1.f) return dv3(vec3{0.f, 0.f, 0.f});
  const vec3 h = normalize(wi + wo);
  const float HoL = fabsf(dot(h, wi));
  const D5 lum = dlum(t.base);
  const DV3 tint = lum.v > 0.f ? t.base * (1.f / lum) : dv3(vec3{1.f, 1.f, 1.f});
  const DV3 cs = dlerp(dv3(vec3{1.f, 1.f, 1.f}), tint, m.sheen_tint);
  const float mm = 1.f - HoL, m2 = mm * mm, m5 = m2 * m2 * mm;
  const float NoL = saturate(dot(n, wi));
  const D5 w = (1.f - t.metal) * (1.f - m.transmission);
  return cs * (m.sheen * m5 * w * NoL);
}

// f * cos of bsdf_eval (disney_bsdf.h:288-307) and its parameter derivatives.
PHX_DEV DV3 bsdf_eval_d(const phx_material &m, vec3 wo, vec3 wi, vec3 n, vec3 sn) {
  if (dot(wo, n) <= 0.f || dot(wi, n) <= 0.f) return dv3(vec3{0.f, 0.f, 0.f});
  const float NoL = saturate(dot(sn, wi)), NoV = saturate(dot(sn, wo));
  if (NoL <= 0.f || NoV <= 0.f) return dv3(vec3{0.f, 0.f, 0.f});
  const DMat t = dmat(m);
  const D5 a = dmax(t.rough, 0.001f);
  DV3 f = eval_diffuse_d(m, t, NoL, NoV);
  if (!spec_delta(m)) f = f + eval_spec_d(m, t, wo, wi, sn, a);
  f = f + eval_sheen_d(m, t, wo, wi, sn);
  f = f + dv3(eval_clearcoat(m, wo, wi, sn));
  return f;
}

}  // namespace phx
