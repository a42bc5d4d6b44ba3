// Disney BSDF of the gfx950 path tracer: eval (f * cos), pdf and sample.
//
// Lobes and conventions of the reference's disney_bsdf.h:21-521: diffuse with
// the dielectric-specular energy split and an optional subsurface blend, GGX
// reflection (anisotropic D and G in eval only), sheen, a GGX clearcoat, and a
// delta dielectric boundary sampled around a GGX microfacet normal.  The GGX
// alpha is Material::roughness itself (clamped to 0.001).
#pragma once

#include "../../include/phx.h"
#include "dev_math.h"

namespace phx {

struct BsdfSample {
  vec3 wi;
  vec3 f;
  float pdf;
  bool specular;
  bool diffuse;
};

PHX_DEV float dielectric_f0(float ior) {
  const float a = (1.f - ior) / (1.f + ior);
  return a * a;
}
PHX_DEV float schlick_m5(float c) {
  const float m = 1.f - saturate(c), m2 = m * m;
  return m2 * m2 * m;
}
PHX_DEV float fresnel_dielectric(float cos_i, float eta) {
  cos_i = saturate(cos_i);
  const float sin2_t = eta * eta * (1.f - cos_i * cos_i);
  if (sin2_t >= 1.f) return 1.f;
  const float cos_t = sqrtf(1.f - sin2_t);
  const float rp = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
  const float rs = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
  return 0.5f * (rp * rp + rs * rs);
}
PHX_DEV float D_ggx(float NoH, float a) {
  const float a2 = a * a;
  const float d = NoH * NoH * (a2 - 1.f) + 1.f;
  return a2 / (kPi * d * d);
}
PHX_DEV float D_ggx_aniso(float NoH, float XoH, float YoH, float ax, float ay) {
  const float d = (XoH * XoH) / (ax * ax) + (YoH * YoH) / (ay * ay) + NoH * NoH;
  return 1.f / (kPi * ax * ay * d * d);
}
PHX_DEV float G1(float NoV, float a) {
  const float a2 = a * a;
  const float den = NoV + sqrtf(a2 + (1.f - a2) * NoV * NoV);
  return den <= 0.f ? 0.f : (2.f * NoV) / den;
}
PHX_DEV float G2(float NoV, float NoL, float a) { return G1(NoV, a) * G1(NoL, a); }
PHX_DEV float ggx_pdf(float NoH, float VoH, float a) { return D_ggx(NoH, a) * NoH / (4.f * fmaxf(VoH, 1e-7f)); }

PHX_DEV float spec_scale(const phx_material &m) { return fmaxf(2.f * m.specular, 0.f); }
PHX_DEV vec3 base_of(const phx_material &m) { return ld3(m.base_color); }

PHX_DEV void spec_f0_f90(const phx_material &m, vec3 &F0, vec3 &F90) {
  const float scale = spec_scale(m);
  const float opaque = 1.f - saturate(m.transmission);
  const float f0d = fminf(dielectric_f0(m.ior) * scale, 1.f) * opaque;
  const vec3 base = base_of(m);
  vec3 F0d = {f0d, f0d, f0d};
  if (m.specular_tint > 0.f) {
    const float lum = luminance(base);
    const vec3 tint = lum > 0.f ? base * (1.f / lum) : vec3{1.f, 1.f, 1.f};
    F0d = lerp(F0d, tint * f0d, m.specular_tint);
  }
  F0 = lerp(F0d, base, m.metallic);
  const float f90d = fminf(scale, 1.f) * opaque;
  F90 = lerp(vec3{f90d, f90d, f90d}, vec3{1.f, 1.f, 1.f}, m.metallic);
}
PHX_DEV float diffuse_scale(const phx_material &m, float NoV) {
  const float scale = spec_scale(m);
  const float f0d = fminf(dielectric_f0(m.ior) * scale, 1.f);
  const float f90d = fminf(scale, 1.f);
  const float mm = 1.f - saturate(NoV), m2 = mm * mm;
  return 1.f - (f0d + (f90d - f0d) * m2 * m2 * mm);  // product order of disney_bsdf.h:141
}
PHX_DEV bool spec_delta(const phx_material &m) { return m.roughness < 0.001f; }

struct LobeW {
  float d, s, t, c;
};
// Lobe selection weights; F0 and F90 are spec_f0_f90(m).
PHX_DEV LobeW lobe_weights(const phx_material &m, float NoV, vec3 F0, vec3 F90) {
  LobeW w;
  const float opaque = 1.f - saturate(m.transmission);
  const float diel = 1.f - saturate(m.metallic);
  const float de = diel * opaque * diffuse_scale(m, NoV);
  const vec3 base = base_of(m);
  w.d = de > 0.f ? de * fmaxf(luminance(base), 0.05f) : 0.f;
  if (m.sheen > 0.f && diel * opaque > 0.f) w.d += 0.1f * m.sheen;
  const float m5 = schlick_m5(NoV);
  const vec3 F = F0 + (F90 - F0) * m5;
  const float spec = luminance(F);
  w.s = spec > 0.f ? fmaxf(spec, 0.05f) : 0.f;
  w.t = saturate(m.transmission) * diel;
  w.c = 0.f;
  if (m.clearcoat > 0.f) {
    const float mm = 1.f - saturate(NoV), m2 = mm * mm;
    w.c = fmaxf(m.clearcoat * 0.25f * (0.04f + 0.96f * m2 * m2 * mm), 0.02f);
  }
  return w;
}
PHX_DEV LobeW lobe_weights(const phx_material &m, float NoV) {
  vec3 F0, F90;
  spec_f0_f90(m, F0, F90);
  return lobe_weights(m, NoV, F0, F90);
}

PHX_DEV vec3 eval_diffuse(const phx_material &m, float NoL, float NoV) {
  const float w = (1.f - m.metallic) * (1.f - m.transmission) * diffuse_scale(m, NoV);
  if (w <= 0.f) return {0.f, 0.f, 0.f};
  const vec3 base = base_of(m);
  if (m.subsurface > 0.f) {
    const vec3 dif = base * (w * kInvPi * NoL);
    const float mm = 1.f - NoL;
    const float Fss = 1.f + (0.5f - 1.f) * mm * mm * mm * mm * mm;
    const float ss = 1.25f * (Fss * (1.f / (NoL + 0.001f) - 0.5f) + 0.5f);
    const vec3 ssc = lerp(base, ld3(m.subsurface_color), m.subsurface);
    return lerp(dif, ssc * (w * kInvPi * NoL * ss), m.subsurface);
  }
  return base * (w * kInvPi * NoL);
}
PHX_DEV vec3 eval_spec(const phx_material &m, vec3 wo, vec3 wi, vec3 n, float a) {
  const float NoV = saturate(dot(n, wo)), NoL = saturate(dot(n, wi));
  if (NoV <= 0.f || NoL <= 0.f) return {0.f, 0.f, 0.f};
  const vec3 h = normalize(wi + wo);
  const float NoH = saturate(dot(n, h)), VoH = saturate(dot(wo, h));
  vec3 F0, F90;
  spec_f0_f90(m, F0, F90);
  const float m5 = schlick_m5(VoH);
  const vec3 F = F0 + (F90 - F0) * m5;
  float D, G;
  if (m.anisotropic > 0.f) {
    const float aspect = sqrtf(1.f - 0.9f * m.anisotropic);
    const float ax = fmaxf(a / aspect, 0.001f), ay = fmaxf(a * aspect, 0.001f);
    vec3 t, b;
    onb(n, t, b);
    D = D_ggx_aniso(NoH, dot(t, h), dot(b, h), ax, ay);
    G = sqrtf(G2(NoV, NoL, ax) * G2(NoV, NoL, ay));
  } else {
    D = D_ggx(NoH, a);
    G = G2(NoV, NoL, a);
  }
  return F * (D * G / (4.f * fmaxf(NoV, 1e-7f)));
}
PHX_DEV vec3 eval_sheen(const phx_material &m, vec3 wo, vec3 wi, vec3 n) {
  if (m.sheen <= 0.f || m.metallic >= 1.f) return {0.f, 0.f, 0.f};
  const vec3 h = normalize(wi + wo);
  const float HoL = fabsf(dot(h, wi));
  const vec3 base = base_of(m);
  const float lum = luminance(base);
  const vec3 tint = lum > 0.f ? base * (1.f / lum) : vec3{1.f, 1.f, 1.f};
  const vec3 cs = lerp(vec3{1.f, 1.f, 1.f}, tint, m.sheen_tint);
  const float mm = 1.f - HoL, m2 = mm * mm, m5 = m2 * m2 * mm;
  const float NoL = saturate(dot(n, wi));
  const float w = (1.f - m.metallic) * (1.f - m.transmission);
  return cs * (m.sheen * m5 * w * NoL);
}
PHX_DEV vec3 eval_clearcoat(const phx_material &m, vec3 wo, vec3 wi, vec3 n) {
  if (m.clearcoat <= 0.f) return {0.f, 0.f, 0.f};
  const float NoV = saturate(dot(n, wo)), NoL = saturate(dot(n, wi));
  if (NoV <= 0.f || NoL <= 0.f) return {0.f, 0.f, 0.f};
  const float a = fmaxf(m.clearcoat_roughness, 0.001f);
  const vec3 h = normalize(wi + wo);
  const float NoH = saturate(dot(n, h)), VoH = saturate(dot(wo, h));
  const float F0 = dielectric_f0(1.5f);
  const float m5 = schlick_m5(VoH);
  const float F = F0 + (1.f - F0) * m5;
  const float D = D_ggx(NoH, a), G = G2(NoV, NoL, a);
  const float w = m.clearcoat * 0.25f;
  const float s = w * D * G / (4.f * fmaxf(NoV, 1e-7f));
  return vec3{F * s, F * s, F * s};
}

// f * cos(theta_i) of the reflection lobes (disney_bsdf.h:288-307).
PHX_DEV vec3 bsdf_eval(const phx_material &m, vec3 wo, vec3 wi, vec3 n, vec3 sn) {
  if (dot(wo, n) <= 0.f || dot(wi, n) <= 0.f) return {0.f, 0.f, 0.f};
  const float NoL = saturate(dot(sn, wi)), NoV = saturate(dot(sn, wo));
  if (NoL <= 0.f || NoV <= 0.f) return {0.f, 0.f, 0.f};
  const float a = fmaxf(m.roughness, 0.001f);
  vec3 f = eval_diffuse(m, NoL, NoV);
  if (!spec_delta(m)) f = f + eval_spec(m, wo, wi, sn, a);
  f = f + eval_sheen(m, wo, wi, sn);
  f = f + eval_clearcoat(m, wo, wi, sn);
  return f;
}
// Solid-angle density of bsdf_sample's reflection lobes (disney_bsdf.h:310-355).
PHX_DEV float bsdf_pdf(const phx_material &m, vec3 wo, vec3 wi, vec3 n, vec3 sn) {
  if (dot(wo, n) <= 0.f || dot(wi, n) <= 0.f) return 0.f;
  const float NoL = saturate(dot(sn, wi)), NoV = saturate(dot(sn, wo));
  if (NoL <= 0.f || NoV <= 0.f) return 0.f;
  const float a = fmaxf(m.roughness, 0.001f);
  const LobeW w = lobe_weights(m, NoV);
  const float sum = w.d + w.s + w.t + w.c;
  if (sum <= 0.f) return 0.f;
  const float pd = w.d / sum, ps = w.s / sum, pc = w.c / sum;
  float pdf = 0.f;
  if (pd > 0.f) pdf += pd * (NoL * kInvPi);
  if (ps > 0.f && !spec_delta(m)) {
    const vec3 h = normalize(wi + wo);
    pdf += ps * ggx_pdf(saturate(dot(sn, h)), saturate(dot(wo, h)), a);
  }
  if (pc > 0.f) {
    const vec3 h = normalize(wi + wo);
    pdf += pc * ggx_pdf(saturate(dot(sn, h)), saturate(dot(wo, h)), fmaxf(m.clearcoat_roughness, 0.001f));
  }
  return fmaxf(pdf, 0.f);
}
// bsdf_eval and bsdf_pdf of one direction in one pass: the half vector, the
// GGX distributions and the Fresnel endpoints are computed once.  Every value
// is rounded exactly as in the two separate functions above.
PHX_DEV vec3 bsdf_eval_pdf(const phx_material &m, vec3 wo, vec3 wi, vec3 n, vec3 sn, float &pdf) {
  pdf = 0.f;
  if (dot(wo, n) <= 0.f || dot(wi, n) <= 0.f) return {0.f, 0.f, 0.f};
  const float NoL = saturate(dot(sn, wi)), NoV = saturate(dot(sn, wo));
  if (NoL <= 0.f || NoV <= 0.f) return {0.f, 0.f, 0.f};
  const float a = fmaxf(m.roughness, 0.001f);
  const bool delta = spec_delta(m);
  const vec3 h = normalize(wi + wo);
  const float NoH = saturate(dot(sn, h)), VoH = saturate(dot(wo, h));
  const float m5 = schlick_m5(VoH);
  vec3 F0, F90;
  spec_f0_f90(m, F0, F90);
  vec3 f = eval_diffuse(m, NoL, NoV);
  float D_spec = -1.f;  // the isotropic GGX D of the specular lobe, once known
  if (!delta) {
    const vec3 F = F0 + (F90 - F0) * m5;
    float D, G;
    if (m.anisotropic > 0.f) {
      const float aspect = sqrtf(1.f - 0.9f * m.anisotropic);
      const float ax = fmaxf(a / aspect, 0.001f), ay = fmaxf(a * aspect, 0.001f);
      vec3 t, b;
      onb(sn, t, b);
      D = D_ggx_aniso(NoH, dot(t, h), dot(b, h), ax, ay);
      G = sqrtf(G2(NoV, NoL, ax) * G2(NoV, NoL, ay));
    } else {
      D = D_ggx(NoH, a);
      G = G2(NoV, NoL, a);
      D_spec = D;
    }
    f = f + F * (D * G / (4.f * fmaxf(NoV, 1e-7f)));
  }
  if (m.sheen > 0.f && m.metallic < 1.f) {
    const float HoL = fabsf(dot(h, wi));
    const vec3 base = base_of(m);
    const float lum = luminance(base);
    const vec3 tint = lum > 0.f ? base * (1.f / lum) : vec3{1.f, 1.f, 1.f};
    const vec3 cs = lerp(vec3{1.f, 1.f, 1.f}, tint, m.sheen_tint);
    const float mm = 1.f - HoL, m2 = mm * mm, sm5 = m2 * m2 * mm;
    const float w = (1.f - m.metallic) * (1.f - m.transmission);
    f = f + cs * (m.sheen * sm5 * w * NoL);
  }
  float D_cc = 0.f;
  const float a_cc = fmaxf(m.clearcoat_roughness, 0.001f);
  if (m.clearcoat > 0.f) {
    const float F0c = dielectric_f0(1.5f);
    const float F = F0c + (1.f - F0c) * m5;
    D_cc = D_ggx(NoH, a_cc);
    const float G = G2(NoV, NoL, a_cc);
    const float w = m.clearcoat * 0.25f;
    const float sc = w * D_cc * G / (4.f * fmaxf(NoV, 1e-7f));
    f = f + vec3{F * sc, F * sc, F * sc};
  }
  const LobeW w = lobe_weights(m, NoV, F0, F90);
  const float sum = w.d + w.s + w.t + w.c;
  if (sum <= 0.f) return f;
  const float pd = w.d / sum, ps = w.s / sum, pc = w.c / sum;
  float p = 0.f;
  if (pd > 0.f) p += pd * (NoL * kInvPi);
  if (ps > 0.f && !delta) {
    const float D = D_spec >= 0.f ? D_spec : D_ggx(NoH, a);
    p += ps * (D * NoH / (4.f * fmaxf(VoH, 1e-7f)));
  }
  if (pc > 0.f) p += pc * (D_cc * NoH / (4.f * fmaxf(VoH, 1e-7f)));
  pdf = fmaxf(p, 0.f);
  return f;
}
PHX_DEV vec3 ggx_half(vec3 n, float a, Rng &rng) {
  vec3 t, b;
  onb(n, t, b);
  const float r1 = rng.next(), r2 = rng.next();
  const float a2 = a * a;
  const float ct = sqrtf((1.f - r1) / (1.f + (a2 - 1.f) * r1));
  const float st = sqrtf(fmaxf(0.f, 1.f - ct * ct));
  const float phi = 2.f * kPi * r2;
  return normalize(t * (st * __cosf(phi)) + b * (st * __sinf(phi)) + n * ct);
}
// disney_bsdf.h:381-521.  Delta samples carry pdf = 1 and the path weight in f.
PHX_DEV BsdfSample bsdf_sample(const phx_material &m, vec3 wo, vec3 n, vec3 sn, Rng &rng) {
  BsdfSample out{{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, 0.f, false, false};
  const bool front = dot(wo, n) > 0.f;
  if (!front && m.transmission <= 0.f) return out;
  const float rough = fmaxf(m.roughness, 0.001f);
  const float NoV = saturate(fabsf(dot(sn, wo)));
  LobeW w;
  float te = 1.f;
  if (!front) {
    w = LobeW{0.f, 0.f, 1.f, 0.f};
  } else {
    w = lobe_weights(m, NoV);
    te = w.t;
  }
  const float sum = w.d + w.s + w.t + w.c;
  if (sum <= 0.f) return out;
  const float pd = w.d / sum, ps = w.s / sum, pt = w.t / sum, pc = w.c / sum;
  const float xi = rng.next();
  // The reflection lobes only choose the direction; f and pdf are evaluated
  // once below, so a wave whose lanes chose different lobes evaluates once.
  bool evaluate = false;
  const BsdfSample none{{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, 0.f, false, false};
  if (xi < pd) {
    out.diffuse = true;
    out.wi = cosine_hemisphere(sn, rng);
    evaluate = true;
  } else if (xi < pd + ps) {
    if (spec_delta(m)) {
      out.wi = reflect(-wo, sn);
      if (dot(out.wi, n) <= 0.f) out.wi = reflect(-wo, n);
      if (dot(out.wi, n) <= 0.f) return none;
      vec3 F0, F90;
      spec_f0_f90(m, F0, F90);
      const float m5 = schlick_m5(saturate(dot(wo, sn)));
      out.f = (F0 + (F90 - F0) * m5) * (1.f / ps);
      out.pdf = 1.f;
      out.specular = true;
      return out;
    }
    const vec3 h = ggx_half(sn, rough, rng);
    out.wi = reflect(-wo, h);
    if (dot(out.wi, n) <= 0.f) return none;
    evaluate = true;
  } else if (xi < pd + ps + pt && m.transmission > 0.f) {
    const vec3 nn = front ? n : n * -1.f;
    vec3 snf = front ? sn : sn * -1.f;
    if (dot(snf, wo) <= 0.f) snf = nn;
    vec3 h = snf;
    if (m.roughness >= 0.001f) {
      h = ggx_half(snf, rough, rng);
      if (dot(h, wo) <= 0.f) h = snf;
    }
    const bool refractive = (m.flags & PHX_MATERIAL_FLAG_REFRACTIVE) != 0 && m.thin == 0;
    const float ior = fmaxf(m.ior, 1.0001f);
    const float eta = refractive ? (front ? (1.f / ior) : ior) : 1.f;
    const float cos_i = saturate(dot(wo, h));
    float F;
    if (refractive) {
      F = fresnel_dielectric(cos_i, eta);
    } else {
      const float f0 = dielectric_f0(m.ior);
      const float m1 = 1.f - cos_i, m2 = m1 * m1;  // (1 - cos_i)^5
      F = f0 + (1.f - f0) * (m2 * m2 * m1);
    }
    const float lw = te / pt;
    out.pdf = 1.f;
    out.specular = true;
    if (rng.next() < F) {
      out.wi = reflect(-wo, h);
      if (dot(out.wi, nn) <= 0.f) out.wi = reflect(-wo, nn);
      if (dot(out.wi, nn) <= 0.f) return none;
      out.f = vec3{lw, lw, lw};
      return out;
    }
    if (refractive) {
      out.wi = normalize(refract(-wo, h, eta));
      if (dot(out.wi, nn) >= 0.f) out.wi = normalize(refract(-wo, snf, eta));
      if (dot(out.wi, nn) >= 0.f) return none;
    } else {
      out.wi = normalize(-wo);
    }
    out.f = base_of(m) * lw;
    return out;
  } else if (pc > 0.f) {
    const vec3 h = ggx_half(sn, fmaxf(m.clearcoat_roughness, 0.001f), rng);
    out.wi = reflect(-wo, h);
    if (dot(out.wi, n) <= 0.f) return none;
    evaluate = true;
  }
  if (evaluate) out.f = bsdf_eval_pdf(m, wo, out.wi, n, sn, out.pdf);
  return out;
}

}  // namespace phx
