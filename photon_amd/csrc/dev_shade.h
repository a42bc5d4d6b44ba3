// One shading event of a path: pt_shade (pathtracer.cpp:765-1090) as a device
// function shared by the streaming integrator and the reference-structure
// baseline.  Inputs are the path's ray and state and its closest hit; outputs
// are the continuing ray (if any), the shadow ray (if any) and the emission /
// background added to the sample's radiance.
#pragma once

#include "dev_scene.h"

namespace phx {

constexpr uint32_t kEmissionOnly = 0x80000000u;
// The queue entry holds no path: its pixel stopped sampling (adaptive
// sampling), so nothing is traced, shaded or added for it.  Bit 7 of the
// bounce-kind word, which the 7-bit counters leave free.
constexpr uint32_t kNoPath = 0x80u;

struct ShadeParams {
  phx_camera cam;
  uint32_t width, height, npix;
  uint32_t max_depth;
  uint32_t max_it;
  uint32_t spp;
  uint32_t sample_offset;
  int32_t bottom_up;
  float region[4];
  vec3 bg;
  float bg_alpha;
  vec3 ambient;
  float throughput_clamp, clamp_direct, clamp_indirect;
  int32_t max_diffuse, max_glossy, max_transmission, max_volume;
  const float4 *bg_image;  // RGBA rows from the bottom of the frame, or null
  uint32_t bg_w, bg_h;
};

struct PathState {
  vec3 ro, rd;
  float tmin, tmax;
  vec3 tp;
  float prev_pdf;
  uint32_t depth;
  uint32_t path_it;  // bounce-loop iteration of this path (the RNG dimension)
  uint32_t kinds;
  bool primary;
  float path_dist;
  uint32_t item;  // the path's item (view, spp range) in the run's item table
  uint32_t pixel; // the path's pixel in its view (set by the caller; not queued)
};

struct ShadeOut {
  bool cont;
  PathState next;
  bool shadow;
  vec3 so, sd;
  float stmax;
  vec3 sc;
  vec3 add;      // radiance added to the sample
  float add_alpha;
  bool primary_event;  // the camera ray ended here (AOV source)
};

PHX_DEV bool count_bounce(uint32_t &kinds, uint32_t shift, int limit) {
  uint32_t c = (kinds >> shift) & 0x7fu;
  if (c < 0x7fu) ++c;
  kinds = (kinds & ~(0x7fu << shift)) | (c << shift);
  return limit >= 0 && int(c) > limit;
}
PHX_DEV vec3 clamp_contribution(vec3 c, int light_bounce, float ds, float is) {
  const float limit = light_bounce > 0 ? is : ds;
  if (!(limit > 0.f)) return c;
  const float sum = fabsf(c.x) + fabsf(c.y) + fabsf(c.z);
  return sum > limit ? c * (limit / sum) : c;
}

// Surface attributes at a hit (mesh_intersector.h:126-178).
PHX_DEV SurfaceHit surface_at(const DevScene &s, uint32_t tri, float t, float b1, float b2, vec3 ro, vec3 rd) {
  SurfaceHit h;
  const uint4 ta = s.tri_attr[tri];
  if (s.tri_ng) {
    const float4 nr = s.tri_ng[tri];
    h.prim = __float_as_uint(nr.w);
    h.ng = vec3{nr.x, nr.y, nr.z};
  } else {
    const TriRec tr = s.tris[tri];
    h.prim = __float_as_uint(tr.v0.w);
    h.ng = tri_normal(tr);
  }
  h.p = ro + rd * t;
  const uint32_t i0 = ta.x, i1 = ta.y, i2 = ta.z;
  const float w0 = 1.f - b1 - b2;
  h.sn = h.ng;
  if (s.normals) {
    h.sn = normalize(ld3(s.normals + size_t(i0) * 3) * w0 + ld3(s.normals + size_t(i1) * 3) * b1 +
                     ld3(s.normals + size_t(i2) * 3) * b2);
    if (dot(h.sn, h.ng) < 0.f) h.sn = h.sn * -1.f;
  }
  h.uv = vec2{0.f, 0.f};
  if (s.texcoords) {
    const float *a = s.texcoords + size_t(i0) * 2, *b = s.texcoords + size_t(i1) * 2, *c = s.texcoords + size_t(i2) * 2;
    h.uv = {a[0] * w0 + b[0] * b1 + c[0] * b2, a[1] * w0 + b[1] * b1 + c[1] * b2};
  }
  h.has_color = false;
  h.color = vec3{0.f, 0.f, 0.f};
  if (s.colors && s.colors[size_t(i0) * 3] >= 0.f) {
    h.color = ld3(s.colors + size_t(i0) * 3) * w0 + ld3(s.colors + size_t(i1) * 3) * b1 +
              ld3(s.colors + size_t(i2) * 3) * b2;
    h.has_color = true;
  }
  h.mid = ta.w;  // the material id (0 without material ids), as material_ids[prim]
  h.b1 = b1;
  h.b2 = b2;
  // pathtracer.cpp:906-910
  if (!(dot(h.sn, h.sn) > 0.25f)) h.sn = h.ng;
  else if (dot(h.sn, h.ng) < 0.f) h.sn = h.sn * -1.f;
  return h;
}

// Bilinear lookup in the background image at image-plane position (bs, bt),
// edge texels clamped (sample_background_image, pathtracer.cpp:370-389).
PHX_DEV float4 background_at(const ShadeParams &w, float bs, float bt) {
  const int W = int(w.bg_w), H = int(w.bg_h);
  const float fx = bs * float(W) - 0.5f, fy = bt * float(H) - 0.5f;
  const float flx = floorf(fx), fly = floorf(fy);
  const float tx = fx - flx, ty = fy - fly;
  auto cl = [](int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); };
  const int x0 = cl(int(flx), W), x1 = cl(int(flx) + 1, W), y0 = cl(int(fly), H), y1 = cl(int(fly) + 1, H);
  const float4 p00 = w.bg_image[y0 * W + x0], p01 = w.bg_image[y0 * W + x1];
  const float4 p10 = w.bg_image[y1 * W + x0], p11 = w.bg_image[y1 * W + x1];
  const float ux = 1.f - tx, uy = 1.f - ty;
  const float4 a = make_float4(p00.x * ux + p01.x * tx, p00.y * ux + p01.y * tx, p00.z * ux + p01.z * tx,
                               p00.w * ux + p01.w * tx);
  const float4 b = make_float4(p10.x * ux + p11.x * tx, p10.y * ux + p11.y * tx, p10.z * ux + p11.z * tx,
                               p10.w * ux + p11.w * tx);
  return make_float4(a.x * uy + b.x * ty, a.y * uy + b.y * ty, a.z * uy + b.z * ty, a.w * uy + b.w * ty);
}

// Observers of the weights that depend on the material parameters, called by
// shade_path at next-event estimation (f * cos toward the light and the
// constant factor Le * w_mis / pdf_light) and at the BSDF sample that
// continues the path.  The primal kernels pass NoHooks; the path-replay
// adjoint records derivatives there.
struct NoHooks {
  PHX_DEV void nee(const phx_material &, const SurfaceHit &, vec3, vec3, vec3, vec3, vec3, vec3) {}
  PHX_DEV void bsdf(const phx_material &, const SurfaceHit &, vec3, vec3, vec3, vec3, const BsdfSample &) {}
};

// Shading tiers: the code a run's scenes and views need, so a kernel carries
// no branch that could not run in it (a smaller kernel keeps more of the rest
// in registers).  kTierLean: no volumes, textures, parameter sources or
// analytic lights, no background image; kTierStandard: volumes, textures and
// analytic lights; kTierFull: also parameter sources (attributes, samplers,
// sourced normal maps and emission) and the background image.
constexpr int kTierLean = 0, kTierStandard = 1, kTierFull = 2;

// VOL_PRE: the volume event was sampled ahead (k_volume) and is given by
// vol_ev (scatter distance or -1, albedo) and vol_rng (the RNG state after its
// draws); otherwise it is sampled here.  The draws are the same either way.
template <class Hooks, bool VOL_PRE = false, int TIER = kTierFull>
PHX_DEV void shade_path(const DevScene &s, const ShadeParams &w, const PathState &ps, bool hr_hit, uint32_t htri,
                        float ht, float b1, float b2, Rng &rng, ShadeOut &out, Hooks &hooks,
                        float4 vol_ev = make_float4(-1.f, 0.f, 0.f, 0.f), uint32_t vol_rng = 0u) {
  out.cont = false;
  out.shadow = false;
  out.add = vec3{0.f, 0.f, 0.f};
  out.add_alpha = 0.f;
  out.primary_event = false;
  const vec3 ro = ps.ro, rd = ps.rd;
  const uint32_t depth = ps.depth;
  bool primary = ps.primary;
  uint32_t kinds = ps.kinds;
  float path_dist = ps.path_dist;
  const vec3 tp = ps.tp;
  const float prev_pdf = ps.prev_pdf;
  const bool emission_only = (kinds & kEmissionOnly) != 0u;
  // The continuing path inherits everything; fields are overwritten below.
  out.next = ps;
  out.next.path_it = ps.path_it + 1u;

  constexpr bool STD = TIER >= kTierStandard, FULL = TIER >= kTierFull;
  if (STD && s.vol_count > 0) {
    // Scattering inside a volume in front of the hit (:789-857).
    vec3 alb{0.f, 0.f, 0.f};
    float ts;
    if (VOL_PRE) {
      ts = vol_ev.x;
      alb = vec3{vol_ev.y, vol_ev.z, vol_ev.w};
      rng.s = vol_rng;
    } else {
      ts = vol_sample_scatter(s, ro, rd, ps.tmin, hr_hit ? ht : ps.tmax, rng, alb);
    }
    if (ts >= 0.f) {
      const vec3 p = ro + rd * ts;
      if (primary) {
        primary = false;
        out.add_alpha += 1.f;
        out.primary_event = true;
      }
      if (depth >= w.max_depth || emission_only) return;
      const vec3 tpv = tp * alb;
      const float phase = 0.25f * kInvPi;
      const Direct ds = sample_direct<STD, FULL>(s, p, rng);
      if (ds.valid) {
        const float wm = ds.mis ? power_heuristic(ds.pdf, phase) : 1.f;
        out.sc = clamp_contribution(tpv * ds.L * (phase * wm / ds.pdf), int(depth), w.clamp_direct, w.clamp_indirect);
        out.shadow = true;
        out.so = p;
        out.sd = ds.wi;
        out.stmax = ds.dist < 1e29f ? ds.dist * (1.f - 1e-4f) : 1e30f;
      }
      vec3 tn = tpv;
      if (depth >= 2) {
        const float surv = fminf(fmaxf(max3(tpv), 0.05f), 0.95f);
        if (rng.next() >= surv) return;
        tn = tpv * (1.f / surv);
      }
      if (count_bounce(kinds, 24u, w.max_volume)) kinds |= kEmissionOnly;
      out.cont = true;
      out.next.ro = p;
      out.next.rd = uniform_sphere(rng);
      out.next.tmin = 0.f;
      out.next.tmax = 1e30f;
      out.next.tp = tn;
      out.next.prev_pdf = phase;
      out.next.depth = depth + 1u;
      out.next.kinds = kinds;
      out.next.primary = false;
      return;
    }
  }

  if (!hr_hit) {
    // Miss: background for camera rays, environment with MIS (:860-898).
    vec3 Lm{0.f, 0.f, 0.f};
    if (primary) {
      if (s.has_env && s.env_visible) {
        Lm = env_eval(s, rd);
        out.add_alpha += 1.f;
      } else if (FULL && w.bg_image) {
        const uint32_t px = ps.pixel % w.width, py = ps.pixel / w.width;
        const float bs = (float(px) + 0.5f) / float(w.width), bt = (float(py) + 0.5f) / float(w.height);
        const float4 bg = background_at(w, bs, w.bottom_up ? bt : 1.f - bt);
        Lm = vec3{bg.x, bg.y, bg.z} * bg.w;
        out.add_alpha += bg.w;
      } else {
        Lm = w.bg * w.bg_alpha;
        out.add_alpha += w.bg_alpha;
      }
    } else {
      Lm = w.ambient;
      if (s.has_env) {
        float wm = 1.f;
        if (prev_pdf > 0.f) wm = power_heuristic(prev_pdf, env_pdf(s, rd) / float(s.strategy_count));
        Lm = Lm + env_eval(s, rd) * wm;
      }
    }
    vec3 fm = tp * Lm;
    if (!primary || (s.has_env && s.env_visible))
      fm = clamp_contribution(fm, int(depth) - 1, w.clamp_direct, w.clamp_indirect);
    out.add = fm;
    return;
  }

  const SurfaceHit sh = surface_at(s, htri, ht, b1, b2, ro, rd);
  const vec3 p = sh.p, ng = sh.ng, wo = -rd;
  float alpha = 1.f;
  const phx_material mat = resolve_material<STD, FULL>(s, sh, alpha);
  const bool hidden = (mat.flags & PHX_MATERIAL_FLAG_HIDDEN_TO_CAMERA) != 0 && primary;
  if (hidden || (alpha < 1.f && rng.next() >= alpha)) {
    // Stochastic transparency (:929-951): continue behind the surface.
    float ntmin = s.ray_eps, ntmax = 1e30f;
    path_dist += ht;
    if (primary) {
      float cn, cf;
      clip_range(w.cam, rd, cn, cf);
      if (cf < 1e30f) {
        ntmax = cf - path_dist;
        if (ntmax < s.ray_eps) {
          ntmin = 1.f;
          ntmax = 0.f;
        }
      }
    }
    out.cont = true;
    out.next.ro = p;
    out.next.tmin = ntmin;
    out.next.tmax = ntmax;
    out.next.prev_pdf = -1.f;
    out.next.path_dist = path_dist;
    return;
  }
  vec3 n = ng, sn = sh.sn;
  const bool front = dot(n, wo) >= 0.f;
  if (!(mat.transmission > 0.f) && !front) {
    n = n * -1.f;
    sn = sn * -1.f;
  }
  if (STD && mat.normal_tex >= 0 && s.tex_count > 0) {
    vec3 nm = tex_sample(s, mat.normal_tex, sh.uv.x, sh.uv.y);
    nm = nm * 2.f - vec3{1.f, 1.f, 1.f};
    vec3 ta, bi;
    onb(sn, ta, bi);
    sn = normalize(ta * nm.x + bi * nm.y + sn * nm.z);
  }
  if (FULL && mat.normal_src.kind != PHX_SOURCE_CONSTANT) sn = apply_normal_map(s, mat.normal_src, sh, sn);
  if (primary) {
    primary = false;
    out.add_alpha += 1.f;
    out.primary_event = true;
  }
  if (material_may_emit(mat) && emits_on_side(mat, front)) {
    float wm = 1.f;
    if (prev_pdf > 0.f && s.em_count > 0) {
      const float pa = s.prim_light_pdf[sh.prim];
      const float cl = fabsf(dot(ng, wo));
      if (pa > 0.f && cl > 0.f) wm = power_heuristic(prev_pdf, pa * ht * ht / cl / float(s.strategy_count));
    }
    out.add = clamp_contribution(tp * material_emission(mat) * wm, int(depth) - 1, w.clamp_direct, w.clamp_indirect);
  }
  if ((mat.flags & PHX_MATERIAL_FLAG_EMITTER_ONLY) != 0 || depth >= w.max_depth || emission_only) return;
  {
    // Next event estimation (:1006-1035).
    const Direct ds = sample_direct<STD, FULL>(s, p, rng);
    if (ds.valid) {
      float bp;
      const vec3 f = bsdf_eval_pdf(mat, wo, ds.wi, n, sn, bp);
      if (f.x > 0.f || f.y > 0.f || f.z > 0.f) {
        float wm = 1.f;
        if (ds.mis) wm = power_heuristic(ds.pdf, bp);
        out.sc = clamp_contribution(tp * f * ds.L * (wm / ds.pdf), int(depth), w.clamp_direct, w.clamp_indirect);
        hooks.nee(mat, sh, wo, ds.wi, n, sn, tp, ds.L * (wm / ds.pdf));
        const vec3 sorg = p + n * s.ray_eps;
        vec3 sdir = ds.wi;
        float stmax = 1e30f;
        if (ds.dist < 1e29f) {
          const vec3 tl = p + ds.wi * ds.dist - sorg;
          const float sdist = length(tl);
          sdir = sdist > 0.f ? tl * (1.f / sdist) : ds.wi;
          stmax = sdist * (1.f - 1e-3f);
        }
        out.shadow = true;
        out.so = sorg;
        out.sd = sdir;
        out.stmax = stmax;
      }
    }
  }
  // BSDF sample, Russian roulette, next ray (:1037-1089).
  const BsdfSample bs = bsdf_sample(mat, wo, n, sn, rng);
  if (bs.pdf <= 0.f || (bs.f.x <= 0.f && bs.f.y <= 0.f && bs.f.z <= 0.f)) return;
  vec3 tpn = tp * (bs.f * (1.f / bs.pdf));
  if (depth >= 2) {
    const float surv = fminf(fmaxf(max3(tpn), 0.05f), 0.95f);
    if (rng.next() >= surv) return;
    tpn = tpn * (1.f / surv);
  }
  if (w.throughput_clamp > 0.f) {
    const float mm = max3(tpn);
    if (mm > w.throughput_clamp) tpn = tpn * (w.throughput_clamp / mm);
  }
  if (!finite3(tpn)) return;
  const vec3 wi = normalize(bs.wi);
  const float side = dot(wi, ng) < 0.f ? -1.f : 1.f;
  const bool through = (dot(wi, ng) < 0.f) != (dot(wo, ng) < 0.f);
  const bool exceeded = through ? count_bounce(kinds, 16u, w.max_transmission)
                                : (bs.diffuse ? count_bounce(kinds, 0u, w.max_diffuse)
                                              : count_bounce(kinds, 8u, w.max_glossy));
  if (exceeded) kinds |= kEmissionOnly;
  hooks.bsdf(mat, sh, wo, n, sn, tp, bs);
  out.cont = true;
  out.next.ro = p + ng * (side * s.ray_eps);
  out.next.rd = wi;
  out.next.tmin = 0.f;
  out.next.tmax = 1e30f;
  out.next.tp = tpn;
  out.next.prev_pdf = bs.specular ? -1.f : bs.pdf;
  out.next.depth = depth + 1u;
  out.next.kinds = kinds;
  out.next.primary = false;
}
PHX_DEV void shade_path(const DevScene &s, const ShadeParams &w, const PathState &ps, bool hr_hit, uint32_t htri,
                        float ht, float b1, float b2, Rng &rng, ShadeOut &out) {
  NoHooks h;
  shade_path(s, w, ps, hr_hit, htri, ht, b1, b2, rng, out, h);
}

}  // namespace phx
