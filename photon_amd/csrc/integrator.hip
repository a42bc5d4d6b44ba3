// Wavefront path-tracing kernels for gfx950.
//
// Replaces PathTracer::render's per-bounce kernels (pathtracer.cpp:688-1155):
//   k_generate   <- pt_init_sample (:691-760)
//   k_extend     <- trace_closest (kokkos_backend.cpp:21-61), over the
//                   compacted ray queue instead of all W*H slots
//   k_shade      <- pt_shade (:765-1090), pushing continuing paths and shadow
//                   rays into compacted queues
//   k_shadow     <- trace_occluded + pt_nee_accum_pack (:1092-1123), with no
//                   host readback: empty queues end the loop on the device
//   k_accumulate / k_resolve <- pt_resolve (:1143-1155)
//
// A pass renders k samples of every pixel at once: path slot = s * npix + pixel.
// Each slot owns its radiance accumulator, so no atomics touch image data and
// the result does not depend on scheduling.  RNG streams are keyed by
// (pixel, sample, dimension) exactly as the reference keys them, so a path
// here draws the same numbers as the same path in the reference.
#include "integrator.h"

#include "dev_queue.h"
#include "dev_trace.h"

namespace phx {

constexpr uint32_t kEmissionOnly = 0x80000000u;
constexpr uint32_t kPrimaryBit = 0x10000u;

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
PHX_DEV vec3 xyz(float4 v) { return {v.x, v.y, v.z}; }
PHX_DEV float4 f4(vec3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

__global__ void k_reset(WaveQueues q, uint32_t nslots) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kCounterWords) q.counters[i] = (i == kRayCount) ? nslots : 0u;
}

__global__ __launch_bounds__(256) void k_generate(WaveParams w, WaveQueues q) {
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= w.nslots) return;
  const uint32_t pixel = slot % w.npix, sl = slot / w.npix;
  const uint32_t sample = w.sample_base + sl;
  const uint32_t x = pixel % w.width, y = pixel / w.width;
  Rng rng = make_rng(pixel, sample, 0u);
  float u = (float(x) + rng.next()) / float(w.width);
  float v = (float(y) + rng.next()) / float(w.height);
  if (!w.bottom_up) v = 1.f - v;
  u = w.region[0] + (w.region[2] - w.region[0]) * u;
  v = w.region[1] + (w.region[3] - w.region[1]) * v;
  vec3 o, d;
  camera_ray(w.cam, u, v, rng, o, d);
  float tn, tf;
  clip_range(w.cam, d, tn, tf);
  q.o[0][slot] = f4(o, tn);
  q.d[0][slot] = f4(d, tf);
  q.tp[0][slot] = make_float4(1.f, 1.f, 1.f, -1.f);
  q.st[0][slot] = make_uint4(slot, kPrimaryBit, 0u, 0u);
  q.slot_L[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
  q.slot_aov1[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Closest hit over the ray queue of iteration `it`: persistent waves fetch
// 64 rays at a time.
__global__ __launch_bounds__(256) void k_extend(DevScene s, WaveQueues q, uint32_t it) {
  __shared__ int lds[kShortStack * 256];
  const uint32_t count = q.counters[kRayCount + it];
  const int buf = it & 1;
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  LaneStack st{&lds[threadIdx.x], 256, q.spill + size_t(gtid) * (kMaxStack - kShortStack), 0};
  const uint32_t lane = lane_id();
  while (true) {
    const uint32_t base = wave_fetch(&q.counters[kFetchExtend + it]);
    if (base >= count) break;
    const uint32_t i = base + lane;
    const bool valid = i < count;
    bool live = false;
    if (valid) {
      const float4 o = q.o[buf][i], d = q.d[buf][i];
      live = d.w > o.w;
      TraceHit h;
      traverse<false>(s, xyz(o), xyz(d), o.w, d.w, st, h);
      q.hit[i] = make_float4(h.t, h.u, h.v, __uint_as_float(h.tri));
    }
    wave_count(&q.stats[kStatClosest], live);
  }
}

// pt_shade (pathtracer.cpp:765-1090) on the queue of iteration `it`.
__global__ __launch_bounds__(256) void k_shade(DevScene s, WaveParams w, WaveQueues q, uint32_t it) {
  const uint32_t count = q.counters[kRayCount + it];
  if (blockIdx.x * blockDim.x >= count) return;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < count;
  const int buf = it & 1, nbuf = buf ^ 1;

  bool cont = false, shadow = false;
  float4 no = {}, nd = {}, ntp = {};
  uint4 nst = {};
  float4 so = {}, sd = {}, sc = {};

  if (valid) {
    const float4 o4 = q.o[buf][i], d4 = q.d[buf][i], tp4 = q.tp[buf][i], hit = q.hit[i];
    const uint4 st = q.st[buf][i];
    const vec3 ro = xyz(o4), rd = xyz(d4);
    const float rtmin = o4.w, rtmax = d4.w;
    const uint32_t slot = st.x;
    const uint32_t depth = st.y & 0xffffu;
    bool primary = (st.y & kPrimaryBit) != 0u;
    uint32_t kinds = st.z;
    float path_dist = __uint_as_float(st.w);
    vec3 tp = xyz(tp4);
    const float prev_pdf = tp4.w;
    const uint32_t pixel = slot % w.npix;
    const uint32_t sample = w.sample_base + slot / w.npix;
    const bool emission_only = (kinds & kEmissionOnly) != 0u;
    const uint32_t htri = __float_as_uint(hit.w);
    const bool hr_hit = htri != 0xffffffffu;
    const float ht = hit.x;
    Rng rng = make_rng(pixel, sample, 1u + it);
    float4 L = q.slot_L[slot];
    bool Ldirty = false;
    bool done = false;

    if (s.vol_count > 0) {
      vec3 alb{0.f, 0.f, 0.f};
      const float t_end = hr_hit ? ht : rtmax;
      const float ts = vol_sample_scatter(s, ro, rd, rtmin, t_end, rng, alb);
      if (ts >= 0.f) {
        done = true;
        const vec3 p = ro + rd * ts;
        if (primary) {
          primary = false;
          L.w += 1.f;
          Ldirty = true;
          q.slot_aov0[slot] = f4(-rd, path_dist + ts);
          q.slot_aov1[slot] = f4(alb, 1.f);
        }
        if (!(depth >= w.max_depth || emission_only)) {
          const vec3 tpv = tp * alb;
          const float phase = 0.25f * kInvPi;
          const Direct ds = sample_direct(s, p, rng);
          if (ds.valid) {
            const float wm = ds.mis ? power_heuristic(ds.pdf, phase) : 1.f;
            const vec3 c = clamp_contribution(tpv * ds.L * (phase * wm / ds.pdf), int(depth), w.clamp_direct,
                                              w.clamp_indirect);
            shadow = true;
            so = f4(p, 0.f);
            sd = f4(ds.wi, ds.dist < 1e29f ? ds.dist * (1.f - 1e-4f) : 1e30f);
            sc = f4(c, __uint_as_float(slot));
          }
          vec3 tn = tpv;
          bool killed = false;
          if (depth >= 2) {
            const float surv = fminf(fmaxf(max3(tpv), 0.05f), 0.95f);
            if (rng.next() >= surv) killed = true;
            else tn = tpv * (1.f / surv);
          }
          if (!killed) {
            if (count_bounce(kinds, 24u, w.max_volume)) kinds |= kEmissionOnly;
            cont = true;
            no = f4(p, 0.f);
            nd = f4(uniform_sphere(rng), 1e30f);
            ntp = f4(tn, phase);
            nst = make_uint4(slot, depth + 1u, kinds, __float_as_uint(path_dist));
          }
        }
      }
    }

    if (!done && !hr_hit) {
      done = true;
      vec3 Lm{0.f, 0.f, 0.f};
      if (primary) {
        if (s.has_env && s.env_visible) {
          Lm = env_eval(s, rd);
          L.w += 1.f;
        } else {
          Lm = w.bg * w.bg_alpha;
          L.w += w.bg_alpha;
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
      L = f4(xyz(L) + fm, L.w);
      Ldirty = true;
    }

    if (!done) {
      // Surface hit: attributes as intersect_mesh_bvh computes them
      // (mesh_intersector.h:126-178).
      const TriRec tr = s.tris[htri];
      const uint32_t prim = __float_as_uint(tr.v0.w);
      const float b1 = hit.y, b2 = hit.z;
      const vec3 p = ro + rd * ht;
      const vec3 ng = normalize(cross(xyz(tr.e1), xyz(tr.e2)));
      const uint32_t i0 = s.indices[prim * 3], i1 = s.indices[prim * 3 + 1], i2 = s.indices[prim * 3 + 2];
      const float w0 = 1.f - b1 - b2;
      vec3 hsn = ng;
      if (s.normals) {
        hsn = normalize(ld3(s.normals + size_t(i0) * 3) * w0 + ld3(s.normals + size_t(i1) * 3) * b1 +
                        ld3(s.normals + size_t(i2) * 3) * b2);
        if (dot(hsn, ng) < 0.f) hsn = hsn * -1.f;
      }
      vec2 uv{0.f, 0.f};
      if (s.texcoords) {
        const float *a = s.texcoords + size_t(i0) * 2, *b = s.texcoords + size_t(i1) * 2,
                    *c = s.texcoords + size_t(i2) * 2;
        uv = {a[0] * w0 + b[0] * b1 + c[0] * b2, a[1] * w0 + b[1] * b1 + c[1] * b2};
      }
      bool has_color = false;
      vec3 color{0.f, 0.f, 0.f};
      if (s.colors && s.colors[size_t(i0) * 3] >= 0.f) {
        color = ld3(s.colors + size_t(i0) * 3) * w0 + ld3(s.colors + size_t(i1) * 3) * b1 +
                ld3(s.colors + size_t(i2) * 3) * b2;
        has_color = true;
      }
      const uint32_t mid = s.material_ids ? s.material_ids[prim] : 0u;
      const vec3 wo = -rd;
      // pathtracer.cpp:906-910
      if (!(dot(hsn, hsn) > 0.25f)) hsn = ng;
      else if (dot(hsn, ng) < 0.f) hsn = hsn * -1.f;
      float alpha = 1.f;
      const phx_material mat = resolve_material(s, mid, uv, has_color, color, alpha);
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
        cont = true;
        no = f4(p, ntmin);
        nd = f4(rd, ntmax);
        ntp = f4(tp, -1.f);
        nst = make_uint4(slot, depth | (primary ? kPrimaryBit : 0u), kinds, __float_as_uint(path_dist));
      } else {
        vec3 n = ng, sn = hsn;
        const bool front = dot(n, wo) >= 0.f;
        if (!(mat.transmission > 0.f) && !front) {
          n = n * -1.f;
          sn = sn * -1.f;
        }
        if (mat.normal_tex >= 0 && s.tex_count > 0) {
          vec3 nm = tex_sample(s, mat.normal_tex, uv.x, uv.y);
          nm = nm * 2.f - vec3{1.f, 1.f, 1.f};
          vec3 ta, bi;
          onb(sn, ta, bi);
          sn = normalize(ta * nm.x + bi * nm.y + sn * nm.z);
        }
        if (primary) {
          primary = false;
          L.w += 1.f;
          Ldirty = true;
          q.slot_aov0[slot] = f4(normalize(sn), path_dist + ht);
          q.slot_aov1[slot] = f4(base_of(mat), 1.f);
        }
        if (material_may_emit(mat) && emits_on_side(mat, front)) {
          float wm = 1.f;
          if (prev_pdf > 0.f && s.em_count > 0) {
            const float pa = s.prim_light_pdf[prim];
            const float cl = fabsf(dot(ng, wo));
            if (pa > 0.f && cl > 0.f) wm = power_heuristic(prev_pdf, pa * ht * ht / cl / float(s.strategy_count));
          }
          L = f4(xyz(L) + clamp_contribution(tp * material_emission(mat) * wm, int(depth) - 1, w.clamp_direct,
                                             w.clamp_indirect),
                 L.w);
          Ldirty = true;
        }
        if (!((mat.flags & PHX_MATERIAL_FLAG_EMITTER_ONLY) != 0 || depth >= w.max_depth || emission_only)) {
          // Next event estimation (:1006-1035).
          const Direct ds = sample_direct(s, p, rng);
          if (ds.valid) {
            const vec3 f = bsdf_eval(mat, wo, ds.wi, n, sn);
            if (f.x > 0.f || f.y > 0.f || f.z > 0.f) {
              float wm = 1.f;
              if (ds.mis) wm = power_heuristic(ds.pdf, bsdf_pdf(mat, wo, ds.wi, n, sn));
              const vec3 c = clamp_contribution(tp * f * ds.L * (wm / ds.pdf), int(depth), w.clamp_direct,
                                                w.clamp_indirect);
              const vec3 sorg = p + n * s.ray_eps;
              vec3 sdir = ds.wi;
              float stmax = 1e30f;
              if (ds.dist < 1e29f) {
                const vec3 tl = p + ds.wi * ds.dist - sorg;
                const float sdist = length(tl);
                sdir = sdist > 0.f ? tl * (1.f / sdist) : ds.wi;
                stmax = sdist * (1.f - 1e-3f);
              }
              shadow = true;
              so = f4(sorg, 0.f);
              sd = f4(sdir, stmax);
              sc = f4(c, __uint_as_float(slot));
            }
          }
          // BSDF sample, Russian roulette, next ray (:1037-1089).
          const BsdfSample bs = bsdf_sample(mat, wo, n, sn, rng);
          if (!(bs.pdf <= 0.f || (bs.f.x <= 0.f && bs.f.y <= 0.f && bs.f.z <= 0.f))) {
            vec3 tpn = tp * (bs.f * (1.f / bs.pdf));
            bool alive = true;
            if (depth >= 2) {
              const float surv = fminf(fmaxf(max3(tpn), 0.05f), 0.95f);
              if (rng.next() >= surv) alive = false;
              else tpn = tpn * (1.f / surv);
            }
            if (alive && w.throughput_clamp > 0.f) {
              const float mm = max3(tpn);
              if (mm > w.throughput_clamp) tpn = tpn * (w.throughput_clamp / mm);
            }
            if (alive && finite3(tpn)) {
              const vec3 wi = normalize(bs.wi);
              const float side = dot(wi, ng) < 0.f ? -1.f : 1.f;
              const bool through = (dot(wi, ng) < 0.f) != (dot(wo, ng) < 0.f);
              const bool exceeded = through ? count_bounce(kinds, 16u, w.max_transmission)
                                            : (bs.diffuse ? count_bounce(kinds, 0u, w.max_diffuse)
                                                          : count_bounce(kinds, 8u, w.max_glossy));
              if (exceeded) kinds |= kEmissionOnly;
              cont = true;
              no = f4(p + ng * (side * s.ray_eps), 0.f);
              nd = f4(wi, 1e30f);
              ntp = f4(tpn, bs.specular ? -1.f : bs.pdf);
              nst = make_uint4(slot, depth + 1u, kinds, __float_as_uint(path_dist));
            }
          }
        }
      }
    }
    if (Ldirty) q.slot_L[slot] = L;
  }

  wave_count(&q.stats[kStatShade], valid);
  const uint32_t ci = wave_alloc(&q.counters[kRayCount + it + 1], cont);
  if (cont) {
    q.o[nbuf][ci] = no;
    q.d[nbuf][ci] = nd;
    q.tp[nbuf][ci] = ntp;
    q.st[nbuf][ci] = nst;
  }
  const uint32_t si = wave_alloc(&q.counters[kShadowCount + it], shadow);
  if (shadow) {
    q.sh_o[si] = so;
    q.sh_d[si] = sd;
    q.sh_c[si] = sc;
  }
}

// trace_occluded + pt_nee_accum_pack (pathtracer.cpp:1092-1123).
__global__ __launch_bounds__(256) void k_shadow(DevScene s, WaveParams w, WaveQueues q, uint32_t it) {
  __shared__ int lds[kShortStack * 256];
  const uint32_t count = q.counters[kShadowCount + it];
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  LaneStack st{&lds[threadIdx.x], 256, q.spill + size_t(gtid) * (kMaxStack - kShortStack), 0};
  const uint32_t lane = lane_id();
  while (true) {
    const uint32_t base = wave_fetch(&q.counters[kFetchShadow + it]);
    if (base >= count) break;
    const uint32_t i = base + lane;
    const bool valid = i < count;
    bool live = false;
    if (valid) {
      const float4 o = q.sh_o[i], d = q.sh_d[i], c = q.sh_c[i];
      live = d.w > o.w;
      TraceHit h;
      const bool occ = traverse<true>(s, xyz(o), xyz(d), o.w, d.w, st, h);
      if (!occ) {
        const uint32_t slot = __float_as_uint(c.w);
        vec3 contrib = xyz(c);
        if (s.vol_count > 0) {
          const uint32_t pixel = slot % w.npix;
          const uint32_t sample = w.sample_base + slot / w.npix;
          Rng rng = make_rng(pixel, sample, 0x10000u + it);
          contrib = contrib * vol_transmittance(s, xyz(o), xyz(d), o.w, d.w, rng);
        }
        if (finite3(contrib)) {
          float4 L = q.slot_L[slot];
          q.slot_L[slot] = make_float4(L.x + contrib.x, L.y + contrib.y, L.z + contrib.z, L.w);
        }
      }
    }
    wave_count(&q.stats[kStatShadow], live);
  }
}

// Folds the pass's k samples of each pixel into the frame, sample order.
__global__ __launch_bounds__(256) void k_accumulate(WaveParams w, WaveQueues q, FrameBuffers f, uint32_t first_pass) {
  const uint32_t pixel = blockIdx.x * blockDim.x + threadIdx.x;
  if (pixel >= w.npix) return;
  float4 acc = first_pass ? make_float4(0.f, 0.f, 0.f, 0.f) : f.accum[pixel];
  float4 a0 = first_pass ? make_float4(0.f, 0.f, 0.f, 1e30f) : f.aov0[pixel];
  float4 a1 = first_pass ? make_float4(0.f, 0.f, 0.f, 0.f) : f.aov1[pixel];
  for (uint32_t sl = 0; sl < w.k; ++sl) {
    const uint32_t slot = sl * w.npix + pixel;
    const float4 L = q.slot_L[slot];
    acc.x += L.x;
    acc.y += L.y;
    acc.z += L.z;
    acc.w += L.w;
    if (a1.w == 0.f) {
      const float4 s1 = q.slot_aov1[slot];
      if (s1.w != 0.f) {
        a0 = q.slot_aov0[slot];
        a1 = s1;
      }
    }
  }
  f.accum[pixel] = acc;
  f.aov0[pixel] = a0;
  f.aov1[pixel] = a1;
}

// pt_resolve (pathtracer.cpp:1143-1155): mean over the samples, AOVs.
__global__ __launch_bounds__(256) void k_resolve(uint32_t npix, uint32_t ns, FrameBuffers f) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  const float inv = ns > 0 ? 1.f / float(ns) : 0.f;
  const float4 a = f.accum[p], a0 = f.aov0[p], a1 = f.aov1[p];
  f.color[p * 3] = a.x * inv;
  f.color[p * 3 + 1] = a.y * inv;
  f.color[p * 3 + 2] = a.z * inv;
  f.alpha[p] = a.w * inv;
  f.depth[p] = a0.w;
  f.normal[p * 3] = a0.x;
  f.normal[p * 3 + 1] = a0.y;
  f.normal[p * 3 + 2] = a0.z;
  f.albedo[p * 3] = a1.x;
  f.albedo[p * 3 + 1] = a1.y;
  f.albedo[p * 3 + 2] = a1.z;
}

// ---- host launchers -----------------------------------------------------------
hipError_t launch_reset(const WaveQueues &q, uint32_t nslots, hipStream_t st) {
  hipLaunchKernelGGL(k_reset, dim3((kCounterWords + 255) / 256), dim3(256), 0, st, q, nslots);
  return hipGetLastError();
}
hipError_t launch_generate(const WaveParams &w, const WaveQueues &q, hipStream_t st) {
  hipLaunchKernelGGL(k_generate, dim3((w.nslots + 255) / 256), dim3(256), 0, st, w, q);
  return hipGetLastError();
}
hipError_t launch_extend(const DevScene &s, const WaveQueues &q, uint32_t it, uint32_t grid, hipStream_t st) {
  hipLaunchKernelGGL(k_extend, dim3(grid), dim3(256), 0, st, s, q, it);
  return hipGetLastError();
}
hipError_t launch_shade(const DevScene &s, const WaveParams &w, const WaveQueues &q, uint32_t it, hipStream_t st) {
  hipLaunchKernelGGL(k_shade, dim3((w.nslots + 255) / 256), dim3(256), 0, st, s, w, q, it);
  return hipGetLastError();
}
hipError_t launch_shadow(const DevScene &s, const WaveParams &w, const WaveQueues &q, uint32_t it, uint32_t grid,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_shadow, dim3(grid), dim3(256), 0, st, s, w, q, it);
  return hipGetLastError();
}
hipError_t launch_accumulate(const WaveParams &w, const WaveQueues &q, const FrameBuffers &f, bool first,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_accumulate, dim3((w.npix + 255) / 256), dim3(256), 0, st, w, q, f, first ? 1u : 0u);
  return hipGetLastError();
}
hipError_t launch_resolve(uint32_t npix, uint32_t ns, const FrameBuffers &f, hipStream_t st) {
  hipLaunchKernelGGL(k_resolve, dim3((npix + 255) / 256), dim3(256), 0, st, npix, ns, f);
  return hipGetLastError();
}

}  // namespace phx
