// Path-replay adjoint of the wavefront integrator: gradients of an image loss
// with respect to the trainable material parameters (base colour, GGX
// roughness, metallic) and the albedo texels of the texture atlas.
//
// Estimator.  Sampling is detached: every random decision (light pick, BSDF
// lobe and direction, Russian roulette, transparency) is the primal's, made
// with the current parameters; the parameters enter only through the BSDF
// weights f * cos at next-event estimation and at the sample that continues a
// path.  A path's radiance is L = sum_k C_k, each C_k a product of such weights
// and parameter-free factors, so
//   dL/dtheta = sum_j  tp_j (.) df^nee_j/dtheta (.) K_j      (visible NEE at j)
//             + sum_j  R_j (.) (df_j/dtheta) / f_j           (continuing sample at j)
// where R_j is the radiance the path collects after vertex j.  The forward
// pass (the ordinary kernels) stores each sample's L; the replay regenerates
// the same paths from their seeds and walks them again, subtracting each
// contribution from the stored L so that R_j is known at vertex j, and tests
// NEE visibility inline.  Nothing per vertex is stored.
#include "dev_bsdf_ad.h"
#include "dev_queue.h"
#include "dev_stream.h"

namespace phx {

// Derivative observer of shade_path for one sample (see NoHooks).
struct AdjHooks {
  vec3 delta;  // dLoss / dL of this sample
  bool nee_on, bsdf_on;
  float nee_g[kGradParams];    // delta-contracted NEE derivative, before visibility
  float u[3][kGradParams];     // delta_c (df_c/dp) / f_c of the continuing sample
  uint32_t mid;
  vec2 uv;
  int tex_id;
  bool has_color;
  PHX_DEV void note(const phx_material &m, const SurfaceHit &sh) {
    mid = sh.mid;
    uv = sh.uv;
    tex_id = m.base_color_tex;
    has_color = sh.has_color;
  }
  PHX_DEV void nee(const phx_material &m, const SurfaceHit &sh, vec3 wo, vec3 wi, vec3 n, vec3 sn, vec3 tp, vec3 K) {
    const DV3 f = bsdf_eval_d(m, wo, wi, n, sn);
    const vec3 c = delta * tp * K;
#pragma unroll
    for (int k = 0; k < kGradParams; ++k) nee_g[k] = c.x * f.x.d[k] + c.y * f.y.d[k] + c.z * f.z.d[k];
    nee_on = true;
    note(m, sh);
  }
  PHX_DEV void bsdf(const phx_material &m, const SurfaceHit &sh, vec3 wo, vec3 n, vec3 sn, vec3,
                    const BsdfSample &bs) {
    if (bs.specular) return;  // delta and transmission weights are not differentiated
    const DV3 f = bsdf_eval_d(m, wo, bs.wi, n, sn);
    const float ix = bs.f.x != 0.f ? delta.x / bs.f.x : 0.f;
    const float iy = bs.f.y != 0.f ? delta.y / bs.f.y : 0.f;
    const float iz = bs.f.z != 0.f ? delta.z / bs.f.z : 0.f;
#pragma unroll
    for (int k = 0; k < kGradParams; ++k) {
      u[0][k] = ix * f.x.d[k];
      u[1][k] = iy * f.y.d[k];
      u[2][k] = iz * f.z.d[k];
    }
    bsdf_on = true;
    note(m, sh);
  }
};

PHX_DEV void grad_add(double *lds, uint32_t lds_mats, double *global, uint32_t mid, int k, float g) {
  if (g == 0.f) return;
  if (mid < lds_mats) atomicAdd(&lds[mid * kGradParams + k], double(g));
  else unsafeAtomicAdd(&global[size_t(mid) * kGradParams + k], double(g));
}

// Routes one vertex's parameter gradient: base colour to the material or, for
// a texture-mapped base colour, to the four texels the lookup blends;
// roughness and metallic to the material while they lie inside the (0, 1)
// range resolve_material clamps them to.  A parameter that comes from a
// vertex colour, an attribute or a sampler is not the material's: it gets none.
PHX_DEV void grad_route(const float g[kGradParams], const AdjHooks &h, const DevScene &s, const AdjointArgs &a,
                        double *lds) {
  const phx_material &raw = s.materials[h.mid];
  const bool tex = s.tex_count > 0;
  if (h.tex_id >= 0 && tex) {
    if (a.grad_tex) {
      const TexTaps t = tex_taps(s, h.tex_id, h.uv.x, h.uv.y);
      if (t.valid) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int c = 0; c < 3; ++c)
            if (g[c] != 0.f && t.w[j] != 0.f)
              unsafeAtomicAdd(&a.grad_tex[size_t(t.idx[j]) * 3 + c], double(g[c] * t.w[j]));
      }
    }
  } else if (!h.has_color && raw.base_color_src.kind == PHX_SOURCE_CONSTANT) {
    for (int c = 0; c < 3; ++c) grad_add(lds, a.lds_mats, a.grad_mat, h.mid, c, g[c]);
  }
  if (!(raw.roughness_tex >= 0 && tex) && raw.roughness_src.kind == PHX_SOURCE_CONSTANT && raw.roughness > 0.f &&
      raw.roughness < 1.f)
    grad_add(lds, a.lds_mats, a.grad_mat, h.mid, 3, g[3]);
  if (!(raw.metallic_tex >= 0 && tex) && raw.metallic_src.kind == PHX_SOURCE_CONSTANT && raw.metallic > 0.f &&
      raw.metallic < 1.f)
    grad_add(lds, a.lds_mats, a.grad_mat, h.mid, 4, g[4]);
}

// Replay of one wavefront iteration: pt_shade with derivative observers,
// inline NEE visibility, gradient accumulation (per work-group in LDS for the
// first lds_mats materials, flushed once), continuing paths to the next queue.
// The work-group steps through tiles of B rays together: one striped fetch
// and one slot claim per tile, not one per wave.  Two waves per SIMD (256
// registers): uncapped, the standard tier took 256 registers and 3 AGPRs,
// one wave per SIMD, and its pass was 26 % slower than with a 12-byte spill.
template <class P, int B, int TIER>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(2))) void k_shade_adjoint(P v, StreamQueues q,
                                                                                             AdjointArgs a) {
  const uint32_t it = q.counters[kIter];
  const uint32_t count = q.counters[ray_count_word(it)];
  const uint32_t units = fetch_units(count, B);
  if (blockIdx.x >= units) return;  // work-groups beyond the fetchers take no part (count == 0 included)
  __shared__ int lds[kShortStack * B];
  __shared__ double glds[kAdjLdsMaterials * kGradParams];
  for (uint32_t k = threadIdx.x; k < a.lds_mats * kGradParams; k += B) glds[k] = 0.0;
  __syncthreads();
  const int buf = it & 1, nbuf = buf ^ 1;
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  LaneStack st{&lds[threadIdx.x], B, q.spill + size_t(gtid) * (kMaxStack - kShortStack), 0};
  // the replay runs no shadow kernel: its fetch set serves this loop
  uint32_t *set = &q.counters[fetch_set(kFetchShadow, it)];
  uint32_t stripe = blockIdx.x % kStripes, left = fetch_walk(min(units, gridDim.x));  // thread 0's cursor
  __shared__ uint32_t s_tile[2];
  while (true) {
    if (threadIdx.x == 0) fetch_tile(set, count, units, B, stripe, left, s_tile[0], s_tile[1]);
    __syncthreads();
    const uint32_t base = s_tile[0], end = s_tile[1];
    if (base == 0xffffffffu) break;  // uniform over the work-group
    const uint32_t i = base + threadIdx.x;
    const bool valid = i < end;
    ShadeOut out;
    out.cont = false;
    out.shadow = false;
    uint32_t g = 0;
    if (valid) {
      const uint4 stv = q.st[buf][i];
      g = stv.x;
      const RunItem &r = v.item_at(stv.y >> kItemShift);
      const DevScene &s = v.scene_of(r);
      PathState ps = unpack(q.o[buf][i], q.d[buf][i], q.tp[buf][i], stv);
      const float4 hit = q.hit[i];
      const uint32_t htri = hit_tri(__float_as_uint(hit.w));
      const uint32_t l = g - r.first;
      const uint32_t pixel = l % r.sp.npix;
      ps.pixel = pixel;
      const uint32_t sample = r.sample_base + l / r.sp.npix;
      AdjHooks h;
      h.nee_on = false;
      h.bsdf_on = false;
      const float *dp = a.adjoint + size_t(r.pix_base + pixel) * 3;
      h.delta = vec3{dp[0], dp[1], dp[2]} * (1.f / float(r.sp.spp));
      Rng rng = make_rng(pixel, sample, 1u + ps.path_it);
      shade_path<AdjHooks, false, TIER>(s, r.sp, ps, htri != 0xffffffffu, htri, hit.x, hit.y, hit.z, rng, out, h);
      const float4 Lr = q.sample_L[g];
      vec3 rem = xyz(Lr) - out.add;
      float gv[kGradParams];
#pragma unroll
      for (int k = 0; k < kGradParams; ++k) gv[k] = 0.f;
      if (out.shadow) {
        const DevScene sc = trace_scene(v, stv.y >> kItemShift);
        TraceHit th;
        if (!traverse<true>(sc, out.so, out.sd, 0.f, out.stmax, st, th)) {
          vec3 contrib = out.sc;
          float tr = 1.f;
          if (s.vol_count > 0) {
            Rng rv = make_rng(pixel, sample, 0x10000u + (out.next.path_it - 1u));
            tr = vol_transmittance(s, out.so, out.sd, 0.f, out.stmax, rv);
            contrib = contrib * tr;
          }
          if (finite3(contrib)) {
            rem = rem - contrib;
            if (h.nee_on) {
#pragma unroll
              for (int k = 0; k < kGradParams; ++k) gv[k] += h.nee_g[k] * tr;
            }
          }
        }
      }
      if (h.bsdf_on) {
#pragma unroll
        for (int k = 0; k < kGradParams; ++k) gv[k] += h.u[0][k] * rem.x + h.u[1][k] * rem.y + h.u[2][k] * rem.z;
      }
      q.sample_L[g] = make_float4(rem.x, rem.y, rem.z, Lr.w);
      if (h.nee_on || h.bsdf_on) grad_route(gv, h, s, a, glds);
      if (out.next.path_it >= r.sp.max_it) out.cont = false;
    }
    // (s_tile is read above before this claim's barriers; thread 0 rewrites it after them)
    uint32_t ci = 0, unused = 0;
    block_alloc_pair<B / 64>(reinterpret_cast<unsigned long long *>(&q.counters[ray_count_word(it + 1)]), out.cont, ci,
                             false, unused);
    if (out.cont) {
      const PathState &n = out.next;
      q.o[nbuf][ci] = f4(n.ro, n.tmin);
      q.d[nbuf][ci] = f4(n.rd, n.tmax);
      q.tp[nbuf][ci] = f4(n.tp, n.prev_pdf);
      q.st[nbuf][ci] = make_uint4(g, pack_flags(n), n.kinds, __float_as_uint(n.path_dist));
    }
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < a.lds_mats * kGradParams; k += B)
    if (glds[k] != 0.0) unsafeAtomicAdd(&a.grad_mat[k], glds[k]);
}

__global__ void k_grad_out(const double *src, float *dst, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] += float(src[i]);
}
__global__ void k_clear_f64(double *p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// The parameters of materials 0..count-1 from a device array of count * 5
// floats (base r, g, b, roughness, metallic).
__global__ void k_set_material_params(phx_material *mats, const float *p, uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  mats[i].base_color[0] = p[i * 5];
  mats[i].base_color[1] = p[i * 5 + 1];
  mats[i].base_color[2] = p[i * 5 + 2];
  mats[i].roughness = p[i * 5 + 3];
  mats[i].metallic = p[i * 5 + 4];
}

hipError_t launch_shade_adjoint(const ViewArgs &v, const StreamQueues &q, const AdjointArgs &a, const LoopShape &ls,
                                hipStream_t st) {
  auto go = [&](const auto &p) {
    using P = std::decay_t<decltype(p)>;
    // runs of the lean tier replay with the standard tier's code (one fewer instantiation)
    const bool full = ls.tier == kTierFull;
    const dim3 g(ls.trace_grid);
    if (ls.trace_block == 128 && full) hipLaunchKernelGGL((k_shade_adjoint<P, 128, kTierFull>), g, dim3(128), 0, st, p, q, a);
    else if (ls.trace_block == 128) hipLaunchKernelGGL((k_shade_adjoint<P, 128, kTierStandard>), g, dim3(128), 0, st, p, q, a);
    else if (full) hipLaunchKernelGGL((k_shade_adjoint<P, 256, kTierFull>), g, dim3(256), 0, st, p, q, a);
    else hipLaunchKernelGGL((k_shade_adjoint<P, 256, kTierStandard>), g, dim3(256), 0, st, p, q, a);
    return hipGetLastError();
  };
  switch (v.kind) {
    case 0: return go(v.one);
    case 1: return go(v.shared);
    default: return go(v.many);
  }
}
hipError_t launch_grad_out(const double *src, float *dst, uint32_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_grad_out, dim3((n + 255) / 256), dim3(256), 0, st, src, dst, n);
  return hipGetLastError();
}
hipError_t launch_clear_f64(double *p, uint32_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_clear_f64, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
  return hipGetLastError();
}
hipError_t launch_set_material_params(phx_material *mats, const float *p, uint32_t count, hipStream_t st) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_set_material_params, dim3((count + 255) / 256), dim3(256), 0, st, mats, p, count);
  return hipGetLastError();
}

}  // namespace phx
