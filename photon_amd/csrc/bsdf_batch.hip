// Batch Disney BSDF entry points: the device BSDF of the shading kernels on
// caller-given (material, wo, wi, n, sn) tuples, for parity checks against
// disney_bsdf_eval / disney_bsdf_pdf / disney_bsdf_sample (disney_bsdf.h:288-521).
#include "context.h"
#include "dev_bsdf.h"
#include "dev_bsdf_ad.h"

namespace {

using namespace phx;
using namespace phx_detail;

PHX_DEV vec3 ldv(const float *p, uint32_t i) { return {p[i * 3], p[i * 3 + 1], p[i * 3 + 2]}; }

__global__ void k_bsdf_eval(const phx_material *mats, const float *wo, const float *wi, const float *ng,
                            const float *sn, uint32_t n, float *f, float *pdf) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const phx_material m = mats[i];
  const vec3 a = ldv(wo, i), b = ldv(wi, i), c = ldv(ng, i), d = ldv(sn, i);
  const vec3 r = bsdf_eval(m, a, b, c, d);
  f[i * 3] = r.x;
  f[i * 3 + 1] = r.y;
  f[i * 3 + 2] = r.z;
  pdf[i] = bsdf_pdf(m, a, b, c, d);
}

__global__ void k_bsdf_sample(const phx_material *mats, const float *wo, const float *ng, const float *sn,
                              const uint32_t *seeds, uint32_t n, float *wi, float *f, float *pdf, uint32_t *flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Rng rng{seeds[i] ? seeds[i] : 1u};
  const BsdfSample s = bsdf_sample(mats[i], ldv(wo, i), ldv(ng, i), ldv(sn, i), rng);
  wi[i * 3] = s.wi.x;
  wi[i * 3 + 1] = s.wi.y;
  wi[i * 3 + 2] = s.wi.z;
  f[i * 3] = s.f.x;
  f[i * 3 + 1] = s.f.y;
  f[i * 3 + 2] = s.f.z;
  pdf[i] = s.pdf;
  flags[i] = (s.specular ? 1u : 0u) | (s.diffuse ? 2u : 0u);
}

__global__ void k_bsdf_eval_grad(const phx_material *mats, const float *wo, const float *wi, const float *ng,
                                 const float *sn, uint32_t n, float *f, float *df) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const phx_material m = mats[i];
  const DV3 r = bsdf_eval_d(m, ldv(wo, i), ldv(wi, i), ldv(ng, i), ldv(sn, i));
  f[i * 3] = r.x.v;
  f[i * 3 + 1] = r.y.v;
  f[i * 3 + 2] = r.z.v;
  for (int k = 0; k < kGradParams; ++k) {
    df[i * 15 + k * 3] = r.x.d[k];
    df[i * 15 + k * 3 + 1] = r.y.d[k];
    df[i * 15 + k * 3 + 2] = r.z.d[k];
  }
}

}  // namespace

extern "C" {

int phx_bsdf_eval_grad(phx_context *c, const phx_material *mats, const float *wo, const float *wi, const float *ng,
                       const float *sn, uint32_t n, float *f, float *df, void *stream) {
  if (!c || (n && (!mats || !wo || !wi || !ng || !sn || !f || !df))) return PHX_ERR_INVALID;
  if (n == 0) return PHX_OK;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  hipLaunchKernelGGL(k_bsdf_eval_grad, dim3((n + 255) / 256), dim3(256), 0, st, mats, wo, wi, ng, sn, n, f, df);
  HIP_TRY(c, hipGetLastError());
  return PHX_OK;
}

int phx_bsdf_eval(phx_context *c, const phx_material *mats, const float *wo, const float *wi, const float *ng,
                  const float *sn, uint32_t n, float *f, float *pdf, void *stream) {
  if (!c || (n && (!mats || !wo || !wi || !ng || !sn || !f || !pdf))) return PHX_ERR_INVALID;
  if (n == 0) return PHX_OK;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  hipLaunchKernelGGL(k_bsdf_eval, dim3((n + 255) / 256), dim3(256), 0, st, mats, wo, wi, ng, sn, n, f, pdf);
  HIP_TRY(c, hipGetLastError());
  return PHX_OK;
}

int phx_bsdf_sample(phx_context *c, const phx_material *mats, const float *wo, const float *ng, const float *sn,
                    const uint32_t *seeds, uint32_t n, float *wi, float *f, float *pdf, uint32_t *flags,
                    void *stream) {
  if (!c || (n && (!mats || !wo || !ng || !sn || !seeds || !wi || !f || !pdf || !flags))) return PHX_ERR_INVALID;
  if (n == 0) return PHX_OK;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  hipLaunchKernelGGL(k_bsdf_sample, dim3((n + 255) / 256), dim3(256), 0, st, mats, wo, ng, sn, seeds, n, wi, f, pdf,
                     flags);
  HIP_TRY(c, hipGetLastError());
  return PHX_OK;
}

}  // extern "C"
