// Display transforms on the GPU: exposure, the reference's ACES and Reinhard
// tone maps (tone_mapping.h:7-21), the sRGB or 1/2.2 transfer curve
// (tone_mapping.h:23-28, color.h:9-12) and the 8-bit quantisation of
// photon_render (photon_render.cpp:24-35); and the denoiser seam, which has
// nothing to run without Open Image Denoise (denoiser.cpp:8-14).
//
// One thread per pixel: the pass reads 12 or 16 bytes and writes 12 + 3 per
// pixel, so it runs at the HBM's pace; the arithmetic is a few dozen
// operations and two powf per channel at most.
#include <cmath>

#include "context.h"

namespace {

using namespace phx_detail;

__device__ __forceinline__ float clamp01f(float x) { return x < 0.f ? 0.f : (x > 1.f ? 1.f : x); }

__device__ __forceinline__ float tone(float c, int op) {
  if (op == PHX_TONEMAP_ACES) {
    // aces_tonemap: c (2.51 c + 0.03) / (c (2.43 c + 0.59) + 0.14), the
    // products and sums in the reference's order, no contraction
    const float num = c * (c * 2.51f + 0.03f);
    const float den = c * (c * 2.43f + 0.59f) + 0.14f;
    return num / den;
  }
  if (op == PHX_TONEMAP_REINHARD) return c / (1.f + c);
  return c;
}

__device__ __forceinline__ float curve(float x, int kind) {
  if (kind == PHX_CURVE_SRGB) return x <= 0.0031308f ? 12.92f * x : 1.055f * powf(x, 1.f / 2.4f) - 0.055f;
  if (kind == PHX_CURVE_GAMMA22) return powf(x, 1.f / 2.2f);
  return x;
}

__global__ __launch_bounds__(256) void k_display(const float *in, uint32_t channels, uint64_t n, float scale,
                                                 int op, int kind, int clamp, float *out, uint8_t *out8) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    float c[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float v = in[i * channels + k] * scale;  // exposure_adjust
      v = tone(v, op);
      if (clamp) v = clamp01f(v);
      v = curve(v, kind);
      if (clamp) v = clamp01f(v);
      c[k] = v;
    }
    if (out) {
#pragma unroll
      for (int k = 0; k < 3; ++k) out[i * 3 + k] = c[k];
    }
    if (out8) {
#pragma unroll
      for (int k = 0; k < 3; ++k) out8[i * 3 + k] = uint8_t(clamp01f(c[k]) * 255.999f);
    }
  }
}

}  // namespace

extern "C" {

int phx_display_transform(phx_context *c, const float *d_rgb, uint32_t channels, uint64_t pixels,
                          const phx_display *t, float *d_out, uint8_t *d_out8, void *stream) {
  if (!c || !t) return PHX_ERR_INVALID;
  if (channels != 3 && channels != 4) return fail(c, PHX_ERR_INVALID, "display_transform: 3 or 4 input channels");
  if (t->tonemap < PHX_TONEMAP_NONE || t->tonemap > PHX_TONEMAP_REINHARD)
    return fail(c, PHX_ERR_INVALID, "display_transform: unknown tone map");
  if (t->curve < PHX_CURVE_LINEAR || t->curve > PHX_CURVE_GAMMA22)
    return fail(c, PHX_ERR_INVALID, "display_transform: unknown transfer curve");
  if (pixels == 0 || (!d_out && !d_out8)) return PHX_OK;
  if (!d_rgb) return PHX_ERR_INVALID;
  // exposure_adjust's factor pow(2, exposure), once on the host
  const float scale = std::pow(2.f, t->exposure);
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  const uint64_t blocks = std::min<uint64_t>((pixels + 255) / 256, uint64_t(c->cu_count) * 32u);
  hipLaunchKernelGGL(k_display, dim3(uint32_t(blocks)), dim3(256), 0, st, d_rgb, channels, pixels, scale,
                     int(t->tonemap), int(t->curve), int(t->clamp), d_out, d_out8);
  HIP_TRY(c, hipGetLastError());
  return PHX_OK;
}

int phx_denoiser_is_available(void) { return 0; }

int phx_denoise_buffer(float *color, const float *albedo, const float *normal, uint32_t width, uint32_t height) {
  // Denoiser::denoise_buffer without OIDN leaves the colour as it is
  (void)albedo;
  (void)normal;
  if (!color && uint64_t(width) * height) return PHX_ERR_INVALID;
  return PHX_OK;
}

}  // extern "C"
