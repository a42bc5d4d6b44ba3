// Wavefront state of the integrator and the launchers of its kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "dev_scene.h"

namespace phx {

// Counter words (uint32) of one pass.  Iteration `it` reads the ray queue
// counted by kRayCount + it and the shadow queue counted by kShadowCount + it;
// the persistent traversal kernels fetch work through kFetch* + it.  All are
// cleared by k_reset at the start of a pass, so no host read or memset sits
// inside the bounce loop.
constexpr uint32_t kMaxIterations = 255;
constexpr uint32_t kRayCount = 0;
constexpr uint32_t kShadowCount = kRayCount + kMaxIterations + 1;
constexpr uint32_t kFetchExtend = kShadowCount + kMaxIterations + 1;
constexpr uint32_t kFetchShadow = kFetchExtend + kMaxIterations + 1;
constexpr uint32_t kCounterWords = kFetchShadow + kMaxIterations + 1;

constexpr uint32_t kStatClosest = 0;
constexpr uint32_t kStatShadow = 1;
constexpr uint32_t kStatShade = 2;
constexpr uint32_t kStatWords = 8;

struct WaveParams {
  phx_camera cam;
  uint32_t width, height, npix;
  uint32_t nslots;       // npix * k
  uint32_t k;            // samples per pixel in this pass
  uint32_t sample_base;  // sample index of slot group 0 (includes sample_offset)
  uint32_t max_depth;
  int32_t bottom_up;
  float region[4];
  vec3 bg;
  float bg_alpha;
  vec3 ambient;
  float throughput_clamp, clamp_direct, clamp_indirect;
  int32_t max_diffuse, max_glossy, max_transmission, max_volume;
};

// Queue storage (capacity = path slots of a pass).  Ray queue entries:
// o = (origin, tmin), d = (direction, tmax), tp = (throughput, prev_pdf),
// st = (slot, depth | primary << 16, bounce kinds, path distance bits).
struct WaveQueues {
  float4 *o[2];
  float4 *d[2];
  float4 *tp[2];
  uint4 *st[2];
  float4 *hit;      // (t, b1, b2, leaf triangle bits)
  float4 *sh_o;     // (origin, tmin)
  float4 *sh_d;     // (direction, tmax)
  float4 *sh_c;     // (NEE contribution, slot bits)
  float4 *slot_L;   // (radiance, alpha) per slot
  float4 *slot_aov0;  // (normal, depth)
  float4 *slot_aov1;  // (albedo, written)
  uint32_t *counters;
  unsigned long long *stats;
  int *spill;       // per-thread traversal stack spill
};

struct FrameBuffers {
  float4 *accum;  // (radiance sum, alpha sum)
  float4 *aov0;   // (normal, depth)
  float4 *aov1;   // (albedo, written)
  float *color, *alpha, *depth, *normal, *albedo;
};

hipError_t launch_reset(const WaveQueues &q, uint32_t nslots, hipStream_t st);
hipError_t launch_generate(const WaveParams &w, const WaveQueues &q, hipStream_t st);
hipError_t launch_extend(const DevScene &s, const WaveQueues &q, uint32_t it, uint32_t grid, hipStream_t st);
hipError_t launch_shade(const DevScene &s, const WaveParams &w, const WaveQueues &q, uint32_t it, hipStream_t st);
hipError_t launch_shadow(const DevScene &s, const WaveParams &w, const WaveQueues &q, uint32_t it, uint32_t grid,
                         hipStream_t st);
hipError_t launch_accumulate(const WaveParams &w, const WaveQueues &q, const FrameBuffers &f, bool first,
                             hipStream_t st);
hipError_t launch_resolve(uint32_t npix, uint32_t ns, const FrameBuffers &f, hipStream_t st);

// Seam A kernels (trace.hip).
hipError_t launch_trace_compact(const DevScene &s, const float *org, const float *dir, const float *tmin,
                                const float *tmax, uint32_t n, phx_hit_compact *out, uint32_t *fetch, int *spill,
                                uint32_t grid, hipStream_t st);
hipError_t launch_trace_full(const DevScene &s, const float *org, const float *dir, const float *tmin,
                             const float *tmax, uint32_t n, phx_hit *out, uint32_t *fetch, int *spill, uint32_t grid,
                             hipStream_t st);
hipError_t launch_trace_occluded(const DevScene &s, const float *org, const float *dir, const float *tmin,
                                 const float *tmax, uint32_t n, uint32_t *out, uint32_t *fetch, int *spill,
                                 uint32_t grid, hipStream_t st);
hipError_t launch_clear_u32(uint32_t *p, uint32_t n, hipStream_t st);

}  // namespace phx
