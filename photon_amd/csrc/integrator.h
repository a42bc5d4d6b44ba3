// Streaming wavefront integrator: state layout and kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include "dev_shade.h"

namespace phx {

// Counter words of one streaming run, in rings of kRing entries indexed by
// ring(it) for wavefront iteration `it` (a global step of the wavefront, not
// of a path): the ray queue of `it` is counted by kRayCount + ring(it), the
// shadow queue it fills by kShadowCount + ring(it); the persistent traversal
// kernels fetch work through kFetch* + ring(it).  kNextSample is the next
// sample to start.  k_reset clears everything at the start of a run and
// k_regen_commit clears the entries of it + 1 before they are used, so no
// host write or memset sits inside the bounce loop.
constexpr uint32_t kMaxIterations = 255;  // per path (RenderParams bound)
constexpr uint32_t kRing = 16;
constexpr uint32_t kRayCount = 0;
constexpr uint32_t kShadowCount = kRayCount + kRing;
constexpr uint32_t kFetchExtend = kShadowCount + kRing;
constexpr uint32_t kFetchShadow = kFetchExtend + kRing;
constexpr uint32_t kRegenTake = kFetchShadow + kRing;
constexpr uint32_t kNextSample = kRegenTake + kRing;
constexpr uint32_t kCounterWords = kNextSample + 4;
__host__ __device__ inline uint32_t ring(uint32_t it) { return it & (kRing - 1); }

constexpr uint32_t kStatClosest = 0;
constexpr uint32_t kStatShadow = 1;
constexpr uint32_t kStatShade = 2;
constexpr uint32_t kStatWords = 8;

// A streaming run renders samples [0, total) of the chunk: sample id
// sid = s * npix + pixel, with sample index sample_base + s.
struct StreamParams {
  ShadeParams sp;
  uint32_t total;        // samples in this run
  uint32_t capacity;     // path slots (queue length)
  uint32_t sample_base;  // sample index of s = 0 (sample_offset + first spp of the chunk)
  uint32_t chunk_s0;     // first spp index of the chunk (AOV bookkeeping)
  uint32_t regen_when_empty;  // 1: start new samples only into an empty queue (pass by pass)
};

// Queue entries: o = (origin, tmin), d = (direction, tmax),
// tp = (throughput, prev_pdf), st = (sample id, depth | path_it << 8 |
// primary << 16, bounce kinds, camera-ray distance bits).
struct StreamQueues {
  float4 *o[2];
  float4 *d[2];
  float4 *tp[2];
  uint4 *st[2];
  float4 *hit;       // (t, b1, b2, leaf triangle bits)
  float4 *sh_o;      // (origin, path_it bits); shadow rays start at t = 0
  float4 *sh_d;      // (direction, tmax)
  float4 *sh_c;      // (NEE contribution, sample id bits)
  float4 *sample_L;  // per sample of the run: (radiance, alpha)
  uint32_t *aov_s;   // per pixel: first sample index with a camera-ray event
  uint32_t *counters;
  unsigned long long *stats;
  int *spill;        // per-thread traversal stack spill
};

struct FrameBuffers {
  float4 *accum;  // (radiance sum, alpha sum)
  float *color, *alpha, *depth, *normal, *albedo;
};

// RenderParams -> the shading constants of one frame (render.cpp).
ShadeParams make_shade_params(const phx_camera &cam, const phx_render_params &p, uint32_t max_depth, uint32_t max_it);

hipError_t launch_reset(const StreamQueues &q, hipStream_t st);
hipError_t launch_regen(const DevScene &s, const StreamParams &w, const StreamQueues &q, uint32_t it, hipStream_t st);
hipError_t launch_extend(const DevScene &s, const StreamQueues &q, uint32_t it, uint32_t grid, hipStream_t st);
hipError_t launch_shade(const DevScene &s, const StreamParams &w, const StreamQueues &q, uint32_t it, hipStream_t st);
hipError_t launch_shadow(const DevScene &s, const StreamParams &w, const StreamQueues &q, uint32_t it, uint32_t grid,
                         hipStream_t st);
hipError_t launch_reduce(const StreamParams &w, const StreamQueues &q, const FrameBuffers &f, bool first,
                         hipStream_t st);
hipError_t launch_aov_init(uint32_t npix, const StreamQueues &q, hipStream_t st);
hipError_t launch_resolve(const DevScene &s, const ShadeParams &sp, const StreamQueues &q, const FrameBuffers &f,
                          uint32_t ns, uint32_t *fetch, uint32_t grid, hipStream_t st);

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
