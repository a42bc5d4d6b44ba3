// phx_render: PathTracer::render (pathtracer.cpp:564-1159) on the streaming
// wavefront kernels of integrator.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "context.h"

namespace {

using namespace phx;
using namespace phx_detail;

// Per-sample radiance buffer budget of one streaming run (16 B per sample):
// larger frames are rendered in several runs of whole samples per pixel.
constexpr uint64_t kSampleBudget = uint64_t(1) << 27;
// Iterations launched between two reads of the queue counters.
constexpr uint32_t kIterBatch = 8;
// Every path ends within max_it iterations and at least one sample starts per
// iteration, so a run needs at most total + max_it iterations; this guard only
// turns a logic error into an error code instead of a hang.
constexpr uint32_t kIterationGuard = 1u << 26;

}  // namespace

namespace phx {

ShadeParams make_shade_params(const phx_camera &cam, const phx_render_params &p, uint32_t max_depth, uint32_t max_it) {
  ShadeParams w{};
  w.cam = cam;
  w.width = p.width;
  w.height = p.height;
  w.npix = p.width * p.height;
  w.max_depth = max_depth;
  w.max_it = max_it;
  w.spp = std::max<uint32_t>(p.samples_per_pixel, 1);
  w.sample_offset = p.sample_offset;
  w.bottom_up = p.bottom_up;
  for (int i = 0; i < 4; ++i) w.region[i] = p.image_region[i];
  w.bg = vec3{p.background_color[0], p.background_color[1], p.background_color[2]};
  w.bg_alpha = p.background_alpha;
  w.ambient = vec3{p.ambient_color[0] * p.ambient_radiance, p.ambient_color[1] * p.ambient_radiance,
                   p.ambient_color[2] * p.ambient_radiance};
  w.throughput_clamp = p.throughput_clamp;
  w.clamp_direct = p.clamp_direct_sum;
  w.clamp_indirect = p.clamp_indirect_sum;
  w.max_diffuse = p.max_diffuse_bounces;
  w.max_glossy = p.max_glossy_bounces;
  w.max_transmission = p.max_transmission_bounces;
  w.max_volume = p.max_volume_bounces;
  return w;
}

}  // namespace phx

extern "C" int phx_render(phx_context *c, const phx_camera *cam, const phx_render_params *p) {
  if (!c || !cam || !p) return PHX_ERR_INVALID;
  if (!c->have_scene) return fail(c, PHX_ERR_NO_SCENE, "render before set_scene");
  if (p->width == 0 || p->height == 0) return fail(c, PHX_ERR_INVALID, "render: empty frame");
  if (p->adaptive_sampling) return fail(c, PHX_ERR_UNSUPPORTED, "render: adaptive sampling is not offered");
  const uint64_t npix64 = uint64_t(p->width) * p->height;
  if (npix64 > (uint64_t(1) << 26)) return fail(c, PHX_ERR_INVALID, "render: frame larger than 2^26 pixels");
  const uint32_t npix = uint32_t(npix64);
  const uint32_t spp = std::max<uint32_t>(p->samples_per_pixel, 1);
  const uint32_t max_depth = p->max_depth > 0 ? p->max_depth : 1u;
  const uint32_t max_it = max_depth + 1u + 8u;  // pathtracer.cpp:682
  if (max_it > kMaxIterations) return fail(c, PHX_ERR_INVALID, "render: max_depth too large");
  HIP_TRY(c, hipSetDevice(c->device));
  const auto t0 = std::chrono::steady_clock::now();

  const ShadeParams sp = make_shade_params(*cam, *p, max_depth, max_it);
  const uint32_t spp_chunk = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(spp, kSampleBudget / npix)));
  const uint64_t chunk_total = uint64_t(npix) * spp_chunk;
  // 2^25 path slots (about 6 GB of queues): fewer, fuller launches; HBM is
  // plentiful on MI355X.
  const uint32_t want = c->tuning.paths_in_flight ? c->tuning.paths_in_flight : (1u << 25);
  const uint32_t cap = uint32_t(std::min<uint64_t>(want, chunk_total));

  if (cap > c->q_capacity) {
    for (int b = 0; b < 2; ++b) {
      HIP_TRY(c, c->q_o[b].reserve(size_t(cap) * 16));
      HIP_TRY(c, c->q_d[b].reserve(size_t(cap) * 16));
      HIP_TRY(c, c->q_tp[b].reserve(size_t(cap) * 16));
      HIP_TRY(c, c->q_st[b].reserve(size_t(cap) * 16));
    }
    HIP_TRY(c, c->q_hit.reserve(size_t(cap) * 16));
    HIP_TRY(c, c->sh_o.reserve(size_t(cap) * 16));
    HIP_TRY(c, c->sh_d.reserve(size_t(cap) * 16));
    HIP_TRY(c, c->sh_c.reserve(size_t(cap) * 16));
    c->q_capacity = cap;
  }
  HIP_TRY(c, c->sample_L.reserve(size_t(chunk_total) * 16));
  HIP_TRY(c, c->aov_s.reserve(size_t(npix) * 4));
  HIP_TRY(c, c->counters.reserve(kCounterWords * 4));
  HIP_TRY(c, c->stats_buf.reserve(kStatWords * 8));
  if (!c->pinned) HIP_TRY(c, hipHostMalloc(reinterpret_cast<void **>(&c->pinned), 64, hipHostMallocDefault));
  int r = ensure_trace_scratch(c);
  if (r) return r;
  if (c->f_w != p->width || c->f_h != p->height || !c->f_accum.p) {
    HIP_TRY(c, c->f_accum.reserve(size_t(npix) * 16));
    HIP_TRY(c, c->f_color.reserve(size_t(npix) * 12));
    HIP_TRY(c, c->f_alpha.reserve(size_t(npix) * 4));
    HIP_TRY(c, c->f_depth.reserve(size_t(npix) * 4));
    HIP_TRY(c, c->f_normal.reserve(size_t(npix) * 12));
    HIP_TRY(c, c->f_albedo.reserve(size_t(npix) * 12));
    c->f_w = p->width;
    c->f_h = p->height;
  }
  StreamQueues q{};
  for (int b = 0; b < 2; ++b) {
    q.o[b] = c->q_o[b].as<float4>();
    q.d[b] = c->q_d[b].as<float4>();
    q.tp[b] = c->q_tp[b].as<float4>();
    q.st[b] = c->q_st[b].as<uint4>();
  }
  q.hit = c->q_hit.as<float4>();
  q.sh_o = c->sh_o.as<float4>();
  q.sh_d = c->sh_d.as<float4>();
  q.sh_c = c->sh_c.as<float4>();
  q.sample_L = c->sample_L.as<float4>();
  q.aov_s = c->aov_s.as<uint32_t>();
  q.counters = c->counters.as<uint32_t>();
  q.stats = c->stats_buf.as<unsigned long long>();
  q.spill = c->spill.as<int>();
  FrameBuffers fb{c->f_accum.as<float4>(), c->f_color.as<float>(), c->f_alpha.as<float>(),
                  c->f_depth.as<float>(), c->f_normal.as<float>(), c->f_albedo.as<float>()};

  hipStream_t st = c->stream;
  const bool timed = c->tuning.time_kernels != 0;
  // PHX_TRACE_COUNTS=N: read the queue length every N iterations and print it.
  const char *tc = std::getenv("PHX_TRACE_COUNTS");
  const bool trace_counts = tc && std::atoi(tc) > 0;
  const uint32_t batch = trace_counts ? uint32_t(std::atoi(tc)) : kIterBatch;
  size_t ev = 0;
  auto mark = [&]() -> hipError_t {
    if (!timed) return hipSuccess;
    if (ev == c->events.size()) {
      hipEvent_t e;
      hipError_t err = hipEventCreate(&e);
      if (err != hipSuccess) return err;
      c->events.push_back(e);
    }
    return hipEventRecord(c->events[ev++], st);
  };

  HIP_TRY(c, launch_clear_u32(reinterpret_cast<uint32_t *>(q.stats), kStatWords * 2, st));
  HIP_TRY(c, launch_aov_init(npix, q, st));
  for (uint32_t s0 = 0; s0 < spp; s0 += spp_chunk) {
    StreamParams w{};
    w.sp = sp;
    const uint32_t nspp = std::min(spp_chunk, spp - s0);
    w.total = npix * nspp;
    w.capacity = std::min<uint32_t>(cap, w.total);
    w.sample_base = p->sample_offset + s0;
    w.chunk_s0 = s0;
    w.regen_when_empty = std::getenv("PHX_REGEN_WHEN_EMPTY") ? 1u : 0u;
    HIP_TRY(c, launch_reset(q, st));
    HIP_TRY(c, launch_regen(c->dev, w, q, 0, st));
    uint32_t it = 0;
    while (true) {
      for (uint32_t j = 0; j < batch; ++j, ++it) {
        HIP_TRY(c, mark());
        HIP_TRY(c, launch_extend(c->dev, q, it, c->trace_grid, st));
        HIP_TRY(c, mark());
        HIP_TRY(c, mark());
        HIP_TRY(c, launch_shade(c->dev, w, q, it, st));
        HIP_TRY(c, mark());
        HIP_TRY(c, mark());
        HIP_TRY(c, launch_shadow(c->dev, w, q, it, c->trace_grid, st));
        HIP_TRY(c, mark());
        HIP_TRY(c, launch_regen(c->dev, w, q, it + 1, st));
      }
      HIP_TRY(c, hipMemcpyAsync(&c->pinned[0], q.counters + kRayCount + ring(it), 4, hipMemcpyDeviceToHost, st));
      HIP_TRY(c, hipMemcpyAsync(&c->pinned[1], q.counters + kNextSample, 4, hipMemcpyDeviceToHost, st));
      HIP_TRY(c, hipStreamSynchronize(st));
      if (trace_counts)
        std::fprintf(stderr, "[phx] run %u iteration %u: queue %u, started %u of %u samples\n", s0, it, c->pinned[0],
                     c->pinned[1], w.total);
      if (c->pinned[0] == 0 && c->pinned[1] >= w.total) break;
      if (it >= kIterationGuard)
        return fail(c, PHX_ERR_OVERFLOW, "render: wavefront did not drain within its iteration guard");
    }
    HIP_TRY(c, launch_reduce(w, q, fb, s0 == 0, st));
  }
  HIP_TRY(c, launch_clear_u32(c->fetch.as<uint32_t>(), 1, st));
  HIP_TRY(c, launch_resolve(c->dev, sp, q, fb, spp, c->fetch.as<uint32_t>(), c->trace_grid, st));
  unsigned long long hs[kStatWords];
  HIP_TRY(c, hipMemcpyAsync(hs, q.stats, sizeof hs, hipMemcpyDeviceToHost, st));
  HIP_TRY(c, hipStreamSynchronize(st));
  const auto t1 = std::chrono::steady_clock::now();
  c->stats.closest_rays = hs[kStatClosest];
  c->stats.shadow_rays = hs[kStatShadow];
  c->stats.shade_events = hs[kStatShade];
  c->stats.samples = uint64_t(npix) * spp;
  c->stats.pixels = npix;
  c->stats.last_render_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  c->stats.extend_launches = c->stats.shade_launches = c->stats.shadow_launches = 0;
  c->stats.extend_ms = c->stats.shade_ms = c->stats.shadow_ms = 0.0;
  for (size_t e = 0; e + 5 < ev; e += 6) {
    float a = 0.f, b = 0.f, d = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&a, c->events[e], c->events[e + 1]));
    HIP_TRY(c, hipEventElapsedTime(&b, c->events[e + 2], c->events[e + 3]));
    HIP_TRY(c, hipEventElapsedTime(&d, c->events[e + 4], c->events[e + 5]));
    c->stats.extend_ms += a;
    c->stats.shade_ms += b;
    c->stats.shadow_ms += d;
    c->stats.extend_launches++;
    c->stats.shade_launches++;
    c->stats.shadow_launches++;
  }
  c->have_result = true;
  return PHX_OK;
}
