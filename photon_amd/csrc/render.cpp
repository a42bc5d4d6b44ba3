// phx_render and phx_render_batch: PathTracer::render (pathtracer.cpp:564-1159)
// on the streaming wavefront kernels of integrator.hip, for one view or for
// many views sharing the path queues; phx_autotune picks their launch shape.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <utility>

#include "context.h"
#include "dev_queue.h"

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
  const bool bg = p.background_image && p.background_width > 0 && p.background_height > 0;
  w.bg_image = bg ? reinterpret_cast<const float4 *>(p.background_image) : nullptr;
  w.bg_w = bg ? p.background_width : 0u;
  w.bg_h = bg ? p.background_height : 0u;
  return w;
}

}  // namespace phx

namespace {

using namespace phx;
using namespace phx_detail;

// Per-sample radiance buffer budget of one streaming run (16 B per sample,
// 4 GB at most): larger jobs are rendered in several runs of whole samples
// per pixel, each paying the queue's drain at its end.
constexpr uint64_t kSampleBudget = uint64_t(1) << 28;
// Iterations launched (or replayed as one graph) between two reads of the
// queue counters.
constexpr uint32_t kIterBatch = 8;
// Every path ends within max_it iterations and at least one sample starts per
// iteration, so a run needs at most total + max_it iterations; this guard only
// turns a logic error into an error code instead of a hang.
constexpr uint32_t kIterationGuard = 1u << 26;
constexpr uint32_t kDefaultPaths = 1u << 26;
constexpr uint32_t kDefaultRefill = 16;
// Without a tuned refill threshold, a run's traversal refills at 12 idle lanes
// when its deepest BVH has at least 6 levels and at 40 otherwise: idle lanes
// cost more on long traversals, refills more on short ones.  Measured: the
// 737K-triangle frame (depth 9) 1 % faster at 12 than at 16; at 40 instead of
// 20, the Cornell box (depth 2) 4.3 %, the textured Cornell box 4.6 % and the
// fog box 1.5 % faster.
constexpr uint32_t kRefillDeep = 12, kRefillShallow = 40, kDeepTree = 6;
constexpr uint32_t kDefaultChunk = 128;
constexpr uint32_t kGradParamsAbi = 5;  // base r, g, b, roughness, metallic

// One view of a render or a batch.
struct ViewJob {
  const phx_context *scene;
  uint32_t scene_index;
  ShadeParams sp;
  uint32_t spp;
  uint32_t pix_base;
  bool adaptive;  // RenderParams::adaptive_sampling
};

// Item of a run on the host side: the run item plus what the reduce needs.
struct HostItem {
  RunItem r;
  uint32_t nspp;
};
struct Run {
  std::vector<HostItem> items;
  uint32_t total = 0;
};

int check_params(phx_context *c, const phx_render_params *p, const char *who) {
  if (p->width == 0 || p->height == 0) return fail(c, PHX_ERR_INVALID, std::string(who) + ": empty frame");
  if (uint64_t(p->width) * p->height > (uint64_t(1) << 26))
    return fail(c, PHX_ERR_INVALID, std::string(who) + ": frame larger than 2^26 pixels");
  const uint32_t max_depth = p->max_depth > 0 ? p->max_depth : 1u;
  if (max_depth + 9u > kMaxIterations) return fail(c, PHX_ERR_INVALID, std::string(who) + ": max_depth too large");
  if (p->background_image && (p->background_width == 0 || p->background_height == 0))
    return fail(c, PHX_ERR_INVALID, std::string(who) + ": background image without a size");
  if (p->background_image && (reinterpret_cast<uintptr_t>(p->background_image) & 15u))
    return fail(c, PHX_ERR_INVALID, std::string(who) + ": background image not 16-byte aligned");
  return PHX_OK;
}

ViewJob make_job(const phx_context *scene, uint32_t scene_index, const phx_camera &cam, const phx_render_params &p,
                 uint32_t pix_base) {
  const uint32_t max_depth = p.max_depth > 0 ? p.max_depth : 1u;
  const uint32_t max_it = max_depth + 1u + 8u;  // pathtracer.cpp:682
  return ViewJob{scene, scene_index, make_shade_params(cam, p, max_depth, max_it),
                 std::max<uint32_t>(p.samples_per_pixel, 1), pix_base, p.adaptive_sampling != 0};
}

// A view of at least this many samples keeps the GPU's path slots full by
// itself: in a batch it streams in runs of its own, through the single-view
// kernels (scene and view in kernel arguments, scalar loads), as phx_render
// would run it.  Only the smaller views share runs through the item tables,
// whose per-lane view and scene loads cost 10-18 % on the kernels of views
// large enough not to need the sharing (the 16-view draw: its five views of
// 2^22 samples or more took 93.4 ms together in one batch against 83.6 ms as
// plain renders).
constexpr uint64_t kSoloSamples = uint64_t(1) << 22;

// Samples one run may hold: kSampleBudget, or less under PHX_SAMPLE_BUDGET=N
// (diagnostic), to exercise the splitting of large jobs at small sizes.
uint64_t run_budget() {
  uint64_t budget = kSampleBudget;
  if (const char *e = std::getenv("PHX_SAMPLE_BUDGET")) {
    const long long b = std::atoll(e);
    if (b > 0) budget = std::min<uint64_t>(budget, uint64_t(b));
  }
  return budget;
}

// Splits the views' samples into runs of at most run_budget() samples, each
// a list of (view, first spp, spp count) items in view order; views of
// kSoloSamples or more get runs of their own, split as phx_render splits them.
std::vector<Run> plan_runs(const std::vector<ViewJob> &views);

// Runs of an adaptively sampled view (pathtracer.cpp:700-723).  A pixel's
// decision to take sample s >= 16 needs only its own samples before s, so the
// first 16 samples stream as usual and the later ones in groups of
// kAdaptiveGroup sample indices, whose regeneration skips the pixels that
// stopped before the group.  The reduce applies the stopping rule sample by
// sample inside a group and drops the samples after a pixel's stop: the
// result is the reference's for any group size; a pixel that stops inside a
// group costs the rest of the group's samples in work only.  (The reference
// decides for every pixel between two sample indices, one barrier each.)
// The rule tests the variance of single samples, which does not shrink as
// samples add up, so a pixel's estimate settles early: pixels stop at 16 or
// take every sample, and stops in between are rare (4 of 2304 pixels of a
// 48x48 Cornell view at depth 1).  One group for all later samples is then
// the fastest: on the 737K-triangle view at 1024x1024x64 spp (3.2 % of the
// pixels stop, all at 16) groups of 1, 8, 16 and 48 took 155.7, 55.4, 48.4 and
// 44.7 ms against 42.7 ms for the plain frame.
constexpr uint32_t kAdaptiveGroup = 0xffffffffu;
std::vector<Run> plan_adaptive_runs(const ViewJob &j) {
  ViewJob head = j;
  head.spp = std::min<uint32_t>(j.spp, 16u);
  head.adaptive = false;
  std::vector<Run> runs = plan_runs({head});
  // PHX_ADAPTIVE_GROUP=N (diagnostic): another group size, to time the
  // trade between fewer runs and samples past the stops
  uint64_t group = kAdaptiveGroup;
  if (const char *e = std::getenv("PHX_ADAPTIVE_GROUP")) {
    const long long g = std::atoll(e);
    if (g > 0) group = uint64_t(g);
  }
  group = std::max<uint64_t>(1, std::min<uint64_t>(group, run_budget() / j.sp.npix));
  for (uint32_t s = head.spp; s < j.spp;) {
    const uint32_t n = uint32_t(std::min<uint64_t>(group, j.spp - s));
    Run run;
    HostItem h{};
    h.r.sp = j.sp;
    h.r.scene = j.scene_index;
    h.r.first = 0;
    h.r.sample_base = j.sp.sample_offset + s;
    h.r.s0 = s;
    h.r.pix_base = j.pix_base;
    h.r.view = 0;
    h.nspp = n;
    run.items.push_back(h);
    run.total = j.sp.npix * n;
    runs.push_back(run);
    s += n;
  }
  return runs;
}

std::vector<Run> plan_runs(const std::vector<ViewJob> &views) {
  const uint64_t budget = run_budget();
  std::vector<Run> runs(1);
  std::vector<uint32_t> order;
  for (uint32_t v = 0; v < views.size(); ++v)
    if (views.size() > 1 && uint64_t(views[v].sp.npix) * views[v].spp >= kSoloSamples) order.push_back(v);
  const size_t nsolo = order.size();
  for (uint32_t v = 0; v < views.size(); ++v)
    if (!(views.size() > 1 && uint64_t(views[v].sp.npix) * views[v].spp >= kSoloSamples)) order.push_back(v);
  for (size_t o = 0; o < order.size(); ++o) {
    const uint32_t v = order[o];
    const ViewJob &j = views[v];
    // a solo view, and the first shared view after them, start a new run
    if ((o < nsolo || o == nsolo) && !runs.back().items.empty()) runs.emplace_back();
    const uint32_t npix = j.sp.npix;
    uint32_t s0 = 0;
    while (s0 < j.spp) {
      Run &run = runs.back();
      const uint64_t room = budget > run.total ? budget - run.total : 0;
      uint32_t nfit = uint32_t(std::min<uint64_t>(j.spp - s0, room / npix));
      if (nfit == 0) {
        if (!run.items.empty()) {
          runs.emplace_back();
          continue;
        }
        nfit = 1;  // a view whose single sample exceeds the budget runs alone
      }
      HostItem h{};
      h.r.sp = j.sp;
      h.r.scene = j.scene_index;
      h.r.first = run.total;
      h.r.sample_base = j.sp.sample_offset + s0;
      h.r.s0 = s0;
      h.r.pix_base = j.pix_base;
      h.r.view = v;
      h.nspp = nfit;
      run.items.push_back(h);
      run.total += npix * nfit;
      s0 += nfit;
    }
  }
  return runs;
}

LoopShape loop_shape(const phx_context *c, uint32_t capacity) {
  const phx_tuning &t = c->tuning;
  LoopShape ls{};
  ls.trace_block = t.trace_block == 128 ? 128u : 256u;
  ls.trace_grid = uint32_t(c->cu_count) * (t.trace_groups_per_cu ? t.trace_groups_per_cu : 8u);
  ls.shade_grid = std::max<uint32_t>(1, (capacity + 255) / 256);
  ls.regen_grid = std::max<uint32_t>(1, std::min<uint32_t>((capacity + 255) / 256, uint32_t(c->cu_count) * 16u));
  ls.chunk = t.queue_chunk ? t.queue_chunk : kDefaultChunk;
  ls.refill = t.refill_lanes ? t.refill_lanes : kDefaultRefill;
  ls.groups = std::min<uint32_t>(t.sort_by_material, kMaxGroups);
  return ls;
}

// Per-kernel timing of a run: event pairs around the closest-hit, shading and
// any-hit launches, each on the stream of its launch.
enum TimedKind { kTimedExtend = 0, kTimedShade = 1, kTimedShadow = 2 };
struct Timer {
  phx_context *c;
  bool on;
  size_t n = 0;            // events recorded
  std::vector<int> kinds;  // the kind of event pair k (events 2k and 2k + 1)
  hipError_t begin(int kind, hipStream_t st) {
    if (!on) return hipSuccess;
    kinds.push_back(kind);
    return record(st);
  }
  hipError_t end(hipStream_t st) { return on ? record(st) : hipSuccess; }
  hipError_t record(hipStream_t st) {
    if (n == c->events.size()) {
      hipEvent_t e;
      hipError_t err = hipEventCreate(&e);
      if (err != hipSuccess) return err;
      c->events.push_back(e);
    }
    return hipEventRecord(c->events[n++], st);
  }
};

// The second stream of the bounce loop and its fork and join events.
struct Fork {
  hipStream_t st2;
  hipEvent_t ev_fork, ev_join;
};

// Iterations of the bounce loop.  An iteration runs the any-hit pass over the
// shadow queue that the previous iteration's shading filled (k_shadow) beside
// its own closest-hit pass and shading, then adds that queue's unoccluded
// contributions (k_nee_accumulate) and regenerates.  With `fork` the any-hit
// pass runs on the second stream, forked after the previous regeneration and
// joined before k_nee_accumulate: it shares the GPU with k_extend and k_shade,
// so no pass's tail of last slow rays leaves the GPU idle.  Shading fills the
// other shadow queue of the pair (by iteration parity).  k_nee_accumulate
// must follow both the any-hit pass and shading, since shading adds to the
// same samples' radiance; a sample's sum thus takes the emission found by
// bounce i before the direct light of bounce i - 1, and every schedule (serial
// or forked, graph or not, batched or plain) keeps that order.  Without
// `fork` (the serial diagnostic) the same launches run on `st` alone.  The
// shadow queue of the last iteration is left for trace_pending_shadow.
hipError_t launch_iterations(const ViewArgs &va, const StreamQueues &q, const LoopShape &ls, uint32_t n, Timer *tm,
                             hipStream_t st, const Fork *fork) {
  for (uint32_t j = 0; j < n; ++j) {
    hipError_t e;
    const hipStream_t st_any = fork ? fork->st2 : st;
    if (fork) {
      if ((e = hipEventRecord(fork->ev_fork, st)) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(fork->st2, fork->ev_fork, 0)) != hipSuccess) return e;
    }
    if (tm && (e = tm->begin(kTimedShadow, st_any)) != hipSuccess) return e;
    if ((e = launch_shadow(va, q, ls, st_any)) != hipSuccess) return e;
    if (tm && (e = tm->end(st_any)) != hipSuccess) return e;
    if (fork && (e = hipEventRecord(fork->ev_join, fork->st2)) != hipSuccess) return e;
    if (tm && (e = tm->begin(kTimedExtend, st)) != hipSuccess) return e;
    if ((e = launch_extend(va, q, ls, st)) != hipSuccess) return e;
    if (tm && (e = tm->end(st)) != hipSuccess) return e;
    if (ls.volumes) {  // the shading's volume events, counted as shading
      if (tm && (e = tm->begin(kTimedShade, st)) != hipSuccess) return e;
      if ((e = launch_volume(va, q, ls, st)) != hipSuccess) return e;
      if (tm && (e = tm->end(st)) != hipSuccess) return e;
    }
    if (tm && (e = tm->begin(kTimedShade, st)) != hipSuccess) return e;
    if ((e = launch_shade(va, q, ls, st)) != hipSuccess) return e;
    if (tm && (e = tm->end(st)) != hipSuccess) return e;
    // The shadow queue's contributions follow the shading of the next bounce
    // into the samples' radiance; regeneration then reuses their slots.
    if (fork && (e = hipStreamWaitEvent(st, fork->ev_join, 0)) != hipSuccess) return e;
    if ((e = launch_nee(va, q, ls, st)) != hipSuccess) return e;
    if ((e = launch_regen(va, q, 1, ls, st)) != hipSuccess) return e;
  }
  return hipSuccess;
}

// The any-hit pass over the shadow queue of the last iteration that ran.
hipError_t trace_pending_shadow(const ViewArgs &va, const StreamQueues &q, const LoopShape &ls, Timer *tm,
                                hipStream_t st) {
  hipError_t e;
  if (tm && (e = tm->begin(kTimedShadow, st)) != hipSuccess) return e;
  if ((e = launch_shadow(va, q, ls, st)) != hipSuccess) return e;
  if ((e = launch_nee(va, q, ls, st)) != hipSuccess) return e;
  return tm ? tm->end(st) : hipSuccess;
}

// Iterations of the adjoint replay: closest hit, replayed shading with inline
// NEE visibility and gradients, regeneration that keeps the stored radiance.
hipError_t launch_replay_iterations(const ViewArgs &va, const StreamQueues &q, const AdjointArgs &adj,
                                    const LoopShape &ls, uint32_t n, hipStream_t st) {
  for (uint32_t j = 0; j < n; ++j) {
    hipError_t e;
    if ((e = launch_extend(va, q, ls, st)) != hipSuccess) return e;
    if ((e = launch_shade_adjoint(va, q, adj, ls, st)) != hipSuccess) return e;
    if ((e = launch_regen(va, q, 1, ls, st, false)) != hipSuccess) return e;
  }
  return hipSuccess;
}

using Lane = phx_context::Lane;

// Captures kIterBatch iterations of the bounce loop on lane `ln` as a graph,
// unless the graph it holds was captured with the same launch arguments.
int ensure_graph(phx_context *c, Lane &ln, const ViewArgs &va, const StreamQueues &q, const LoopShape &ls,
                 const Fork *fork) {
  std::vector<unsigned char> key(sizeof va + sizeof q + sizeof ls + 1);
  std::memcpy(key.data(), &va, sizeof va);
  std::memcpy(key.data() + sizeof va, &q, sizeof q);
  std::memcpy(key.data() + sizeof va + sizeof q, &ls, sizeof ls);
  key.back() = fork ? 1 : 0;
  if (ln.graph && key == ln.graph_key) return PHX_OK;
  if (ln.graph) {
    (void)hipGraphExecDestroy(ln.graph);
    ln.graph = nullptr;
  }
  HIP_TRY(c, hipStreamBeginCapture(ln.st, hipStreamCaptureModeThreadLocal));
  const hipError_t le = launch_iterations(va, q, ls, kIterBatch, nullptr, ln.st, fork);
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(ln.st, &g);
  if (le != hipSuccess || ce != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return hip_fail(c, le != hipSuccess ? le : ce, "capturing the bounce loop");
  }
  const hipError_t ie = hipGraphInstantiate(&ln.graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess) {
    ln.graph = nullptr;
    return hip_fail(c, ie, "hipGraphInstantiate");
  }
  ln.graph_key = std::move(key);
  return PHX_OK;
}

// Creates lane `l`'s streams (lane 1 only: lane 0 uses the context's) and
// events on first use.
int ensure_lane(phx_context *c, int l) {
  Lane &ln = c->lanes[l];
  if (!ln.st) HIP_TRY(c, hipStreamCreate(&ln.st));
  if (!ln.st2) HIP_TRY(c, hipStreamCreate(&ln.st2));
  for (hipEvent_t *e : {&ln.ev_fork, &ln.ev_join, &ln.ev_read, &ln.ev_reduced})
    if (!*e) HIP_TRY(c, hipEventCreateWithFlags(e, hipEventDisableTiming));
  return PHX_OK;
}

// Sizes lane `ln` for `cap` path slots, `samples` per-sample radiance records
// and item tables of `pad` entries (none for pad 1).
int reserve_lane(phx_context *c, Lane &ln, uint32_t cap, uint32_t samples, bool volumes, size_t pad) {
  if (cap > ln.q_capacity) {
    for (int b = 0; b < 2; ++b) {
      HIP_TRY(c, ln.q_o[b].reserve(size_t(cap) * 16));
      HIP_TRY(c, ln.q_d[b].reserve(size_t(cap) * 16));
      HIP_TRY(c, ln.q_tp[b].reserve(size_t(cap) * 16));
      HIP_TRY(c, ln.q_st[b].reserve(size_t(cap) * 16));
      HIP_TRY(c, ln.sh_o[b].reserve(size_t(cap) * 16));
      HIP_TRY(c, ln.sh_d[b].reserve(size_t(cap) * 16));
      HIP_TRY(c, ln.sh_c[b].reserve(size_t(cap) * 16));
    }
    HIP_TRY(c, ln.q_hit.reserve(size_t(cap) * 16));
    ln.q_capacity = cap;
  }
  HIP_TRY(c, ln.sample_L.reserve(size_t(samples) * 16));
  HIP_TRY(c, ln.counters.reserve(kCounterWords * 4));
  HIP_TRY(c, ln.stats_buf.reserve(kStatWords * 8));
  if (volumes) {
    HIP_TRY(c, ln.vol_ev.reserve(size_t(cap) * 16));
    HIP_TRY(c, ln.vol_rng.reserve(size_t(cap) * 4));
  }
  if (pad > 1) {
    HIP_TRY(c, ln.bt_items.reserve(pad * sizeof(RunItem)));
    HIP_TRY(c, ln.bt_first.reserve(pad * sizeof(uint32_t)));
    if (pad > ln.h_count) {  // no copy from the old staging is pending: the lane's stream is idle
      if (ln.h_items) (void)hipHostFree(ln.h_items);
      if (ln.h_first) (void)hipHostFree(ln.h_first);
      ln.h_items = nullptr;
      ln.h_first = nullptr;
      ln.h_count = 0;
      HIP_TRY(c, hipHostMalloc(reinterpret_cast<void **>(&ln.h_items), pad * sizeof(RunItem), hipHostMallocDefault));
      HIP_TRY(c, hipHostMalloc(reinterpret_cast<void **>(&ln.h_first), pad * sizeof(uint32_t), hipHostMallocDefault));
      ln.h_count = pad;
    }
  }
  return PHX_OK;
}

// The queue pointers of lane `ln`; its two traversal passes spill at
// spill[0..) and spill[trace_threads * (kMaxStack - kShortStack)..).
StreamQueues lane_queues(const Lane &ln, uint32_t *aov, int *spill, size_t trace_threads, bool volumes,
                         uint32_t *pix_active) {
  StreamQueues q{};
  for (int b = 0; b < 2; ++b) {
    q.o[b] = ln.q_o[b].as<float4>();
    q.d[b] = ln.q_d[b].as<float4>();
    q.tp[b] = ln.q_tp[b].as<float4>();
    q.st[b] = ln.q_st[b].as<uint4>();
    q.sh_o[b] = ln.sh_o[b].as<float4>();
    q.sh_d[b] = ln.sh_d[b].as<float4>();
    q.sh_c[b] = ln.sh_c[b].as<float4>();
  }
  q.hit = ln.q_hit.as<float4>();
  q.sample_L = ln.sample_L.as<float4>();
  q.aov_s = aov;
  q.counters = ln.counters.as<uint32_t>();
  q.stats = ln.stats_buf.as<unsigned long long>();
  q.spill = spill;
  q.spill_shadow = spill + trace_threads * (kMaxStack - kShortStack);
  q.vol_ev = volumes ? ln.vol_ev.as<float4>() : nullptr;
  q.vol_rng = volumes ? ln.vol_rng.as<uint32_t>() : nullptr;
  q.pix_active = pix_active;
  return q;
}

// The launch arguments of `run` on lane `ln`: scene and item in kernel
// arguments for a run of one item, else the lane's item tables, uploaded on
// its stream after the kernels of its previous run, which read them.  first[]
// is padded to `pad` (a power of two) with ~0u so the per-lane binary search
// has a fixed length.  The staging is rewritten only once the lane's previous
// run has ended on the host, which follows its upload on the stream.
int view_args(phx_context *c, Lane &ln, const Run &run, const std::vector<const phx_context *> &scenes, size_t pad,
              ViewArgs &va) {
  std::memset(&va, 0, sizeof va);
  if (run.items.size() == 1) {
    va.kind = 0;
    va.one.scene = scenes[run.items[0].r.scene]->dev;
    va.one.item = run.items[0].r;
    return PHX_OK;
  }
  for (size_t k = 0; k < pad; ++k) {
    ln.h_items[k] = k < run.items.size() ? run.items[k].r : run.items.back().r;
    ln.h_first[k] = k < run.items.size() ? run.items[k].r.first : 0xffffffffu;
  }
  HIP_TRY(c, hipMemcpyAsync(ln.bt_items.p, ln.h_items, pad * sizeof(RunItem), hipMemcpyHostToDevice, ln.st));
  HIP_TRY(c, hipMemcpyAsync(ln.bt_first.p, ln.h_first, pad * sizeof(uint32_t), hipMemcpyHostToDevice, ln.st));
  if (scenes.size() == 1) {
    va.kind = 1;
    va.shared.scene = scenes[0]->dev;
    va.shared.items = ln.bt_items.as<RunItem>();
    va.shared.first = ln.bt_first.as<uint32_t>();
    va.shared.count = uint32_t(pad);
  } else {
    va.kind = 2;
    va.many.scenes = c->bt_scenes.as<DevScene>();
    va.many.items = ln.bt_items.as<RunItem>();
    va.many.first = ln.bt_first.as<uint32_t>();
    va.many.count = uint32_t(pad);
  }
  return PHX_OK;
}

// Streams every view's samples through the shared queues and resolves each
// view into `fb` at its pix_base (AOV sample indices in aov[pix_base..]).
// With `adj`, each run is followed by its adjoint replay and nothing is
// resolved: the result is the gradients.
int stream_views(phx_context *c, const std::vector<ViewJob> &views, const std::vector<const phx_context *> &scenes,
                 const FrameBuffers &fb, uint32_t *aov, uint32_t total_pixels, bool use_graph,
                 const AdjointArgs *adj = nullptr) {
  // adaptive sampling: one view (phx_render), the per-pixel state in fb
  const bool adaptive = views.size() == 1 && views[0].adaptive && !adj && fb.nsamp;
  const std::vector<Run> runs = adaptive ? plan_adaptive_runs(views[0]) : plan_runs(views);
  uint32_t max_total = 0;
  size_t max_items = 1;
  for (const Run &r : runs) {
    max_total = std::max(max_total, r.total);
    max_items = std::max(max_items, r.items.size());
  }
  const uint32_t want = c->tuning.paths_in_flight ? c->tuning.paths_in_flight : kDefaultPaths;
  const uint32_t cap = std::max<uint32_t>(1, std::min<uint32_t>(want, max_total));
  // The launch shape of each run: the refill threshold from the depth of its
  // scenes' trees, the shading tier and the volume pass from what its scenes
  // and views hold (a batch's small views do not slow its solo views).
  const LoopShape ls0 = loop_shape(c, cap);
  auto run_shape = [&](const Run &run) {
    LoopShape ls = ls0;
    ls.tier = kTierLean;
    uint32_t depth = 0;
    for (const HostItem &h : run.items) {
      const phx_context *sc = scenes[h.r.scene];
      depth = std::max(depth, sc->bvh_depth);
      ls.volumes |= sc->dev.vol_count > 0 ? 1u : 0u;
      if (sc->dev.vol_count > 0 || sc->dev.tex_count > 0 || sc->dev.light_count > 0)
        ls.tier = std::max<uint32_t>(ls.tier, kTierStandard);
      if (sc->dev.has_sources || h.r.sp.bg_image) ls.tier = kTierFull;
    }
    if (!c->tuning.refill_lanes) ls.refill = depth >= kDeepTree ? kRefillDeep : kRefillShallow;
    ls.tiles = 1;
    if (const char *e = std::getenv("PHX_TILE_ORDER")) ls.tiles = std::atoi(e) != 0;
    return ls;
  };
  bool any_volumes = false;
  for (const phx_context *sc : scenes) any_volumes |= sc->dev.vol_count > 0;

  // PHX_TRACE_COUNTS=N: read the queue length every N iterations and print it.
  const char *tc = std::getenv("PHX_TRACE_COUNTS");
  const bool trace_counts = tc && std::atoi(tc) > 0;
  const uint32_t batch = trace_counts ? uint32_t(std::atoi(tc)) : kIterBatch;
  Timer tm{c, c->tuning.time_kernels != 0};
  // PHX_SERIAL_BOUNCE=1 (diagnostic): every launch of the bounce loop on one
  // stream, the any-hit pass not beside the closest-hit pass.
  const char *sb = std::getenv("PHX_SERIAL_BOUNCE");
  const bool serial = sb && std::atoi(sb) > 0;
  // Two lanes for a batch of several runs (PHX_OVERLAP=0, diagnostic: one).
  const char *ov = std::getenv("PHX_OVERLAP");
  const int nlanes = !adj && !adaptive && !trace_counts && !tm.on && runs.size() > 1 && !(ov && std::atoi(ov) == 0)
                         ? 2 : 1;

  size_t pad = 1;
  while (pad < max_items) pad <<= 1;
  // Two spill areas per lane, one per traversal pass: k_shadow runs beside
  // k_extend on the lane's second stream, and a thread of each may spill at
  // the same index.
  const size_t trace_threads = size_t(ls0.trace_grid) * ls0.trace_block;
  int r = ensure_trace_scratch(c, 2 * nlanes * trace_threads);
  if (r) return r;
  if (!c->pinned) HIP_TRY(c, hipHostMalloc(reinterpret_cast<void **>(&c->pinned), 64, hipHostMallocDefault));
  StreamQueues qs[2]{};
  Fork forks[2]{};
  for (int l = 0; l < nlanes; ++l) {
    if ((r = ensure_lane(c, l))) return r;
    Lane &ln = c->lanes[l];
    if ((r = reserve_lane(c, ln, cap, max_total, any_volumes, max_items > 1 ? pad : 1))) return r;
    qs[l] = lane_queues(ln, aov, c->spill.as<int>() + 2 * l * trace_threads * (kMaxStack - kShortStack),
                        trace_threads, any_volumes, adaptive ? c->f_active.as<uint32_t>() : nullptr);
    forks[l] = Fork{ln.st2, ln.ev_fork, ln.ev_join};
  }
  if (max_items > 1 && scenes.size() > 1) {
    std::vector<DevScene> table(scenes.size());
    for (size_t i = 0; i < scenes.size(); ++i) table[i] = scenes[i]->dev;
    HIP_TRY(c, c->bt_scenes.reserve(table.size() * sizeof(DevScene)));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(c->bt_scenes.p, table.data(), table.size() * sizeof(DevScene), hipMemcpyHostToDevice));
  }

  hipStream_t st = c->stream;
  const StreamQueues &q = qs[0];
  HIP_TRY(c, launch_aov_init(total_pixels, aov, st));
  if (adaptive) HIP_TRY(c, launch_adaptive_init(total_pixels, c->f_active.as<uint32_t>(), st));
  for (int l = 0; l < nlanes; ++l) {
    HIP_TRY(c, launch_clear_u32(reinterpret_cast<uint32_t *>(qs[l].stats), kStatWords * 2, st));
    if (l > 0) {  // lane 1 starts after the set-up above and the caller's work on `st`
      HIP_TRY(c, hipEventRecord(c->lanes[0].ev_reduced, st));
      HIP_TRY(c, hipStreamWaitEvent(c->lanes[l].st, c->lanes[0].ev_reduced, 0));
    }
  }
  auto reduce_run = [&](const Run &run, const StreamQueues &ql, hipStream_t s) -> int {
    for (const HostItem &h : run.items) {
      if (adaptive)
        HIP_TRY(c, launch_reduce_adaptive(h.r.sp.npix, h.nspp, ql.sample_L + h.r.first, fb.accum + h.r.pix_base,
                                          c->f_accum_sq.as<float4>() + h.r.pix_base, fb.nsamp + h.r.pix_base,
                                          c->f_active.as<uint32_t>() + h.r.pix_base, h.r.s0, s));
      else
        HIP_TRY(c, launch_reduce(h.r.sp.npix, h.nspp, ql.sample_L + h.r.first, fb.accum + h.r.pix_base, h.r.s0 == 0,
                                 s));
    }
    return PHX_OK;
  };

  if (nlanes == 1) {
    Lane &ln = c->lanes[0];
    const Fork *fork = serial ? nullptr : &forks[0];
    for (const Run &run : runs) {
      if (run.items.empty()) continue;
      const LoopShape ls = run_shape(run);
      ViewArgs va;
      if ((r = view_args(c, ln, run, scenes, pad, va))) return r;
      const bool graph = use_graph && !adj && va.kind != 0 && !trace_counts && !tm.on;
      if (graph && (r = ensure_graph(c, ln, va, q, ls, fork))) return r;
      for (int pass = 0; pass < (adj ? 2 : 1); ++pass) {
        const bool replay = pass == 1;
        HIP_TRY(c, launch_run_start(q, run.total, std::min(cap, run.total), st));
        HIP_TRY(c, launch_regen(va, q, 0, ls, st, !replay));
        uint32_t it = 0;
        LoopShape lsb = ls;   // the launch shape of the next batch of iterations
        uint32_t nb = batch;  // iterations of the next batch (launched one by one)
        while (true) {
          if (replay) {
            HIP_TRY(c, launch_replay_iterations(va, q, *adj, ls, batch, st));
            it += batch;
          } else if (graph) {
            HIP_TRY(c, hipGraphLaunch(ln.graph, st));
            it += kIterBatch;
          } else {
            HIP_TRY(c, launch_iterations(va, q, lsb, nb, tm.on ? &tm : nullptr, st, fork));
            it += nb;
          }
          // the iteration counter, the samples started, then that iteration's queue
          HIP_TRY(c, hipMemcpyAsync(&c->pinned[0], q.counters + kIter, 4, hipMemcpyDeviceToHost, st));
          HIP_TRY(c, hipMemcpyAsync(&c->pinned[1], q.counters + kNextSample, 4, hipMemcpyDeviceToHost, st));
          HIP_TRY(c, hipMemcpyAsync(&c->pinned[2], q.counters + ray_count_word(it), 4, hipMemcpyDeviceToHost, st));
          HIP_TRY(c, hipStreamSynchronize(st));
          if (c->pinned[0] != it)
            return fail(c, PHX_ERR_OVERFLOW, "render: wavefront iteration counter out of step");
          if (trace_counts)
            std::fprintf(stderr, "[phx] %s iteration %u: queue %u, started %u of %u samples\n",
                         replay ? "replay" : "render", it, c->pinned[2], c->pinned[1], run.total);
          if (c->pinned[2] == 0 && c->pinned[1] >= run.total) break;
          // Once every sample has started, no path is added: the queues of the
          // iterations to come hold at most the one just read, and the shading
          // grid (one thread per queue entry) shrinks to it.  A grid over all
          // path slots costs about 28 us to launch even over an empty queue.
          // The drain that follows takes a few iterations, each bound by its
          // slowest rays; batches of half the size launch fewer iterations over
          // an empty queue at its end (each costs ~60 us) for one more read.
          if (c->pinned[1] >= run.total) {
            lsb.shade_grid = std::max<uint32_t>(1, std::min(lsb.shade_grid, (c->pinned[2] + 255u) / 256u));
            nb = std::max<uint32_t>(1, batch / 2);
          }
          if (it >= kIterationGuard)
            return fail(c, PHX_ERR_OVERFLOW, "render: wavefront did not drain within its iteration guard");
        }
        if (!replay) HIP_TRY(c, trace_pending_shadow(va, q, ls, tm.on ? &tm : nullptr, st));
      }
      if (!adj && (r = reduce_run(run, q, st))) return r;
    }
  } else {
    // Overlapped runs.  Once every sample of a run has started, its drain
    // (a few iterations bound by their slowest rays) leaves most of the GPU
    // idle, and the next run starts on the other lane beside it.  Run k
    // streams on lane k % 2: it starts once run k - 2 has been reduced and run
    // k - 1 drains.  Reduces are issued in run order, each after the previous
    // run's on the other lane, so a view split over several runs adds its
    // samples in the order one lane would; every view's result is the same
    // as with one lane.  The host polls each lane's counter readback.
    struct LaneRun {
      long run = -1;  // the run the lane streams, -1 idle
      ViewArgs va;
      LoopShape ls, lsb;
      bool graph = false, draining = false, finished = false;
      uint32_t it = 0, nb = 0;
    };
    LaneRun lr[2];
    auto launch_batch = [&](int l) -> int {
      LaneRun &s = lr[l];
      Lane &ln = c->lanes[l];
      if (s.graph) {
        HIP_TRY(c, hipGraphLaunch(ln.graph, ln.st));
        s.it += kIterBatch;
      } else {
        HIP_TRY(c, launch_iterations(s.va, qs[l], s.lsb, s.nb, nullptr, ln.st, serial ? nullptr : &forks[l]));
        s.it += s.nb;
      }
      uint32_t *pw = c->pinned + 4 * l;
      HIP_TRY(c, hipMemcpyAsync(&pw[0], qs[l].counters + kIter, 4, hipMemcpyDeviceToHost, ln.st));
      HIP_TRY(c, hipMemcpyAsync(&pw[1], qs[l].counters + kNextSample, 4, hipMemcpyDeviceToHost, ln.st));
      HIP_TRY(c, hipMemcpyAsync(&pw[2], qs[l].counters + ray_count_word(s.it), 4, hipMemcpyDeviceToHost, ln.st));
      HIP_TRY(c, hipEventRecord(ln.ev_read, ln.st));
      return PHX_OK;
    };
    size_t next = 0, reduced = 0;  // the next run to start; runs reduced
    while (reduced < runs.size()) {
      bool progress = false;
      const int ls_next = int(next % 2);
      if (next < runs.size() && lr[ls_next].run < 0) {
        const LaneRun &o = lr[1 - ls_next];
        if (o.run < 0 || o.draining || o.finished) {
          LaneRun &s = lr[ls_next];
          Lane &ln = c->lanes[ls_next];
          const Run &run = runs[next];
          s = LaneRun{};
          s.run = long(next++);
          progress = true;
          if (!run.items.empty()) {
            s.ls = s.lsb = run_shape(run);
            s.nb = batch;
            if ((r = view_args(c, ln, run, scenes, pad, s.va))) return r;
            s.graph = use_graph && s.va.kind != 0;
            if (s.graph && (r = ensure_graph(c, ln, s.va, qs[ls_next], s.ls, serial ? nullptr : &forks[ls_next])))
              return r;
            HIP_TRY(c, launch_run_start(qs[ls_next], run.total, std::min(cap, run.total), ln.st));
            HIP_TRY(c, launch_regen(s.va, qs[ls_next], 0, s.ls, ln.st, true));
            if ((r = launch_batch(ls_next))) return r;
          } else {
            s.finished = true;
          }
        }
      }
      for (int l = 0; l < 2; ++l) {
        LaneRun &s = lr[l];
        if (s.run < 0 || s.finished) continue;
        Lane &ln = c->lanes[l];
        const hipError_t qe = hipEventQuery(ln.ev_read);
        if (qe == hipErrorNotReady) continue;
        HIP_TRY(c, qe);
        progress = true;
        const uint32_t *pw = c->pinned + 4 * l;
        const Run &run = runs[size_t(s.run)];
        if (pw[0] != s.it) return fail(c, PHX_ERR_OVERFLOW, "render: wavefront iteration counter out of step");
        if (pw[2] == 0 && pw[1] >= run.total) {
          HIP_TRY(c, trace_pending_shadow(s.va, qs[l], s.ls, nullptr, ln.st));
          s.finished = true;
          continue;
        }
        if (pw[1] >= run.total) {  // the drain, as with one lane
          s.draining = true;
          s.graph = false;
          s.lsb.shade_grid = std::max<uint32_t>(1, std::min(s.lsb.shade_grid, (pw[2] + 255u) / 256u));
          s.nb = std::max<uint32_t>(1, batch / 2);
        }
        if (s.it >= kIterationGuard)
          return fail(c, PHX_ERR_OVERFLOW, "render: wavefront did not drain within its iteration guard");
        if ((r = launch_batch(l))) return r;
      }
      while (reduced < runs.size()) {
        const int l = int(reduced % 2);
        LaneRun &s = lr[l];
        if (s.run != long(reduced) || !s.finished) break;
        Lane &ln = c->lanes[l];
        if (reduced > 0) HIP_TRY(c, hipStreamWaitEvent(ln.st, c->lanes[1 - l].ev_reduced, 0));
        if ((r = reduce_run(runs[reduced], qs[l], ln.st))) return r;
        HIP_TRY(c, hipEventRecord(ln.ev_reduced, ln.st));
        s = LaneRun{};
        ++reduced;
        progress = true;
      }
      if (!progress) std::this_thread::yield();
    }
    // the resolve and the statistics follow the last run's reduces
    const int last = int((runs.size() - 1) % 2);
    if (last != 0) HIP_TRY(c, hipStreamWaitEvent(st, c->lanes[last].ev_reduced, 0));
  }
  for (const ViewJob &j : views) {
    if (adj) break;
    StreamQueues qv = q;
    qv.aov_s = aov + j.pix_base;
    HIP_TRY(c, launch_clear_u32(c->fetch.as<uint32_t>(), 1, st));
    // the hit records of the path slots are free now: one per pixel, if they fit
    float4 *hits = j.scene->dev.vol_count == 0 && j.sp.npix <= cap ? q.hit : nullptr;
    HIP_TRY(c, launch_resolve(j.scene->dev, j.sp, qv, fb.at(j.pix_base), j.spp, c->fetch.as<uint32_t>(),
                              c->trace_grid, st, hits));
  }
  unsigned long long hs[2][kStatWords];
  for (int l = 0; l < nlanes; ++l)
    HIP_TRY(c, hipMemcpyAsync(hs[l], qs[l].stats, sizeof hs[l], hipMemcpyDeviceToHost, st));
  HIP_TRY(c, hipStreamSynchronize(st));
  c->stats.closest_rays = c->stats.shadow_rays = c->stats.shade_events = 0;
  for (int l = 0; l < nlanes; ++l) {
    c->stats.closest_rays += hs[l][kStatClosest];
    c->stats.shadow_rays += hs[l][kStatShadow];
    c->stats.shade_events += hs[l][kStatShade];
  }
  uint64_t samples = 0;
  for (const ViewJob &j : views) samples += uint64_t(j.sp.npix) * j.spp;
  c->stats.samples = samples;
  c->stats.pixels = total_pixels;
  c->stats.extend_launches = c->stats.shade_launches = c->stats.shadow_launches = 0;
  c->stats.extend_ms = c->stats.shade_ms = c->stats.shadow_ms = 0.0;
  for (size_t k = 0; k < tm.kinds.size(); ++k) {
    float ms = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->events[2 * k], c->events[2 * k + 1]));
    if (tm.kinds[k] == kTimedExtend) {
      c->stats.extend_ms += ms;
      c->stats.extend_launches++;
    } else if (tm.kinds[k] == kTimedShade) {
      c->stats.shade_ms += ms;
      c->stats.shade_launches++;
    } else {
      c->stats.shadow_ms += ms;
      c->stats.shadow_launches++;
    }
  }
  return PHX_OK;
}

int reserve_frame(phx_context *c, DBuf &accum, DBuf &color, DBuf &alpha, DBuf &depth, DBuf &normal, DBuf &albedo,
                  size_t npix) {
  HIP_TRY(c, accum.reserve(npix * 16));
  HIP_TRY(c, color.reserve(npix * 12));
  HIP_TRY(c, alpha.reserve(npix * 4));
  HIP_TRY(c, depth.reserve(npix * 4));
  HIP_TRY(c, normal.reserve(npix * 12));
  HIP_TRY(c, albedo.reserve(npix * 12));
  return PHX_OK;
}

}  // namespace

extern "C" int phx_render(phx_context *c, const phx_camera *cam, const phx_render_params *p) {
  if (!c || !cam || !p) return PHX_ERR_INVALID;
  if (!c->have_scene) return fail(c, PHX_ERR_NO_SCENE, "render before set_scene");
  int r = check_params(c, p, "render");
  if (r) return r;
  HIP_TRY(c, hipSetDevice(c->device));
  const auto t0 = std::chrono::steady_clock::now();
  const uint32_t npix = p->width * p->height;
  if (c->f_w != p->width || c->f_h != p->height || !c->f_accum.p) {
    if ((r = reserve_frame(c, c->f_accum, c->f_color, c->f_alpha, c->f_depth, c->f_normal, c->f_albedo, npix)))
      return r;
    c->f_w = p->width;
    c->f_h = p->height;
  }
  HIP_TRY(c, c->aov_s.reserve(size_t(npix) * 4));
  c->have_result = false;
  c->f_counted = false;
  const std::vector<ViewJob> views{make_job(c, 0, *cam, *p, 0)};
  FrameBuffers fb{c->f_accum.as<float4>(), c->f_color.as<float>(), c->f_alpha.as<float>(),
                  c->f_depth.as<float>(), c->f_normal.as<float>(), c->f_albedo.as<float>()};
  if (p->adaptive_sampling) {
    HIP_TRY(c, c->f_accum_sq.reserve(size_t(npix) * 16));
    HIP_TRY(c, c->f_nsamp.reserve(size_t(npix) * 4));
    HIP_TRY(c, c->f_active.reserve(size_t(npix) * 4));
    fb.nsamp = c->f_nsamp.as<uint32_t>();
  }
  if ((r = stream_views(c, views, {c}, fb, c->aov_s.as<uint32_t>(), npix, false))) return r;
  c->stats.last_render_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  c->have_result = true;
  c->f_counted = p->adaptive_sampling != 0;
  return PHX_OK;
}

extern "C" int phx_render_batch(phx_context *c, const phx_view *views, uint32_t count) {
  if (!c || (count && !views)) return PHX_ERR_INVALID;
  if (count == 0) return fail(c, PHX_ERR_INVALID, "render_batch: no views");
  // a view adds at most one item to a run, and items are numbered in 15 bits
  if (count > kMaxRunItems) return fail(c, PHX_ERR_INVALID, "render_batch: more than 32768 views");
  int r;
  std::vector<const phx_context *> scenes;
  std::vector<ViewJob> jobs;
  uint64_t pix = 0;
  for (uint32_t v = 0; v < count; ++v) {
    const phx_view &w = views[v];
    if (!w.scene) return fail(c, PHX_ERR_INVALID, "render_batch: view without a scene");
    if (!w.scene->have_scene) return fail(c, PHX_ERR_NO_SCENE, "render_batch: a view's scene is not set");
    if (w.scene->device != c->device) return fail(c, PHX_ERR_INVALID, "render_batch: scene on another device");
    if ((r = check_params(c, &w.params, "render_batch"))) return r;
    if (w.params.adaptive_sampling)
      return fail(c, PHX_ERR_UNSUPPORTED, "render_batch: adaptive sampling is offered by phx_render only");
    uint32_t si = 0;
    while (si < scenes.size() && scenes[si] != w.scene) ++si;
    if (si == scenes.size()) scenes.push_back(w.scene);
    jobs.push_back(make_job(w.scene, si, w.camera, w.params, uint32_t(pix)));
    pix += uint64_t(w.params.width) * w.params.height;
    if (pix > (uint64_t(1) << 28)) return fail(c, PHX_ERR_INVALID, "render_batch: more than 2^28 pixels in all");
  }
  HIP_TRY(c, hipSetDevice(c->device));
  const auto t0 = std::chrono::steady_clock::now();
  c->bt_have_result = false;
  if ((r = reserve_frame(c, c->bt_accum, c->bt_color, c->bt_alpha, c->bt_depth, c->bt_normal, c->bt_albedo, pix)))
    return r;
  HIP_TRY(c, c->bt_aov.reserve(pix * 4));
  const FrameBuffers fb{c->bt_accum.as<float4>(), c->bt_color.as<float>(), c->bt_alpha.as<float>(),
                        c->bt_depth.as<float>(), c->bt_normal.as<float>(), c->bt_albedo.as<float>()};
  if ((r = stream_views(c, jobs, scenes, fb, c->bt_aov.as<uint32_t>(), uint32_t(pix), c->tuning.use_graph != 0)))
    return r;
  c->bt_views.clear();
  for (uint32_t v = 0; v < count; ++v)
    c->bt_views.push_back({views[v].params.width, views[v].params.height, jobs[v].pix_base});
  c->stats.last_render_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  c->bt_have_result = true;
  return PHX_OK;
}

extern "C" int phx_read_batch_result(phx_context *c, uint32_t view, float *color, float *alpha, float *depth,
                                     float *normal, float *albedo) {
  if (!c) return PHX_ERR_INVALID;
  if (!c->bt_have_result) return fail(c, PHX_ERR_NO_FRAME, "read_batch_result before render_batch");
  if (view >= c->bt_views.size()) return fail(c, PHX_ERR_INVALID, "read_batch_result: no such view");
  HIP_TRY(c, hipSetDevice(c->device));
  const auto &v = c->bt_views[view];
  const size_t n = size_t(v.width) * v.height, b = v.pix_base;
  if (color) HIP_TRY(c, hipMemcpy(color, c->bt_color.as<float>() + b * 3, n * 12, hipMemcpyDefault));
  if (alpha) HIP_TRY(c, hipMemcpy(alpha, c->bt_alpha.as<float>() + b, n * 4, hipMemcpyDefault));
  if (depth) HIP_TRY(c, hipMemcpy(depth, c->bt_depth.as<float>() + b, n * 4, hipMemcpyDefault));
  if (normal) HIP_TRY(c, hipMemcpy(normal, c->bt_normal.as<float>() + b * 3, n * 12, hipMemcpyDefault));
  if (albedo) HIP_TRY(c, hipMemcpy(albedo, c->bt_albedo.as<float>() + b * 3, n * 12, hipMemcpyDefault));
  return PHX_OK;
}

extern "C" int phx_autotune(phx_context *c, const phx_view *views, uint32_t count, uint32_t repeats, phx_tuning *best,
                            double *best_ms) {
  if (!c) return PHX_ERR_INVALID;
  repeats = std::max<uint32_t>(repeats, 1);
  const phx_tuning start = c->tuning;
  auto time_it = [&](const phx_tuning &t, double &ms) -> int {
    int r = phx_set_tuning(c, &t);
    if (r) return r;
    if ((r = phx_render_batch(c, views, count))) return r;  // warm-up: allocations, graph capture
    ms = 1e300;
    for (uint32_t i = 0; i < repeats; ++i) {
      if ((r = phx_render_batch(c, views, count))) return r;
      ms = std::min(ms, c->stats.last_render_ms);
    }
    return PHX_OK;
  };
  phx_tuning cur = start;
  cur.time_kernels = 0;
  if (!cur.trace_block) cur.trace_block = 256;
  if (!cur.trace_groups_per_cu) cur.trace_groups_per_cu = 8;
  if (!cur.queue_chunk) cur.queue_chunk = kDefaultChunk;
  if (!cur.refill_lanes) cur.refill_lanes = kDefaultRefill;
  if (!cur.paths_in_flight) cur.paths_in_flight = kDefaultPaths;
  double cur_ms;
  int r = time_it(cur, cur_ms);
  if (r) {
    (void)phx_set_tuning(c, &start);
    return r;
  }
  // one pass of coordinate descent over the launch-shape knobs
  const std::vector<std::pair<uint32_t phx_tuning::*, std::vector<uint32_t>>> knobs = {
      {&phx_tuning::trace_block, {128, 256}},
      {&phx_tuning::trace_groups_per_cu, {4, 8, 16, 32}},
      {&phx_tuning::queue_chunk, {64, 128, 256, 512, 1024}},
      {&phx_tuning::refill_lanes, {8, 16, 32, 64}},
      {&phx_tuning::sort_by_material, {0, 4, 8, 15}},
      {&phx_tuning::paths_in_flight, {1u << 23, 1u << 24, 1u << 25, 1u << 26}},
  };
  for (const auto &k : knobs) {
    const uint32_t was = cur.*(k.first);
    for (uint32_t val : k.second) {
      if (val == was) continue;
      phx_tuning t = cur;
      t.*(k.first) = val;
      double ms;
      if ((r = time_it(t, ms))) {
        (void)phx_set_tuning(c, &start);
        return r;
      }
      if (ms < cur_ms) {
        cur = t;
        cur_ms = ms;
      }
    }
  }
  if ((r = phx_set_tuning(c, &cur))) return r;
  if (best) *best = cur;
  if (best_ms) *best_ms = cur_ms;
  return PHX_OK;
}

extern "C" int phx_adjoint(phx_context *c, const phx_camera *cam, const phx_render_params *p, const float *d_adjoint,
                           float *d_grad_materials, float *d_grad_texture) {
  if (!c || !cam || !p || !d_adjoint || !d_grad_materials) return PHX_ERR_INVALID;
  if (!c->have_scene) return fail(c, PHX_ERR_NO_SCENE, "adjoint before set_scene");
  int r = check_params(c, p, "adjoint");
  if (r) return r;
  if (p->adaptive_sampling)
    return fail(c, PHX_ERR_UNSUPPORTED, "adjoint: adaptive sampling is not differentiated (its sample counts are not)");
  HIP_TRY(c, hipSetDevice(c->device));
  const auto t0 = std::chrono::steady_clock::now();
  const uint32_t npix = p->width * p->height;
  const uint32_t nm = c->dev.material_count * kGradParamsAbi;
  const uint32_t nt = d_grad_texture ? c->tex_pixel_count * 3 : 0;
  HIP_TRY(c, c->adj_gm.reserve(size_t(nm) * 8));
  if (nt) HIP_TRY(c, c->adj_gt.reserve(size_t(nt) * 8));
  HIP_TRY(c, c->aov_s.reserve(size_t(npix) * 4));
  hipStream_t st = c->stream;
  HIP_TRY(c, launch_clear_f64(c->adj_gm.as<double>(), nm, st));
  if (nt) HIP_TRY(c, launch_clear_f64(c->adj_gt.as<double>(), nt, st));
  AdjointArgs a{};
  a.adjoint = d_adjoint;
  a.grad_mat = c->adj_gm.as<double>();
  a.grad_tex = nt ? c->adj_gt.as<double>() : nullptr;
  a.lds_mats = std::min<uint32_t>(c->dev.material_count, kAdjLdsMaterials);
  const std::vector<ViewJob> views{make_job(c, 0, *cam, *p, 0)};
  const FrameBuffers fb{};
  if ((r = stream_views(c, views, {c}, fb, c->aov_s.as<uint32_t>(), npix, false, &a))) return r;
  HIP_TRY(c, launch_grad_out(c->adj_gm.as<double>(), d_grad_materials, nm, st));
  if (nt) HIP_TRY(c, launch_grad_out(c->adj_gt.as<double>(), d_grad_texture, nt, st));
  HIP_TRY(c, hipStreamSynchronize(st));
  c->stats.last_render_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return PHX_OK;
}

extern "C" int phx_set_material_params(phx_context *c, const float *d_params, uint32_t count) {
  if (!c || (count && !d_params)) return PHX_ERR_INVALID;
  if (!c->have_scene) return fail(c, PHX_ERR_NO_SCENE, "set_material_params before set_scene");
  if (count > c->dev.material_count) return fail(c, PHX_ERR_INVALID, "set_material_params: more than the scene's materials");
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, launch_set_material_params(c->materials.as<phx_material>(), d_params, count, c->stream));
  HIP_TRY(c, launch_em_records(c->dev, c->em_recs.as<float4>(), c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PHX_OK;
}

extern "C" int phx_set_texture_pixels(phx_context *c, const float *d_pixels, uint32_t first, uint32_t count) {
  if (!c || (count && !d_pixels)) return PHX_ERR_INVALID;
  if (!c->have_scene) return fail(c, PHX_ERR_NO_SCENE, "set_texture_pixels before set_scene");
  if (uint64_t(first) + count > c->tex_pixel_count)
    return fail(c, PHX_ERR_INVALID, "set_texture_pixels: outside the texture atlas");
  if (count == 0) return PHX_OK;
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipMemcpyAsync(c->tex_pixels.as<float>() + size_t(first) * 3, d_pixels, size_t(count) * 12,
                            hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(c, launch_em_records(c->dev, c->em_recs.as<float4>(), c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PHX_OK;
}
