// Streaming wavefront integrator: state layout, the view policies the kernels
// are instantiated for, and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include "dev_queue.h"
#include "dev_shade.h"
#include "dev_trace.h"

namespace phx {

// Counter words of one streaming run, in rings of kRing entries indexed by
// ring(it) for wavefront iteration `it` (a global step of the wavefront, not
// of a path).  Entry it + 1 of the queue pairs counts the ray queue of
// iteration it + 1 (word ray_count_word(it + 1)) and the shadow queue that
// iteration it fills (word shadow_count_word(it + 1)): the two words form one
// 64-bit word, so the shade kernel claims its continuation and shadow slots
// with one atomic per work-group.  The persistent traversal kernels fetch
// work through the fetch sets fetch_set(kFetch*, it) (kFetchSetWords words:
// one chunk counter per stripe of the queue, each in a 128-byte line).  kNextSample is the next sample to start, kIter the
// current iteration, kTotal / kCapacity the run's sample count and path
// slots.  The kernels read the iteration from kIter (so a captured sequence of
// iterations can be replayed unchanged); k_run_start clears everything and
// k_commit clears the entries of it + 1 before they are used, so no host
// write or memset sits inside the bounce loop.  Each counter group has a
// 256-byte block of its own, so the atomics of one kernel do not queue behind
// another group's in the same cache line.
constexpr uint32_t kMaxIterations = 255;  // per path (RenderParams bound)
constexpr uint32_t kRing = 16;
constexpr uint32_t kQueuePairs = 0;  // 2 * kRing words
constexpr uint32_t kRegenTake = 64;
constexpr uint32_t kNextSample = 128;
constexpr uint32_t kIter = kNextSample + 1;
constexpr uint32_t kTotal = kIter + 1;
constexpr uint32_t kCapacity = kTotal + 1;
constexpr uint32_t kFetchExtend = 1024;                           // kRing fetch sets
constexpr uint32_t kFetchShadow = kFetchExtend + kRing * kFetchSetWords;  // kRing fetch sets
constexpr uint32_t kCounterWords = kFetchShadow + kRing * kFetchSetWords;
// Shading by material group: the hit record's top 4 bits carry the hit's
// group (material id mod the group count) and k_shade orders each work-group's
// rays by group (misses last) before shading them.
constexpr uint32_t kMaxGroups = 15;
constexpr uint32_t kMaxClasses = kMaxGroups + 1;
__host__ __device__ inline uint32_t hit_tri(uint32_t w) { return w == 0xffffffffu ? w : (w & 0x0fffffffu); }
__host__ __device__ inline uint32_t ring(uint32_t it) { return it & (kRing - 1); }
__host__ __device__ inline uint32_t ray_count_word(uint32_t it) { return kQueuePairs + 2u * ring(it); }
__host__ __device__ inline uint32_t fetch_set(uint32_t base, uint32_t it) { return base + ring(it) * kFetchSetWords; }
__host__ __device__ inline uint32_t shadow_count_word(uint32_t it) { return kQueuePairs + 2u * ring(it) + 1u; }

constexpr uint32_t kStatClosest = 0;
constexpr uint32_t kStatShadow = 1;
constexpr uint32_t kStatShade = 2;
constexpr uint32_t kStatWords = 8;

// One view's samples [s0, s0 + nspp) inside a streaming run.  A run streams
// the run-global sample ids [0, total); item k owns [first, first + npix *
// nspp) with local id l = s * npix + pixel and sample index sample_base + s.
struct RunItem {
  ShadeParams sp;        // the view's shading constants
  uint32_t scene;        // index into the scene table
  uint32_t first;        // run-global id of the item's first sample
  uint32_t sample_base;  // sample index of s = 0 (sample_offset + s0)
  uint32_t s0;           // first spp index of the item (AOV bookkeeping)
  uint32_t pix_base;     // first pixel of the view in the frame buffers
  uint32_t view;         // view index (diagnostics)
};

// A path's item index rides in the flags word of its queue entry (bits
// kItemShift and up) and in its shadow ray's origin w (bits 8 and up), so the
// kernels index the item table directly instead of searching it by sample id.
constexpr uint32_t kItemShift = 17;
constexpr uint32_t kMaxRunItems = 1u << (32 - kItemShift);

// View policies.  OneItem: a single view and scene passed as kernel
// arguments (uniform, scalar loads).  ItemTable: many views in device tables,
// looked up per lane from the sample id; kOneScene keeps the scene a kernel
// argument when every view renders the same scene.
struct OneItem {
  DevScene scene;
  RunItem item;
  PHX_DEV uint32_t find(uint32_t) const { return 0; }
  PHX_DEV const RunItem &item_at(uint32_t) const { return item; }
  PHX_DEV const DevScene &scene_of(const RunItem &) const { return scene; }
};
template <bool kOneScene>
struct ItemTable {
  DevScene scene;          // used when kOneScene
  const DevScene *scenes;  // used otherwise
  const RunItem *items;
  const uint32_t *first;   // items[k].first, ascending, first[0] = 0
  uint32_t count;
  PHX_DEV uint32_t find(uint32_t g) const {
    uint32_t lo = 0, hi = count;  // largest k with first[k] <= g
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (first[mid] <= g) lo = mid;
      else hi = mid;
    }
    return lo;
  }
  PHX_DEV const RunItem &item_at(uint32_t k) const { return items[k]; }
  PHX_DEV const DevScene &scene_of(const RunItem &r) const { return kOneScene ? scene : scenes[r.scene]; }
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
  // Two shadow queues by iteration parity: the shading of iteration it fills
  // queue (it + 1) & 1 while the any-hit pass of iteration it drains it & 1.
  float4 *sh_o[2];   // (origin, path_it bits); shadow rays start at t = 0
  float4 *sh_d[2];   // (direction, tmax)
  float4 *sh_c[2];   // (NEE contribution, sample id bits)
  float4 *sample_L;  // per sample of the run: (radiance, alpha)
  uint32_t *aov_s;   // per pixel: first sample index with a camera-ray event
  uint32_t *counters;
  unsigned long long *stats;
  int *spill;        // per-thread traversal stack spill
  int *spill_shadow; // the same for k_shadow, which runs beside k_extend
  // Adaptive sampling: per pixel of the frame (pix_base + pixel), nonzero
  // while the pixel takes samples; null when every sample is taken.
  const uint32_t *pix_active;
  // Per ray of the current queue, from k_volume (runs whose scenes hold
  // volumes): the volume event (scatter distance or -1, albedo) and the RNG
  // state after its draws.
  float4 *vol_ev;
  uint32_t *vol_rng;
};

struct FrameBuffers {
  float4 *accum;  // (radiance sum, alpha sum)
  float *color, *alpha, *depth, *normal, *albedo;
  uint32_t *nsamp = nullptr;  // samples taken per pixel (adaptive sampling); null: the view's spp
  FrameBuffers at(uint32_t pix) const {
    return {accum + pix, color + size_t(pix) * 3, alpha + pix, depth + pix, normal + size_t(pix) * 3,
            albedo + size_t(pix) * 3, nsamp ? nsamp + pix : nullptr};
  }
};

// Launch shape of the bounce loop (phx_tuning).
struct LoopShape {
  uint32_t trace_block;  // 128 or 256
  uint32_t trace_grid;   // work-groups of the persistent traversal kernels
  uint32_t shade_grid;   // work-groups (of 256) of the shade kernel: covers the path slots
  uint32_t regen_grid;   // work-groups (of 256) of the regeneration and NEE-accumulation kernels
  uint32_t chunk;        // work items a wave takes per fetch (multiple of 64)
  uint32_t refill;       // idle lanes of a traversal wave that trigger a refill (1..64)
  uint32_t groups;       // material groups of sorted shading (0: queue order)
  uint32_t volumes;      // a scene of the run holds volumes: k_volume runs before k_shade
  uint32_t tier;         // shading tier (kTierLean/Standard/Full) the run's scenes and views need
  uint32_t tiles;        // camera paths dealt in 8x8 pixel tiles instead of row order
};

// RenderParams -> the shading constants of one frame (render.cpp).
ShadeParams make_shade_params(const phx_camera &cam, const phx_render_params &p, uint32_t max_depth, uint32_t max_it);

// The view arguments of one launch: exactly one of the three is used.
struct ViewArgs {
  int kind;  // 0: OneItem, 1: ItemTable<true>, 2: ItemTable<false>
  OneItem one;
  ItemTable<true> shared;
  ItemTable<false> many;
};

// Fills the emitter records of s's emissive CDF (s.em_count entries of four float4).
hipError_t launch_em_records(const DevScene &s, float4 *out, hipStream_t st);
hipError_t launch_run_start(const StreamQueues &q, uint32_t total, uint32_t capacity, hipStream_t st);
// ahead = 0: fill the queue of the current iteration (start of a run);
// ahead = 1: fill the next iteration's queue and advance the iteration.
// forward = false (the adjoint replay) leaves the shading events uncounted.
// A new sample's stored radiance is not touched: its first shading event
// writes it.
hipError_t launch_regen(const ViewArgs &v, const StreamQueues &q, uint32_t ahead, const LoopShape &ls,
                        hipStream_t st, bool forward = true);
hipError_t launch_extend(const ViewArgs &v, const StreamQueues &q, const LoopShape &ls, hipStream_t st);
// With ls.groups > 0 the rays of each work-group are shaded in material-group order.
hipError_t launch_shade(const ViewArgs &v, const StreamQueues &q, const LoopShape &ls, hipStream_t st);
// The any-hit pass over the current iteration's shadow queue, and (runs with
// volumes) the transmittance along its unoccluded rays.
hipError_t launch_shadow(const ViewArgs &v, const StreamQueues &q, const LoopShape &ls, hipStream_t st);
// The unoccluded contributions of that queue into their samples' radiance.
hipError_t launch_nee(const ViewArgs &v, const StreamQueues &q, const LoopShape &ls, hipStream_t st);
hipError_t launch_volume(const ViewArgs &v, const StreamQueues &q, const LoopShape &ls, hipStream_t st);
// Sums the item's samples of each pixel into accum (first: overwrite).
// Adaptive sampling (pathtracer.cpp:700-723, 1129-1139): the reduce of one
// run of samples [s0, s0 + nspp) of a view that also keeps each pixel's sum
// of squared sample values and sample count, applies the stopping rule
// before each of those samples and decides whether the pixel takes sample
// s0 + nspp.  `active` is the mask the run's regeneration used.
hipError_t launch_adaptive_init(uint32_t npix, uint32_t *active, hipStream_t st);
hipError_t launch_reduce_adaptive(uint32_t npix, uint32_t nspp, const float4 *sample_L, float4 *accum,
                                  float4 *accum_sq, uint32_t *nsamp, uint32_t *active, uint32_t s0, hipStream_t st);
hipError_t launch_reduce(uint32_t npix, uint32_t nspp, const float4 *sample_L, float4 *accum, bool first,
                         hipStream_t st);
hipError_t launch_aov_init(uint32_t npix, uint32_t *aov_s, hipStream_t st);
// `hits`: room for one hit record per pixel of the view (scenes without
// volumes): the AOV rays are then traced ahead by the streaming traversal;
// null: the resolve traces them itself.  `fetch`: kFetchSetWords words.
hipError_t launch_resolve(const DevScene &s, const ShadeParams &sp, const StreamQueues &q, const FrameBuffers &f,
                          uint32_t ns, uint32_t *fetch, uint32_t grid, hipStream_t st, float4 *hits = nullptr);

// Path-replay adjoint (adjoint.hip).  Gradients are accumulated in double:
// grad_mat[material * 5 + k] for k = base r, g, b, roughness, metallic, and
// grad_tex[atlas pixel * 3 + c].  `adjoint` is dLoss/dImage (RGB per pixel,
// frame-buffer layout at each view's pix_base).
constexpr uint32_t kAdjLdsMaterials = 128;  // materials accumulated per work-group in LDS
struct AdjointArgs {
  const float *adjoint;
  double *grad_mat;
  double *grad_tex;  // null: texel gradients are not wanted
  uint32_t lds_mats;  // min(material count, kAdjLdsMaterials)
};
hipError_t launch_shade_adjoint(const ViewArgs &v, const StreamQueues &q, const AdjointArgs &a, const LoopShape &ls,
                                hipStream_t st);
hipError_t launch_grad_out(const double *src, float *dst, uint32_t n, hipStream_t st);
hipError_t launch_clear_f64(double *p, uint32_t n, hipStream_t st);
hipError_t launch_set_material_params(phx_material *mats, const float *p, uint32_t count, hipStream_t st);

// Seam A kernels (trace.hip): persistent grids of `grid` work-groups of 256
// with `spill_stride` ints of traversal-stack spill per thread in `spill`
// and the chunk counters in `fetch` (kFetchSetWords words, zeroed first).
constexpr uint32_t kFullSpill = kMaxStack - kShortStack;  // spill for any BVH the builder accepts
hipError_t launch_trace_compact(const DevScene &s, const float *org, const float *dir, const float *tmin,
                                const float *tmax, uint32_t n, phx_hit_compact *out, uint32_t *fetch, int *spill,
                                uint32_t grid, hipStream_t st, uint32_t spill_stride = kFullSpill);
hipError_t launch_trace_full(const DevScene &s, const float *org, const float *dir, const float *tmin,
                             const float *tmax, uint32_t n, phx_hit *out, uint32_t *fetch, int *spill, uint32_t grid,
                             hipStream_t st, uint32_t spill_stride = kFullSpill);
hipError_t launch_trace_occluded(const DevScene &s, const float *org, const float *dir, const float *tmin,
                                 const float *tmax, uint32_t n, uint32_t *out, uint32_t *fetch, int *spill,
                                 uint32_t grid, hipStream_t st, uint32_t spill_stride = kFullSpill);
hipError_t launch_clear_u32(uint32_t *p, uint32_t n, hipStream_t st);
// out[i] = (normalize(cross(e1, e2)), primitive id bits) of tris[i], i < n.
hipError_t launch_tri_normals(const TriRec *tris, float4 *out, uint32_t n, hipStream_t st);

}  // namespace phx
