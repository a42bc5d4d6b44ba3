// Internal: the context object behind the C ABI and its error helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/phx.h"
#include "dev_trace.h"
#include "integrator.h"

namespace phx_detail {

using namespace phx;

// Device allocation owned by the context.
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t reserve(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    release();
    if (n == 0) return hipSuccess;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    bytes = n;
    return hipSuccess;
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

}  // namespace phx_detail

struct phx_context {
  using DBuf = phx_detail::DBuf;
  int device = 0;
  hipStream_t stream = nullptr;
  // The bounce loop's second stream: an iteration's any-hit pass runs there
  // beside the next iteration's closest-hit pass, forked from and joined to
  // `stream` by the two events.
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int cu_count = 256;
  std::string error;
  phx_tuning tuning{};
  phx_stats stats{};

  // geometry (seam A)
  bool have_accel = false;
  DBuf nodes, tris, tri_attr, tri_ng, indices, positions, normals, texcoords, colors, material_ids;
  uint32_t triangle_count = 0, vertex_count = 0;
  float bounds_lo[3]{}, bounds_hi[3]{};
  uint32_t bvh_depth = 0;
  size_t scene_bytes = 0;

  // shading data
  bool have_scene = false;
  DBuf materials, lights, em_ids, em_cdf, em_guide, em_recs, prim_light_pdf, env_pixels, env_marginal, env_cond, env_guide_m, env_guide_c,
      tex_pixels, tex_infos, attributes, samplers, sampler_texels,
      volumes, vol_values, vol_tf, majorants;
  phx::DevScene dev{};
  uint32_t tex_pixel_count = 0;  // texels of the atlas

  // adjoint gradient accumulators (double)
  DBuf adj_gm, adj_gt;

  // Wavefront storage of one run of the bounce loop: its path and shadow
  // queues, per-sample radiance, counters, item tables and captured loop.
  // Lane 0 streams on `stream`/`stream2`; lane 1 (created on first use) lets
  // a batch start its next run while the previous one drains.
  struct Lane {
    DBuf q_o[2], q_d[2], q_tp[2], q_st[2], q_hit, sh_o[2], sh_d[2], sh_c[2], sample_L, counters, stats_buf;
    DBuf vol_ev, vol_rng;  // k_volume's results, sized for the queue when a run's scenes hold volumes
    DBuf bt_items, bt_first;
    uint32_t q_capacity = 0;
    hipStream_t st = nullptr, st2 = nullptr;  // lane 0: aliases of `stream` and `stream2`
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_read = nullptr, ev_reduced = nullptr;  // counter readback landed; the run's reduces done
    phx::RunItem *h_items = nullptr;  // pinned staging of the item tables (hipHostMalloc)
    uint32_t *h_first = nullptr;
    size_t h_count = 0;
    // captured bounce loop and the launch arguments it was captured with
    hipGraphExec_t graph = nullptr;
    std::vector<unsigned char> graph_key;
  };
  Lane lanes[2];
  DBuf aov_s, spill, fetch;
  uint32_t trace_grid = 0;  // persistent grid of the seam-A trace kernels and the resolve pass
  size_t spill_threads = 0;
  std::vector<hipEvent_t> events;  // kernel timing (phx_tuning.time_kernels)
  uint32_t *pinned = nullptr;      // host-visible counter readback (hipHostMalloc)

  // resolved frame (RenderResult) and its accumulator
  DBuf f_accum, f_color, f_alpha, f_depth, f_normal, f_albedo;
  // adaptive sampling: per pixel, the sum of squared sample values, the
  // samples taken and whether the pixel takes the next one
  DBuf f_accum_sq, f_nsamp, f_active;
  bool f_counted = false;  // the frame was adaptively sampled: f_nsamp holds its counts
  uint32_t f_w = 0, f_h = 0;
  bool have_result = false;

  // batched views (phx_render_batch): run tables, concatenated frame buffers
  DBuf bt_scenes, bt_accum, bt_color, bt_alpha, bt_depth, bt_normal, bt_albedo, bt_aov;
  struct BatchView {
    uint32_t width, height, pix_base;
  };
  std::vector<BatchView> bt_views;
  bool bt_have_result = false;

  // render bridge (seam B)
  DBuf b_color, b_alpha, b_depth, b_normal, b_albedo, b_bg;
  uint32_t b_w = 0, b_h = 0, b_samples = 0, b_counter = 0;
  std::vector<float> b_bg_host;  // the background image last uploaded to b_bg
  uint32_t b_bg_w = 0;

  // reference-structure baseline (proxy.hip): host mesh copy, its median-split
  // BVH (built on first use) and full-width per-slot state
  std::vector<float> host_positions;
  std::vector<uint32_t> host_indices;
  std::vector<uint32_t> host_material_ids;  // range-checked, empty without material ids
  bool px_built = false;
  uint32_t px_root = 0, px_slots = 0;
  DBuf px_nodes, px_prims, px_tris, px_tri_attr, px_o, px_d, px_tp, px_st, px_hit, px_sh_o, px_sh_d, px_sh_c, px_flags, px_accum,
      px_count;
};

namespace phx_detail {

phx_material default_material();
// noalloc.hip: looks up the static traversal scratch of `device` (current).
int noalloc_setup(int device);
phx_material convert_bridge_material(const phx_bridge_material &in);

inline int fail(phx_context *c, int code, const std::string &m) {
  if (c) c->error = m;
  return code;
}
inline int hip_fail(phx_context *c, hipError_t e, const char *what) {
  return fail(c, PHX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIP_TRY(c, expr)                                  \
  do {                                                    \
    hipError_t e_ = (expr);                               \
    if (e_ != hipSuccess) return hip_fail((c), e_, #expr); \
  } while (0)

template <class T>
inline int upload(phx_context *c, phx_detail::DBuf &b, const T *src, size_t count) {
  if (!src || count == 0) {
    b.release();
    return PHX_OK;
  }
  HIP_TRY(c, b.reserve(sizeof(T) * count));
  HIP_TRY(c, hipMemcpyAsync(b.p, src, sizeof(T) * count, hipMemcpyHostToDevice, c->stream));
  c->scene_bytes += sizeof(T) * count;
  return PHX_OK;
}

// Persistent traversal grid of the seam-A kernels and the resolve pass (8
// work-groups of 256 per CU), and a per-thread stack spill area for at least
// `threads` threads (0: that grid).
inline int ensure_trace_scratch(phx_context *c, size_t threads = 0) {
  if (c->trace_grid == 0) c->trace_grid = uint32_t(c->cu_count) * 8u;
  threads = std::max(threads, size_t(c->trace_grid) * 256);
  if (threads > c->spill_threads) {
    HIP_TRY(c, c->spill.reserve(threads * (phx::kMaxStack - phx::kShortStack) * sizeof(int)));
    c->spill_threads = threads;
  }
  HIP_TRY(c, c->fetch.reserve(phx::kFetchSetWords * 4));
  return PHX_OK;
}

}  // namespace phx_detail
