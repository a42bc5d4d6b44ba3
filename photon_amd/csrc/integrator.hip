// Streaming wavefront path tracer for gfx950.
//
// Replaces PathTracer::render (pathtracer.cpp:564-1159).  The reference runs,
// for each sample, max_depth + 9 bounce iterations over all W*H pixel slots
// (trace_closest, pt_shade, trace_occluded, pt_nee_accum_pack with a host
// readback).  Here a fixed pool of path slots streams through all
// (pixel, sample) pairs of the frame:
//   k_regen    starts camera paths (pt_init_sample, :691-760) in the free
//              slots of the next iteration's queue
//   k_extend   closest hit of the ray queue (trace_closest)
//   k_shade    pt_shade (:765-1090) via shade_path(); continuing paths and
//              shadow rays go to compacted queues (one atomic per work-group)
//   k_shadow   any hit + NEE accumulation (trace_occluded, pt_nee_accum_pack)
//   k_reduce   per-pixel sums over the samples, in sample order
//   k_resolve  mean, alpha and the AOVs (pt_resolve, :1143-1155)
// Each path carries its own bounce-loop iteration, which keys its random
// streams (pixel, sample, 1 + iteration) and (pixel, sample, 0x10000 +
// iteration) exactly as in the reference, so a path draws the same numbers
// whichever wavefront step it runs in.  Radiance is accumulated per sample and
// summed in sample order: results do not depend on scheduling.
#include "integrator.h"

#include "dev_queue.h"
#include "dev_stream.h"

namespace phx {

// Clears the counters of a run and records its size.
__global__ void k_run_start(StreamQueues q, uint32_t total, uint32_t capacity) {
  for (uint32_t i = threadIdx.x; i < kCounterWords; i += blockDim.x)
    q.counters[i] = i == kTotal ? total : (i == kCapacity ? capacity : 0u);
}

// The sample that regeneration slot g (of the slots [lo, hi) dealt out in
// one launch) starts, with samples dealt in 8x8 pixel tiles: the samples of
// each 8-row band of one sample index of a view that the launch deals out
// whole go in tile order, so 64 consecutive slots (one traversal wave) hold
// an 8x8 block of pixels instead of a 64x1 strip; bands dealt in part keep
// row order.  A bijection on [lo, hi) that stays within each band, so every
// sample is still started once, as the same (pixel, sample index).
__device__ __forceinline__ uint32_t tiled_sample(uint32_t g, uint32_t lo, uint32_t hi, const RunItem &r) {
  const uint32_t w = r.sp.width, npix = r.sp.npix;
  const uint32_t p = (g - r.first) % npix;
  const uint32_t band_len = 8u * w;
  const uint32_t band = p / band_len;
  if ((band + 1u) * 8u > r.sp.height) return g;
  const uint32_t base = g - (p - band * band_len);
  if (base < lo || base + band_len > hi) return g;
  const uint32_t o = g - base;
  const uint32_t full = (w >> 3) * 64u;  // positions in whole 8x8 tiles
  uint32_t x, y;
  if (o < full) {
    x = (o >> 6) * 8u + (o & 7u);
    y = (o >> 3) & 7u;
  } else {  // the last w % 8 columns of the band, row by row
    const uint32_t rem = w & 7u, k = o - full;
    y = k / rem;
    x = (w & ~7u) + k % rem;
  }
  return base + y * w + x;
}

// Fills the free slots of a queue with new camera paths (pt_init_sample).
// ahead = 0: the current iteration's queue; 1: the next one.
template <class P>
__global__ __launch_bounds__(256) void k_regen(P v, StreamQueues q, uint32_t ahead, uint32_t tiles) {
  const uint32_t it = q.counters[kIter] + ahead;
  const uint32_t used = q.counters[ray_count_word(it)];
  const uint32_t next = q.counters[kNextSample];
  const uint32_t free_slots = q.counters[kCapacity] - used;
  const uint32_t left = q.counters[kTotal] - next;
  const uint32_t take = free_slots < left ? free_slots : left;
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid == 0) q.counters[kRegenTake + ring(it)] = take;
  const int buf = it & 1;
  for (uint32_t i = tid; i < take; i += gridDim.x * blockDim.x) {
    uint32_t g = next + i;
    if (tiles) g = tiled_sample(g, next, next + take, v.item_at(v.find(g)));
    const uint32_t item = v.find(g);
    const RunItem &r = v.item_at(item);
    const uint32_t l = g - r.first;
    PathState p;
    const uint32_t pixel = l % r.sp.npix;
    camera_path(r.sp, pixel, r.sample_base + l / r.sp.npix, p);
    p.item = item;
    if (q.pix_active && q.pix_active[r.pix_base + pixel] == 0u) {
      // the pixel stopped sampling: an entry that traces and shades nothing
      p.tmin = 1.f;
      p.tmax = 0.f;
      p.kinds = kNoPath;
    }
    const uint32_t slot = used + i;
    q.o[buf][slot] = f4(p.ro, p.tmin);
    q.d[buf][slot] = f4(p.rd, p.tmax);
    q.tp[buf][slot] = f4(p.tp, p.prev_pdf);
    q.st[buf][slot] = make_uint4(g, pack_flags(p), p.kinds, __float_as_uint(p.path_dist));
  }
}

// Transfer-function entries the tracking kernels stage in LDS (one 256-entry
// table).  In a run of one scene whose tables fit, the table lookup that
// follows each density gather is an LDS read instead of a second dependent
// trip to the caches; the values, and so the results, are the same.  Every
// thread of the work-group calls this.
constexpr uint32_t kTfLds = 256;
template <class P>
__device__ __forceinline__ P with_tf_in_lds(const P &v, float4 *s_tf) {
  P w = v;
  if constexpr (std::is_same<P, OneItem>::value || std::is_same<P, ItemTable<true>>::value) {
    const uint32_t n = v.scene.vol_tf_count;
    if (n > 0 && n <= kTfLds) {
      for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) s_tf[j] = v.scene.vol_tf[j];
      __syncthreads();
      w.scene.vol_tf = s_tf;
    }
  }
  return w;
}

// The volume part of pt_shade (pathtracer.cpp:789-857), ahead of k_shade:
// the free-flight sample through the scene's volumes in front of each ray's
// hit, made with the shading event's own RNG stream (its first draws).  The
// delta-tracking loop is a chain of dependent density lookups; here it runs
// in a small kernel at six waves per SIMD (80 registers and a 12-byte spill
// under that cap; uncapped it takes 81 registers and five waves, and the
// 256^3 fog-box frame is 57.2 instead of 54.6 ms; capped at eight it spills
// more and was slower) against k_shade's four, so more of that latency is
// hidden, and k_shade no longer carries the loop.  k_shade picks the event
// and the RNG state up.  The fog-box frame went from 74.1 to 60.9 ms when
// the loop was split out.  Batches with per-lane scenes (ItemTable<false>)
// would spill 148 bytes under the cap and use k_volume_wide.
template <class P>
__device__ __forceinline__ void volume_events(const P &v0, const StreamQueues &q, float4 *s_tf) {
  const P v = with_tf_in_lds(v0, s_tf);
  const uint32_t it = q.counters[kIter];
  const uint32_t count = q.counters[ray_count_word(it)];
  const int buf = it & 1;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < count; i += gridDim.x * 256u) {
    const uint4 stv = q.st[buf][i];
    const RunItem &r = v.item_at(std::is_same<P, OneItem>::value ? 0u : stv.y >> kItemShift);
    const DevScene &s = v.scene_of(r);
    if (s.vol_count == 0) continue;
    const float4 o = q.o[buf][i], d = q.d[buf][i], hit = q.hit[i];
    const uint32_t l = stv.x - r.first;
    Rng rng = make_rng(l % r.sp.npix, r.sample_base + l / r.sp.npix, 1u + ((stv.y >> 8) & 0xffu));
    const bool hr_hit = hit_tri(__float_as_uint(hit.w)) != 0xffffffffu;
    vec3 alb{0.f, 0.f, 0.f};
    const float ts = vol_sample_scatter(s, xyz(o), xyz(d), o.w, hr_hit ? hit.x : d.w, rng, alb);
    q.vol_ev[i] = make_float4(ts, alb.x, alb.y, alb.z);
    q.vol_rng[i] = rng.s;
  }
}
template <class P>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_volume(P v, StreamQueues q) {
  __shared__ float4 s_tf[kTfLds];
  volume_events(v, q, s_tf);
}
template <class P>
__global__ __launch_bounds__(256) void k_volume_wide(P v, StreamQueues q) {
  __shared__ float4 s_tf[kTfLds];
  volume_events(v, q, s_tf);
}

// The emitter records of the emissive CDF (em_record), rebuilt whenever the
// scene's materials or textures change.
__global__ __launch_bounds__(256) void k_em_records(DevScene s, float4 *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < s.em_count) em_record(s, i, out + size_t(i) * 4);
}
hipError_t launch_em_records(const DevScene &s, float4 *out, hipStream_t st) {
  if (s.em_count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_em_records, dim3((s.em_count + 255) / 256), dim3(256), 0, st, s, out);
  return hipGetLastError();
}

// Commits the regenerated paths, clears the ring entries of the following
// iteration and (ahead = 1) advances the iteration.  With count_shade, the
// shading events of the iteration that just ran (its ray queue) go into the
// statistics here, by one thread, rather than by an atomic per work-group of
// the shade kernel on one shared word.
__global__ void k_commit(StreamQueues q, uint32_t ahead, uint32_t count_shade) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (ahead && count_shade) q.stats[kStatShade] += q.counters[ray_count_word(q.counters[kIter])];
    const uint32_t it = q.counters[kIter] + ahead;
    const uint32_t take = q.counters[kRegenTake + ring(it)];
    q.counters[ray_count_word(it)] += take;
    q.counters[kNextSample] += take;
    // entries of iteration it + 1 (last used by it + 1 - kRing)
    q.counters[ray_count_word(it + 1)] = 0u;
    q.counters[shadow_count_word(it + 1)] = 0u;
    for (uint32_t k = 0; k < kStripes; ++k) {
      q.counters[fetch_set(kFetchExtend, it + 1) + k * kStripeStride] = 0u;
      q.counters[fetch_set(kFetchShadow, it + 1) + k * kStripeStride] = 0u;
    }
    q.counters[kIter] = it;
  }
}

// The acceleration structure of a lane's ray: the kernel-argument scene, or the
// scene-table entry of the ray's item.
template <class P>
PHX_DEV void ray_accel(const P &v, uint32_t item, const BvhNode *&nodes, const TriRec *&tris) {
  if constexpr (std::is_same<P, OneItem>::value || std::is_same<P, ItemTable<true>>::value) {
    nodes = v.scene.nodes;
    tris = v.scene.tris;
  } else {
    const DevScene &sc = v.scene_of(v.item_at(item));
    nodes = sc.nodes;
    tris = sc.tris;
  }
}

// Launches whose rays all trace one scene (one item, or items of one scene)
// pass its structure to the traversal once, as wave-uniform pointers.
template <class P>
constexpr bool kOneAccel = std::is_same<P, OneItem>::value || std::is_same<P, ItemTable<true>>::value;
template <class P>
PHX_DEV const BvhNode *one_accel_nodes(const P &v) {
  if constexpr (kOneAccel<P>) return v.scene.nodes;
  else return nullptr;
}
template <class P>
PHX_DEV const TriRec *one_accel_tris(const P &v) {
  if constexpr (kOneAccel<P>) return v.scene.tris;
  else return nullptr;
}

// Material id of primitive `prim` in the scene of the ray's item.
template <class P>
PHX_DEV uint32_t ray_material(const P &v, uint32_t item, uint32_t prim) {
  const uint32_t *mids;
  if constexpr (std::is_same<P, OneItem>::value || std::is_same<P, ItemTable<true>>::value) {
    mids = v.scene.material_ids;
  } else {
    mids = v.scene_of(v.item_at(item)).material_ids;
  }
  return mids ? mids[prim] : 0u;
}

// Closest hit over the ray queue of the current iteration (persistent waves,
// per-lane refill).  With groups > 0 the hit record's top 4 bits carry the
// hit's material group.  Six waves per SIMD: the kernel fits 80 registers
// without spills.
// TAG: the run shades by material group (more than one), so hits carry their
// group; without it the tagging code is left out, which spares registers.
template <class P, int B, bool TAG>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(6))) void k_extend(P v, StreamQueues q,
                                                                                      uint32_t chunk, uint32_t refill,
                                                                                      uint32_t groups) {
  const uint32_t it = q.counters[kIter];
  const uint32_t count = q.counters[ray_count_word(it)];
  if (count == 0) return;
  __shared__ int lds[kShortStack * B];
  const int buf = it & 1;
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  LaneStack st{&lds[threadIdx.x], B, q.spill + size_t(gtid) * (kMaxStack - kShortStack), 0};
  uint32_t live = 0;
  traverse_stream<false, kOneAccel<P>>(
      st, &q.counters[fetch_set(kFetchExtend, it)], count, chunk, refill,
      [&](uint32_t i, vec3 &o, vec3 &d, float &tmin, float &tmax, const BvhNode *&nodes, const TriRec *&tris) {
        const float4 a = q.o[buf][i], b = q.d[buf][i];
        o = xyz(a);
        d = xyz(b);
        tmin = a.w;
        tmax = b.w;
        live += tmax > tmin ? 1u : 0u;
        ray_accel(v, q.st[buf][i].y >> kItemShift, nodes, tris);
      },
      [&](uint32_t i, bool hit, const TraceHit &h) {
        uint32_t w = h.tri;
        // one group: the shade kernel only parts hits from misses, no tag needed
        if (TAG && hit) w |= (ray_material(v, q.st[buf][i].y >> kItemShift, h.prim) % groups) << 28;
        q.hit[i] = make_float4(h.t, h.u, h.v, __uint_as_float(w));
      },
      one_accel_nodes(v), one_accel_tris(v));
  block_sum<B / 64>(&q.stats[kStatClosest], live);
}

template <class P, bool SORTED, int TIER>
__device__ __forceinline__ void shade_tile(const P &v, const StreamQueues &q, uint32_t groups, uint32_t it,
                                           uint32_t count, uint32_t tile) {
  uint32_t i = tile + threadIdx.x;
  const bool valid = i < count;
  if (SORTED) {
    // counting sort of the work-group's 256 rays by material group (misses
    // last) in LDS: each wave then shades one or two groups, reading rays
    // from the group's own 4 KB range of the queue
    __shared__ uint32_t s_cnt[kMaxClasses + 1];
    __shared__ uint32_t s_perm[256];
    if (threadIdx.x <= kMaxClasses) s_cnt[threadIdx.x] = 0u;
    __syncthreads();
    uint32_t key = kMaxClasses;
    if (valid) {
      const uint32_t w = __float_as_uint(q.hit[i].w);
      key = w == 0xffffffffu ? groups : (w >> 28);
    }
    const uint32_t rank = atomicAdd(&s_cnt[key], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t acc = 0;
      for (uint32_t c = 0; c <= kMaxClasses; ++c) {
        const uint32_t n = s_cnt[c];
        s_cnt[c] = acc;
        acc += n;
      }
    }
    __syncthreads();
    s_perm[s_cnt[key] + rank] = threadIdx.x;
    __syncthreads();
    i = tile + s_perm[threadIdx.x];
  }
  const int buf = it & 1, nbuf = buf ^ 1;
  ShadeOut out;
  out.cont = false;
  out.shadow = false;
  uint32_t g = 0;
  uint4 stv = make_uint4(0u, 0u, 0u, 0u);
  if (valid) stv = q.st[buf][i];
  g = stv.x;
  // One shading event.  (Looping over the wave's distinct items to make the
  // item wave-uniform, so the view and scene would load through scalar loads,
  // measured slower: the loop raised the spill traffic.)
  auto shade_one = [&](uint32_t item) {
    if (stv.z & kNoPath) return;  // a sample its pixel did not take
    const RunItem &r = v.item_at(item);
    const DevScene &s = v.scene_of(r);
    const uint32_t l = g - r.first;
    const uint32_t pixel = l % r.sp.npix;
    const uint32_t s_local = l / r.sp.npix;
    PathState ps = unpack(q.o[buf][i], q.d[buf][i], q.tp[buf][i], stv);
    ps.pixel = pixel;
    const float4 hit = q.hit[i];
    const uint32_t htri = hit_tri(__float_as_uint(hit.w));
    Rng rng = make_rng(pixel, r.sample_base + s_local, 1u + ps.path_it);
    float4 vol_ev = make_float4(-1.f, 0.f, 0.f, 0.f);
    uint32_t vol_rng = 0u;
    if (s.vol_count > 0) {  // k_volume ran for every run with a volume scene
      vol_ev = q.vol_ev[i];
      vol_rng = q.vol_rng[i];
    }
    // a path's first event starts its sample's radiance (regeneration leaves
    // the slot as it was); its shadow rays add to it in later iterations
    const bool first_event = ps.path_it == 0u;
    NoHooks hooks;
    shade_path<NoHooks, true, TIER>(s, r.sp, ps, htri != 0xffffffffu, htri, hit.x, hit.y, hit.z, rng, out, hooks,
                                    vol_ev, vol_rng);
    if (first_event) {
      q.sample_L[g] = make_float4(out.add.x, out.add.y, out.add.z, out.add_alpha);
    } else if (out.add.x != 0.f || out.add.y != 0.f || out.add.z != 0.f || out.add_alpha != 0.f) {
      const float4 L = q.sample_L[g];
      q.sample_L[g] = make_float4(L.x + out.add.x, L.y + out.add.y, L.z + out.add.z, L.w + out.add_alpha);
    }
    if (out.primary_event) atomicMin(&q.aov_s[r.pix_base + pixel], r.s0 + s_local);
    // the reference's loop bound (pathtracer.cpp:682,762)
    if (out.next.path_it >= r.sp.max_it) out.cont = false;
  };
  if (valid) shade_one(std::is_same<P, OneItem>::value ? 0u : stv.y >> kItemShift);
  uint32_t ci = 0, si = 0;
  block_alloc_pair<4>(reinterpret_cast<unsigned long long *>(&q.counters[ray_count_word(it + 1)]), out.cont, ci,
                      out.shadow, si);
  if (out.cont) {
    const PathState &n = out.next;
    q.o[nbuf][ci] = f4(n.ro, n.tmin);
    q.d[nbuf][ci] = f4(n.rd, n.tmax);
    q.tp[nbuf][ci] = f4(n.tp, n.prev_pdf);
    q.st[nbuf][ci] = make_uint4(g, pack_flags(n), n.kinds, __float_as_uint(n.path_dist));
  }
  if (out.shadow) {
    // the shading path's iteration keys the transmittance stream
    const uint32_t sb = (it + 1u) & 1u;
    q.sh_o[sb][si] = f4(out.so, __uint_as_float((out.next.path_it - 1u) | (out.next.item << 8)));
    q.sh_d[sb][si] = f4(out.sd, out.stmax);
    q.sh_c[sb][si] = f4(out.sc, __uint_as_float(g));
  }
}

// pt_shade over the ray queue, one thread per path; continuing paths and
// shadow rays go to compacted queues, one atomic per work-group.  The grid covers the
// path slots and work-groups beyond the queue exit at once: a grid-stride loop
// over the queue measured slower, since the loop raises the spill traffic.
template <class P, bool SORTED, int TIER>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_shade(P v, StreamQueues q,
                                                                                      uint32_t groups) {
  const uint32_t it = q.counters[kIter];
  const uint32_t count = q.counters[ray_count_word(it)];
  const uint32_t tile = blockIdx.x * 256u;
  if (tile >= count) return;
  shade_tile<P, SORTED, TIER>(v, q, groups, it, count, tile);
}

// Any hit over the shadow queue (trace_occluded; persistent waves, per-lane
// refill; six waves per SIMD as k_extend).  It runs in the iteration after the
// shade kernel that filled the queue, beside that iteration's k_extend and
// k_shade: the queue is the one counted for the current iteration, and that
// iteration's shading fills the other one.
template <class P, int B>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(6))) void k_shadow(P v, StreamQueues q,
                                                                                      uint32_t chunk, uint32_t refill) {
  const uint32_t it = q.counters[kIter];
  const uint32_t count = q.counters[shadow_count_word(it)];
  if (count == 0) return;
  const float4 *sh_o = q.sh_o[it & 1u], *sh_d = q.sh_d[it & 1u];
  float4 *sh_c = q.sh_c[it & 1u];
  __shared__ int lds[kShortStack * B];
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  LaneStack st{&lds[threadIdx.x], B, q.spill_shadow + size_t(gtid) * (kMaxStack - kShortStack), 0};
  uint32_t live = 0;
  traverse_stream<true, kOneAccel<P>>(
      st, &q.counters[fetch_set(kFetchShadow, it)], count, chunk, refill,
      [&](uint32_t i, vec3 &o, vec3 &d, float &tmin, float &tmax, const BvhNode *&nodes, const TriRec *&tris) {
        const float4 a = sh_o[i], b = sh_d[i];
        o = xyz(a);
        d = xyz(b);
        tmin = 0.f;
        tmax = b.w;
        live += tmax > 0.f ? 1u : 0u;
        ray_accel(v, __float_as_uint(a.w) >> 8, nodes, tris);
      },
      [&](uint32_t i, bool occ, const TraceHit &) {
        // occluded entries lose their sample id; k_nee_accumulate adds the rest
        if (occ) sh_c[i].w = __uint_as_float(kNoRay);
      },
      one_accel_nodes(v), one_accel_tris(v));
  block_sum<B / 64>(&q.stats[kStatShadow], live);
}

// The volume transmittance along the unoccluded shadow rays of the current
// queue (the ratio tracking of pt_nee_accum_pack, pathtracer.cpp:1095-1121),
// folded into their contributions, for runs whose scenes hold volumes.  It
// follows k_shadow on the same stream, beside the next bounce's closest-hit
// and shading passes, so this chain of dependent density lookups is off the
// path that k_nee_accumulate's join puts after shading.  The draws come from
// the shadow ray's own stream, so the result does not depend on where it runs.
template <class P>
__global__ __launch_bounds__(256) void k_shadow_transmit(P v0, StreamQueues q) {
  __shared__ float4 s_tf[kTfLds];
  const P v = with_tf_in_lds(v0, s_tf);
  const uint32_t it = q.counters[kIter];
  const uint32_t count = q.counters[shadow_count_word(it)];
  const uint32_t sb = it & 1u;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < count; i += gridDim.x * 256u) {
    const float4 c = q.sh_c[sb][i];
    const uint32_t g = __float_as_uint(c.w);
    if (g == kNoRay) continue;
    const float4 o = q.sh_o[sb][i];
    const uint32_t key = __float_as_uint(o.w);  // path_it | item << 8
    const RunItem &r = v.item_at(key >> 8);
    const DevScene &s = v.scene_of(r);
    if (s.vol_count == 0) continue;
    const float4 d = q.sh_d[sb][i];
    const uint32_t l = g - r.first;
    Rng rng = make_rng(l % r.sp.npix, r.sample_base + l / r.sp.npix, 0x10000u + (key & 0xffu));
    const vec3 contrib = xyz(c) * vol_transmittance(s, xyz(o), xyz(d), 0.f, d.w, rng);
    q.sh_c[sb][i] = make_float4(contrib.x, contrib.y, contrib.z, c.w);
  }
}

// pt_nee_accum_pack (pathtracer.cpp:1095-1121): the unoccluded shadow rays'
// contributions (with their volume transmittance from k_shadow_transmit) into
// their samples' radiance.  One thread per shadow ray, over the queue in
// strides of the grid; after k_shadow and k_shade, on the same queue.
template <class P>
__global__ __launch_bounds__(256) void k_nee_accumulate(P, StreamQueues q) {
  const uint32_t it = q.counters[kIter];
  const uint32_t count = q.counters[shadow_count_word(it)];
  const uint32_t sb = it & 1u;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < count; i += gridDim.x * 256u) {
    const float4 c = q.sh_c[sb][i];
    const uint32_t g = __float_as_uint(c.w);
    if (g == kNoRay) continue;
    const vec3 contrib = xyz(c);
    if (finite3(contrib)) {
      const float4 L = q.sample_L[g];
      q.sample_L[g] = make_float4(L.x + contrib.x, L.y + contrib.y, L.z + contrib.z, L.w);
    }
  }
}
// Sums an item's samples of each pixel into the frame, in sample order.
__global__ __launch_bounds__(256) void k_reduce(uint32_t npix, uint32_t nspp, const float4 *sample_L, float4 *accum,
                                                uint32_t first) {
  const uint32_t pixel = blockIdx.x * blockDim.x + threadIdx.x;
  if (pixel >= npix) return;
  float4 acc = first ? make_float4(0.f, 0.f, 0.f, 0.f) : accum[pixel];
  for (uint32_t s = 0; s < nspp; ++s) {
    const float4 L = sample_L[size_t(s) * npix + pixel];
    acc.x += L.x;
    acc.y += L.y;
    acc.z += L.z;
    acc.w += L.w;
  }
  accum[pixel] = acc;
}

// Adaptive sampling: before sample s >= 16 a pixel with at least 16 samples
// stops when its largest per-channel variance over its squared luminance is
// below 0.01 (pathtracer.cpp:676-723); the squared sample values are summed
// after each sample from the difference of the running sums (:1129-1139).
constexpr uint32_t kAdaptiveMinSamples = 16;
constexpr float kAdaptiveThreshold = 0.01f;
__device__ __forceinline__ bool adaptive_stop(uint32_t s, uint32_t ns, const float4 &acc, const float4 &sq) {
  if (s < kAdaptiveMinSamples || ns < kAdaptiveMinSamples) return false;
  const float inv = 1.f / float(ns);
  const float mx = acc.x * inv, my = acc.y * inv, mz = acc.z * inv;
  const float vx = sq.x * inv - mx * mx, vy = sq.y * inv - my * my, vz = sq.z * inv - mz * mz;
  const float max_var = fmaxf(vx, fmaxf(vy, vz));
  const float lum = fmaxf(0.2126f * mx + 0.7152f * my + 0.0722f * mz, 1e-4f);
  return max_var / (lum * lum) < kAdaptiveThreshold;
}
__global__ void k_adaptive_init(uint32_t npix, uint32_t *active) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < npix) active[p] = 1u;
}
// Adds samples s0 .. s0+nspp-1 of the pixels that took them, in order, each
// after the stopping check for its index; a pixel that stops keeps none of
// the later ones.  `active` then says whether it takes sample s0+nspp.
__global__ __launch_bounds__(256) void k_reduce_adaptive(uint32_t npix, uint32_t nspp, const float4 *sample_L,
                                                         float4 *accum, float4 *accum_sq, uint32_t *nsamp,
                                                         uint32_t *active, uint32_t s0) {
  const uint32_t pixel = blockIdx.x * blockDim.x + threadIdx.x;
  if (pixel >= npix) return;
  const bool first = s0 == 0;
  float4 acc = first ? make_float4(0.f, 0.f, 0.f, 0.f) : accum[pixel];
  float4 sq = first ? make_float4(0.f, 0.f, 0.f, 0.f) : accum_sq[pixel];
  uint32_t ns = first ? 0u : nsamp[pixel];
  bool go = active[pixel] != 0u;
  for (uint32_t s = 0; s < nspp && go; ++s) {
    if (adaptive_stop(s0 + s, ns, acc, sq)) {
      go = false;
      break;
    }
    const float4 L = sample_L[size_t(s) * npix + pixel];
    const float4 before = acc;
    acc.x += L.x;
    acc.y += L.y;
    acc.z += L.z;
    acc.w += L.w;
    const float dx = acc.x - before.x, dy = acc.y - before.y, dz = acc.z - before.z;
    sq.x += dx * dx;
    sq.y += dy * dy;
    sq.z += dz * dz;
    ++ns;
  }
  accum[pixel] = acc;
  accum_sq[pixel] = sq;
  nsamp[pixel] = ns;
  if (go && adaptive_stop(s0 + nspp, ns, acc, sq)) go = false;
  active[pixel] = go ? 1u : 0u;
}
hipError_t launch_adaptive_init(uint32_t npix, uint32_t *active, hipStream_t st) {
  hipLaunchKernelGGL(k_adaptive_init, dim3((npix + 255) / 256), dim3(256), 0, st, npix, active);
  return hipGetLastError();
}
hipError_t launch_reduce_adaptive(uint32_t npix, uint32_t nspp, const float4 *sample_L, float4 *accum,
                                  float4 *accum_sq, uint32_t *nsamp, uint32_t *active, uint32_t s0, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_adaptive, dim3((npix + 255) / 256), dim3(256), 0, st, npix, nspp, sample_L, accum,
                     accum_sq, nsamp, active, s0);
  return hipGetLastError();
}

__global__ void k_aov_init(uint32_t npix, uint32_t *aov_s) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < npix) aov_s[p] = 0xffffffffu;
}

// The AOV samples' camera rays traced ahead of the resolve (scenes without
// volumes): pixel p's first sample with a camera-ray event (aov_s) is traced
// again by the persistent closest-hit traversal, its hit record stored in
// hits[p] (t, b1, b2, leaf triangle; 0xffffffff: a miss or no such sample).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_aov_trace(DevScene s, ShadeParams w,
                                                                                          StreamQueues q,
                                                                                          uint32_t *fetch,
                                                                                          float4 *hits) {
  __shared__ int lds[kShortStack * 256];
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  LaneStack st{&lds[threadIdx.x], 256, q.spill + size_t(gtid) * (kMaxStack - kShortStack), 0};
  traverse_stream<false, true>(
      st, fetch, w.npix, kChunk, 16u,
      [&](uint32_t p, vec3 &o, vec3 &d, float &tmin, float &tmax, const BvhNode *&, const TriRec *&) {
        const uint32_t s_first = q.aov_s[p];
        if (s_first == 0xffffffffu) {
          o = vec3{0.f, 0.f, 0.f};
          d = vec3{0.f, 0.f, 1.f};
          tmin = 1.f;
          tmax = 0.f;
          return;
        }
        PathState ps;
        camera_path(w, p, w.sample_offset + s_first, ps);
        o = ps.ro;
        d = ps.rd;
        tmin = ps.tmin;
        tmax = ps.tmax;
      },
      [&](uint32_t p, bool hit, const TraceHit &h) {
        hits[p] = make_float4(h.t, h.u, h.v, __uint_as_float(hit ? h.tri : 0xffffffffu));
      },
      s.nodes, s.tris);
}

// pt_resolve plus the AOVs: the first sample whose camera ray ended at a
// surface or a volume event is traced again up to that event
// (pathtracer.cpp:800-809, 974-983) to read depth, normal and albedo.  With
// `hits` (k_aov_trace ran) its first trace is read from there.
__global__ __launch_bounds__(256) void k_resolve(DevScene s, ShadeParams w, StreamQueues q, FrameBuffers f, uint32_t ns,
                                                 uint32_t *fetch, const float4 *hits) {
  __shared__ int lds[kShortStack * 256];
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  LaneStack st{&lds[threadIdx.x], 256, q.spill + size_t(gtid) * (kMaxStack - kShortStack), 0};
  const uint32_t lane = lane_id();
  const float inv = ns > 0 ? 1.f / float(ns) : 0.f;
  WorkCursor cur = first_chunk(kChunk);
  while (true) {
    const uint32_t base = next_batch(cur, fetch, w.npix, kChunk);
    if (base >= w.npix) break;
    const uint32_t p = base + lane;
    if (p >= w.npix) continue;
    const float4 a = f.accum[p];
    const float ip = f.nsamp ? (f.nsamp[p] > 0u ? 1.f / float(f.nsamp[p]) : 0.f) : inv;
    f.color[p * 3] = a.x * ip;
    f.color[p * 3 + 1] = a.y * ip;
    f.color[p * 3 + 2] = a.z * ip;
    f.alpha[p] = a.w * ip;
    float depth = 1e30f;
    vec3 nrm{0.f, 0.f, 0.f}, alb{0.f, 0.f, 0.f};
    const uint32_t s_first = q.aov_s[p];
    if (s_first != 0xffffffffu) {
      const uint32_t sample = w.sample_offset + s_first;
      PathState ps;
      camera_path(w, p, sample, ps);
      float path_dist = 0.f;
      for (uint32_t it = 0; it < w.max_it; ++it) {
        TraceHit h;
        bool hit;
        if (it == 0 && hits) {
          const float4 hv = hits[p];
          h.t = hv.x;
          h.u = hv.y;
          h.v = hv.z;
          h.tri = __float_as_uint(hv.w);
          hit = h.tri != 0xffffffffu;
        } else {
          hit = traverse<false>(s, ps.ro, ps.rd, ps.tmin, ps.tmax, st, h);
        }
        Rng rng = make_rng(p, sample, 1u + it);
        if (s.vol_count > 0) {
          vec3 va{0.f, 0.f, 0.f};
          const float ts = vol_sample_scatter(s, ps.ro, ps.rd, ps.tmin, hit ? h.t : ps.tmax, rng, va);
          if (ts >= 0.f) {
            depth = path_dist + ts;
            nrm = -ps.rd;
            alb = va;
            break;
          }
        }
        if (!hit) break;
        const SurfaceHit sh = surface_at(s, h.tri, h.t, h.u, h.v, ps.ro, ps.rd);
        float alpha = 1.f;
        const phx_material mat = resolve_material(s, sh, alpha);
        const bool hidden = (mat.flags & PHX_MATERIAL_FLAG_HIDDEN_TO_CAMERA) != 0;
        if (hidden || (alpha < 1.f && rng.next() >= alpha)) {
          path_dist += h.t;
          ps.ro = sh.p;
          ps.tmin = s.ray_eps;
          ps.tmax = 1e30f;
          float cn, cf;
          clip_range(w.cam, ps.rd, cn, cf);
          if (cf < 1e30f) {
            ps.tmax = cf - path_dist;
            if (ps.tmax < s.ray_eps) {
              ps.tmin = 1.f;
              ps.tmax = 0.f;
            }
          }
          continue;
        }
        vec3 sn = sh.sn;
        if (!(mat.transmission > 0.f) && !(dot(sh.ng, -ps.rd) >= 0.f)) sn = sn * -1.f;
        if (mat.normal_tex >= 0 && s.tex_count > 0) {
          vec3 nm = tex_sample(s, mat.normal_tex, sh.uv.x, sh.uv.y);
          nm = nm * 2.f - vec3{1.f, 1.f, 1.f};
          vec3 ta, bi;
          onb(sn, ta, bi);
          sn = normalize(ta * nm.x + bi * nm.y + sn * nm.z);
        }
        if (mat.normal_src.kind != PHX_SOURCE_CONSTANT) sn = apply_normal_map(s, mat.normal_src, sh, sn);
        depth = path_dist + h.t;
        nrm = normalize(sn);
        alb = base_of(mat);
        break;
      }
    }
    f.depth[p] = depth;
    f.normal[p * 3] = nrm.x;
    f.normal[p * 3 + 1] = nrm.y;
    f.normal[p * 3 + 2] = nrm.z;
    f.albedo[p * 3] = alb.x;
    f.albedo[p * 3 + 1] = alb.y;
    f.albedo[p * 3 + 2] = alb.z;
  }
}

// ---- host launchers -----------------------------------------------------------
namespace {
template <class F>
hipError_t with_view(const ViewArgs &v, F f) {
  switch (v.kind) {
    case 0: return f(v.one);
    case 1: return f(v.shared);
    default: return f(v.many);
  }
}
}  // namespace

hipError_t launch_run_start(const StreamQueues &q, uint32_t total, uint32_t capacity, hipStream_t st) {
  hipLaunchKernelGGL(k_run_start, dim3(1), dim3(128), 0, st, q, total, capacity);
  return hipGetLastError();
}
hipError_t launch_regen(const ViewArgs &v, const StreamQueues &q, uint32_t ahead, const LoopShape &ls,
                        hipStream_t st, bool forward) {
  return with_view(v, [&](const auto &p) {
    using P = std::decay_t<decltype(p)>;
    hipLaunchKernelGGL(k_regen<P>, dim3(ls.regen_grid), dim3(256), 0, st, p, q, ahead, ls.tiles);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the forward pass counts its shading events, the adjoint replay does not
    hipLaunchKernelGGL(k_commit, dim3(1), dim3(64), 0, st, q, ahead, forward ? 1u : 0u);
    return hipGetLastError();
  });
}
hipError_t launch_extend(const ViewArgs &v, const StreamQueues &q, const LoopShape &ls, hipStream_t st) {
  return with_view(v, [&](const auto &p) {
    using P = std::decay_t<decltype(p)>;
    const bool tag = ls.groups > 1;
    if (ls.trace_block == 128 && tag)
      hipLaunchKernelGGL((k_extend<P, 128, true>), dim3(ls.trace_grid), dim3(128), 0, st, p, q, ls.chunk, ls.refill,
                         ls.groups);
    else if (ls.trace_block == 128)
      hipLaunchKernelGGL((k_extend<P, 128, false>), dim3(ls.trace_grid), dim3(128), 0, st, p, q, ls.chunk, ls.refill,
                         ls.groups);
    else if (tag)
      hipLaunchKernelGGL((k_extend<P, 256, true>), dim3(ls.trace_grid), dim3(256), 0, st, p, q, ls.chunk, ls.refill,
                         ls.groups);
    else
      hipLaunchKernelGGL((k_extend<P, 256, false>), dim3(ls.trace_grid), dim3(256), 0, st, p, q, ls.chunk, ls.refill,
                         ls.groups);
    return hipGetLastError();
  });
}
hipError_t launch_shade(const ViewArgs &v, const StreamQueues &q, const LoopShape &ls, hipStream_t st) {
  return with_view(v, [&](const auto &p) {
    using P = std::decay_t<decltype(p)>;
    const dim3 g(ls.shade_grid), b(256);
    const uint32_t groups = ls.groups;
    if (groups) {
      if (ls.tier == kTierLean) hipLaunchKernelGGL((k_shade<P, true, kTierLean>), g, b, 0, st, p, q, groups);
      else if (ls.tier == kTierStandard) hipLaunchKernelGGL((k_shade<P, true, kTierStandard>), g, b, 0, st, p, q, groups);
      else hipLaunchKernelGGL((k_shade<P, true, kTierFull>), g, b, 0, st, p, q, groups);
    } else {
      if (ls.tier == kTierLean) hipLaunchKernelGGL((k_shade<P, false, kTierLean>), g, b, 0, st, p, q, 0u);
      else if (ls.tier == kTierStandard) hipLaunchKernelGGL((k_shade<P, false, kTierStandard>), g, b, 0, st, p, q, 0u);
      else hipLaunchKernelGGL((k_shade<P, false, kTierFull>), g, b, 0, st, p, q, 0u);
    }
    return hipGetLastError();
  });
}
hipError_t launch_shadow(const ViewArgs &v, const StreamQueues &q, const LoopShape &ls, hipStream_t st) {
  return with_view(v, [&](const auto &p) {
    using P = std::decay_t<decltype(p)>;
    if (ls.trace_block == 128)
      hipLaunchKernelGGL((k_shadow<P, 128>), dim3(ls.trace_grid), dim3(128), 0, st, p, q, ls.chunk, ls.refill);
    else
      hipLaunchKernelGGL((k_shadow<P, 256>), dim3(ls.trace_grid), dim3(256), 0, st, p, q, ls.chunk, ls.refill);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !ls.volumes) return e;
    hipLaunchKernelGGL(k_shadow_transmit<P>, dim3(ls.regen_grid), dim3(256), 0, st, p, q);
    return hipGetLastError();
  });
}
hipError_t launch_nee(const ViewArgs &v, const StreamQueues &q, const LoopShape &ls, hipStream_t st) {
  return with_view(v, [&](const auto &p) {
    using P = std::decay_t<decltype(p)>;
    hipLaunchKernelGGL(k_nee_accumulate<P>, dim3(ls.regen_grid), dim3(256), 0, st, p, q);
    return hipGetLastError();
  });
}
hipError_t launch_volume(const ViewArgs &v, const StreamQueues &q, const LoopShape &ls, hipStream_t st) {
  return with_view(v, [&](const auto &p) {
    using P = std::decay_t<decltype(p)>;
    if constexpr (std::is_same<P, ItemTable<false>>::value)
      hipLaunchKernelGGL(k_volume_wide<P>, dim3(ls.regen_grid), dim3(256), 0, st, p, q);
    else
      hipLaunchKernelGGL(k_volume<P>, dim3(ls.regen_grid), dim3(256), 0, st, p, q);
    return hipGetLastError();
  });
}
hipError_t launch_reduce(uint32_t npix, uint32_t nspp, const float4 *sample_L, float4 *accum, bool first,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_reduce, dim3((npix + 255) / 256), dim3(256), 0, st, npix, nspp, sample_L, accum,
                     first ? 1u : 0u);
  return hipGetLastError();
}
hipError_t launch_aov_init(uint32_t npix, uint32_t *aov_s, hipStream_t st) {
  hipLaunchKernelGGL(k_aov_init, dim3((npix + 255) / 256), dim3(256), 0, st, npix, aov_s);
  return hipGetLastError();
}
hipError_t launch_resolve(const DevScene &s, const ShadeParams &sp, const StreamQueues &q, const FrameBuffers &f,
                          uint32_t ns, uint32_t *fetch, uint32_t grid, hipStream_t st, float4 *hits) {
  if (hits) {
    // the AOV rays first, through the streaming traversal (fetch: its stripe
    // counters, then the resolve's one counter)
    hipError_t e = launch_clear_u32(fetch, kFetchSetWords, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_aov_trace, dim3(grid), dim3(256), 0, st, s, sp, q, fetch, hits);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = launch_clear_u32(fetch, 1, st)) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_resolve, dim3(grid), dim3(256), 0, st, s, sp, q, f, ns, fetch, static_cast<const float4 *>(hits));
  return hipGetLastError();
}

}  // namespace phx
