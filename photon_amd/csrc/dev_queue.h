// Queue helpers of the wavefront kernels: work-group slot claims, statistics
// sums and the striped work fetch of the persistent kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace phx {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Queue slots for a work-group of kWaves waves, in two queues counted side by
// side in the halves of one 64-bit counter: every thread with pred_lo gets a
// slot of the low queue, every thread with pred_hi one of the high queue (the
// others get unspecified values), with one atomic per work-group.  A half
// never carries into the other while it stays below 2^32.  All waves of the
// GPU append to the same counter, and device-scope atomics on one address are
// serialised: at one atomic per wave the shade kernel spent most of its time
// queued on it.  Every thread of the work-group must call this (it
// synchronises the work-group).
template <int kWaves>
__device__ __forceinline__ void block_alloc_pair(unsigned long long *counter, bool pred_lo, uint32_t &slot_lo,
                                                 bool pred_hi, uint32_t &slot_hi) {
  __shared__ uint32_t s_n[2][kWaves];
  __shared__ unsigned long long s_base;
  const uint64_t ml = __ballot(pred_lo), mh = __ballot(pred_hi);
  const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
  if (lane == 0) {
    s_n[0][wave] = uint32_t(__popcll(ml));
    s_n[1][wave] = uint32_t(__popcll(mh));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t nl = 0, nh = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      nl += s_n[0][w];
      nh += s_n[1][w];
    }
    s_base = (nl | nh) ? atomicAdd(counter, (unsigned long long)nl | ((unsigned long long)nh << 32)) : 0ull;
  }
  __syncthreads();
  const unsigned long long base = s_base;
  uint32_t ol = 0, oh = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    ol += uint32_t(w) < wave ? s_n[0][w] : 0u;
    oh += uint32_t(w) < wave ? s_n[1][w] : 0u;
  }
  const uint64_t below = (uint64_t(1) << lane) - 1;
  slot_lo = uint32_t(base) + ol + uint32_t(__popcll(ml & below));
  slot_hi = uint32_t(base >> 32) + oh + uint32_t(__popcll(mh & below));
}

// Adds the sum of `n` over the work-group's threads to a 64-bit statistic,
// one atomic per work-group.  Every thread of the work-group must call it (it
// synchronises the work-group).
template <int kWaves>
__device__ __forceinline__ void block_sum(unsigned long long *counter, uint32_t n) {
  __shared__ uint32_t s_sum[kWaves];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off);
  if (lane_id() == 0) s_sum[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) t += s_sum[w];
    if (t) atomicAdd(counter, (unsigned long long)t);
  }
}

constexpr uint32_t kChunk = 256;  // chunk of the seam-A traversal kernels and the resolve pass

// Striped work fetch of the persistent traversal kernels.  The queue is cut
// into chunks of `chunk` items, dealt round-robin to kStripes stripes (chunk
// k belongs to stripe k mod kStripes), so every stripe holds an even sample
// of the queue: the first iteration's queue is in pixel order, and stripes of
// contiguous ranges differed in cost by whole image bands.  Each stripe has
// its own chunk counter in a 128-byte line of its own (a fetch set of
// kFetchSetWords words), so the fetchers spread their atomics over kStripes
// addresses instead of queueing on one.  A fetcher (a wave, or a work-group's
// leader) starts in its home stripe (its index mod kStripes) and moves on to
// the next stripe when that one is used up.
//
// Every move costs a failed atomic, and device-scope atomics on one address
// are serialised: with all 8192 waves of a persistent grid walking all
// kStripes stripes, a launch over a short queue spent ~80 us in the walk
// alone.  Three things keep the walk short:
//  - only the first fetch_units(count, chunk) fetchers take part, enough to
//    give each a chunk;
//  - a fetcher moves past at most fetch_walk(fetchers) stripes in total,
//    where fetchers = min(units, fetchers in the grid).  Once kStripes or
//    more fetchers take part, every stripe is some fetcher's home, and a
//    fetcher leaves its home stripe only when it is used up, so the home
//    fetchers drain every stripe; the walk only balances the tail.  With
//    fewer fetchers the walk covers every stripe;
//  - the waves of a work-group share the stripes they found used up in LDS.
constexpr uint32_t kStripes = 16;
constexpr uint32_t kStripeStride = 32;  // words between the counters of a set
constexpr uint32_t kFetchSetWords = kStripes * kStripeStride;
constexpr uint32_t kFetchWalk = 4;  // stripes a fetcher may try when every stripe has a home fetcher
__device__ __forceinline__ uint32_t fetch_units(uint32_t count, uint32_t chunk) { return (count + chunk - 1u) / chunk; }
__device__ __forceinline__ uint32_t fetch_walk(uint32_t fetchers) {
  return fetchers >= kStripes ? kFetchWalk : kStripes;
}
// Chunks of the queue's `units` that belong to `stripe`.
__device__ __forceinline__ uint32_t stripe_units(uint32_t units, uint32_t stripe) {
  return units > stripe ? (units - stripe + kStripes - 1u) / kStripes : 0u;
}
// The wave's next chunk [lo, hi) of the queue's `count` items (`units`
// chunks), from its current stripe or the ones after it, trying at most
// `left` more stripes (decremented for each one found used up); lo =
// 0xffffffff when none is left.  `used` is the work-group's LDS mask of
// stripes known to be used up (zero at the start of the fetch): a wave skips
// those without an atomic and marks the ones it finds used up, so the waves of
// a work-group pay one failed atomic per stripe between them.  Wave-uniform:
// every lane of the wave must call it.
__device__ __forceinline__ void fetch_chunk(uint32_t *set, uint32_t *used, uint32_t count, uint32_t units,
                                            uint32_t chunk, uint32_t &stripe, uint32_t &left, uint32_t &lo,
                                            uint32_t &hi) {
  lo = 0xffffffffu;
  hi = 0u;
  for (; left; --left) {
    const uint32_t bit = 1u << stripe;
    if (!(__hip_atomic_load(used, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) & bit)) {
      if (stripe_units(units, stripe)) {
        uint32_t k = 0;
        if (lane_id() == 0) k = atomicAdd(&set[stripe * kStripeStride], 1u);
        k = __shfl(k, 0);
        if (k < stripe_units(units, stripe)) {
          lo = (k * kStripes + stripe) * chunk;
          hi = min(lo + chunk, count);
          return;
        }
      }
      if (lane_id() == 0) __hip_atomic_fetch_or(used, bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    stripe = stripe + 1u == kStripes ? 0u : stripe + 1u;
  }
}
// fetch_chunk for one thread (a work-group's leader fetching the group's
// tiles of n items; no LDS mask, as there is one fetcher per work-group).
__device__ __forceinline__ void fetch_tile(uint32_t *set, uint32_t count, uint32_t units, uint32_t n,
                                           uint32_t &stripe, uint32_t &left, uint32_t &lo, uint32_t &hi) {
  lo = 0xffffffffu;
  hi = 0u;
  for (; left; --left) {
    const uint32_t su = stripe_units(units, stripe);
    if (su) {
      const uint32_t k = atomicAdd(&set[stripe * kStripeStride], 1u);
      if (k < su) {
        lo = (k * kStripes + stripe) * n;
        hi = min(lo + n, count);
        return;
      }
    }
    stripe = stripe + 1u == kStripes ? 0u : stripe + 1u;
  }
}

// Work distribution of the resolve pass: each wave owns a chunk of `chunk`
// items (a multiple of 64) at a time.  The first chunk of every wave is static
// (wave index * chunk); later chunks come from one atomic per wave and chunk on
// a counter that starts after the static ones.  (A plain atomic at L2: a
// coherent pre-read of the counter would bypass L2 and serialize at memory.)
struct WorkCursor {
  uint32_t base;  // first item of the current chunk
  uint32_t sub;   // next 64-item sub-batch within the chunk
};
__device__ __forceinline__ WorkCursor first_chunk(uint32_t chunk) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  return WorkCursor{wave * chunk, 0u};
}
// Base of the wave's next 64-item batch; >= count when the work is done
// (every later batch is then beyond count too).  The counter starts at 0.
__device__ __forceinline__ uint32_t next_batch(WorkCursor &c, uint32_t *counter, uint32_t count, uint32_t chunk) {
  if (c.sub == chunk) {
    uint32_t b = 0;
    if (lane_id() == 0) b = atomicAdd(counter, chunk);
    b = __shfl(b, 0);
    c.base = gridDim.x * blockDim.x / 64u * chunk + b;
    c.sub = 0;
  }
  const uint32_t r = c.base + c.sub;
  c.sub += 64u;
  return r;
}

}  // namespace phx
