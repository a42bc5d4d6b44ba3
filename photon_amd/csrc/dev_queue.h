// Wave-level queue helpers: one atomic per wave (ballot + lane rank).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace phx {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Slot in `counter`'s queue for every lane with `pred`; lanes without pred get
// an unspecified value.  Every active lane of the wave must call it.
__device__ __forceinline__ uint32_t wave_alloc(uint32_t *counter, bool pred) {
  const uint64_t mask = __ballot(pred);
  if (mask == 0) return 0;
  const uint32_t lane = lane_id();
  const uint32_t rank = __popcll(mask & ((uint64_t(1) << lane) - 1));
  const int leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if (int(lane) == leader) base = atomicAdd(counter, uint32_t(__popcll(mask)));
  base = __shfl(base, leader);
  return base + rank;
}

// Adds the wave's count of `pred` lanes to a 64-bit statistic, one atomic per wave.
__device__ __forceinline__ void wave_count(unsigned long long *counter, bool pred) {
  const uint64_t mask = __ballot(pred);
  if (mask == 0) return;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  if (int(lane_id()) == leader) atomicAdd(counter, (unsigned long long)__popcll(mask));
}

// Next batch of 64 work items for the whole wave (persistent kernels).
__device__ __forceinline__ uint32_t wave_fetch(uint32_t *next) {
  uint32_t base = 0;
  if (lane_id() == 0) base = atomicAdd(next, 64u);
  return __shfl(base, 0);
}

}  // namespace phx
