// phx_trace_closest_noalloc: phx_trace_closest with no device allocation, for
// callers that cannot allocate (code under HIP graph capture, tight serving
// loops).  The traversal kernels' scratch - the chunk counters of the
// persistent fetch and the per-thread stack spill - lives in `scratch`, a
// static device array of kScratchBytes that each GPU holds for the life of
// the process.  Its device address is looked up once per device
// (phx_noalloc_setup, called by phx_create with the device current) and only
// that address is passed to kernels.
//
// Every call on one GPU shares that GPU's array: calls whose kernels overlap
// on one GPU (two streams, two host threads, two contexts) race on it.  The
// counters are cleared by a kernel at the start of each call, so a call never
// depends on what the previous call left.
#include <hip/hip_runtime.h>

#include <mutex>

#include "context.h"

namespace {

using namespace phx;
using namespace phx_detail;

// Layout: the fetch counters (kFetchSetWords words) in the first 4 KB, then
// the stack spill of every thread of the grid.
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr size_t kSpillOffset = 4096;
static_assert(kFetchSetWords * 4 <= kSpillOffset, "fetch counters fit before the spill area");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void *g_scratch[kMaxDevices] = {};  // the array's device address on each GPU

void *scratch_of(int device) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
}

}  // namespace

namespace phx_detail {

// Looks up the device address of `scratch` on `device` (which must be
// current), once per device.
int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return PHX_ERR_INVALID;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_scratch[device]) return PHX_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return PHX_ERR_HIP;
  g_scratch[device] = p;
  return PHX_OK;
}

}  // namespace phx_detail

extern "C" int phx_trace_closest_noalloc(phx_context *c, const float *o, const float *d, const float *tmin,
                                         const float *tmax, uint32_t n, phx_hit *out, void *stream) {
  if (!c || (n && (!o || !d || !tmin || !tmax || !out))) return PHX_ERR_INVALID;
  if (!c->have_accel) return fail(c, PHX_ERR_NO_SCENE, "trace before build_accel");
  if (n == 0) return PHX_OK;
  unsigned char *base = static_cast<unsigned char *>(scratch_of(c->device));
  if (!base) return fail(c, PHX_ERR_HIP, "trace_closest_noalloc: the static scratch of this device is not set up");
  // Stack entries beyond the LDS part that a tree of this depth can push
  // (at most kNodeWidth - 1 per level), per thread.
  const int need = (kNodeWidth - 1) * int(c->bvh_depth) + 1 - kShortStack;
  const size_t stride = need > 0 ? size_t(need) : 0;
  const size_t per_group = stride * sizeof(int) * 256;
  const size_t room = kScratchBytes - kSpillOffset;
  // The persistent grid of phx_trace_closest (8 work-groups per CU), or as
  // many work-groups as the scratch holds spill for.
  size_t grid = size_t(c->cu_count) * 8;
  if (per_group) grid = std::min(grid, room / per_group);
  if (grid < size_t(c->cu_count))
    return fail(c, PHX_ERR_OVERFLOW,
                "trace_closest_noalloc: a BVH of depth " + std::to_string(c->bvh_depth) + " needs " +
                    std::to_string(per_group * size_t(c->cu_count) + kSpillOffset) +
                    " bytes of traversal scratch for one work-group per CU; the static scratch holds " +
                    std::to_string(kScratchBytes));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  uint32_t *fetch = reinterpret_cast<uint32_t *>(base);
  int *spill = reinterpret_cast<int *>(base + kSpillOffset);
  HIP_TRY(c, launch_clear_u32(fetch, kFetchSetWords, st));
  HIP_TRY(c, launch_trace_full(c->dev, o, d, tmin, tmax, n, out, fetch, spill, uint32_t(grid), st, uint32_t(stride)));
  return PHX_OK;
}
