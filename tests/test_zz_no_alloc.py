"""phx_trace_closest_noalloc, the entry point that allocates nothing, against
phx_trace_closest on the same rays: every hit record byte for byte, on two
calls in a row, the second with fewer rays.  The 737,290-triangle sphere grid
(BVH depth 9) makes the traversal spill its stack into the static scratch."""
import ctypes as C
import functools

import numpy as np
import pytest

from photon_amd import abi, scenes
from photon_amd._lib import lib

pytestmark = pytest.mark.gpu


@functools.lru_cache(maxsize=None)
def _scene():
    return scenes.sphere_grid(seed=1234)


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


class _Rays:
    """Device copies of a ray batch and an output buffer, made up front."""

    def __init__(self, org, d, tmin, tmax):
        import torch
        self.n = org.shape[0]
        self.args = [_dev(x) for x in (org, d, tmin, tmax)]
        self.buf = torch.full((self.n * C.sizeof(abi.Hit),), 0xAB, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()

    def trace(self, ctx, fn):
        self.buf.fill_(0xAB)
        ctx.synchronize()
        import torch
        torch.cuda.synchronize()
        ctx.check(fn(ctx.h, *[C.c_void_p(a.data_ptr()) for a in self.args], self.n, C.c_void_p(self.buf.data_ptr()),
                     None))
        ctx.synchronize()
        return self.buf.cpu().numpy()


def test_noalloc_trace_equals_trace_closest():
    import torch
    from photon_amd.backend import Context, HipBackend
    ctx = Context(0)
    try:
        s = _scene()
        HipBackend(context=ctx).build_accel(s)
        for n, seed in ((1 << 18, 5), (1000, 6)):  # the second call has fewer rays
            org, d, tmin, tmax = scenes.random_rays(s, n, seed=seed)
            print(f"noalloc trace: {n} rays, seed {seed}", flush=True)
            rays = _Rays(org, d, tmin, tmax)
            free0 = torch.cuda.mem_get_info()[0]
            got = rays.trace(ctx, lib().phx_trace_closest_noalloc)
            free1 = torch.cuda.mem_get_info()[0]
            ref = rays.trace(ctx, lib().phx_trace_closest)
            hits = (abi.Hit * n).from_buffer_copy(got.tobytes())
            frac = np.mean([hits[i].hit for i in range(n)])
            print(f"  hit fraction {frac:.3f}, free device memory change {free0 - free1} bytes", flush=True)
            assert frac > 0.2
            assert np.array_equal(got, ref), "noalloc hit records differ from phx_trace_closest's"
            # the call itself allocates nothing (its kernels' code is loaded by the first call)
            if n < (1 << 18):
                assert free1 == free0
    finally:
        ctx.close()
