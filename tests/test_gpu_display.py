"""phx_display_transform on the GPU against the oracle's restatement of the
reference's tone maps and transfer curves (tone_mapping.h:7-28, color.h:9-12,
photon_render.cpp:24-35), through the C ABI.  Division and products round
as the reference's (IEEE, no contraction), so the tone maps are bit-exact;
the curves call powf, which the GPU's and the host's libm may round an ulp or
two apart, and the sRGB curve's 1.055 p - 0.055 magnifies that near its knee:
floats within 1e-6 relative there, 8-bit values equal but for one-step
differences where a value sits on a rounding boundary."""
import itertools

import numpy as np
import pytest

import oracle
from photon_amd import abi

pytestmark = pytest.mark.gpu


def _hdr(n, seed):
    rng = np.random.default_rng(seed)
    x = np.exp(rng.normal(-1.0, 2.0, (n, 4))).astype(np.float32)
    x[::7] *= -1.0          # negative radiance
    x[::11] = 0.0
    x[::13, 0] = 1e30       # huge
    x[::17, 1] = 0.0031308  # the sRGB knee
    return x


@pytest.mark.parametrize("tonemap,curve", list(itertools.product(
    [abi.PHX_TONEMAP_NONE, abi.PHX_TONEMAP_ACES, abi.PHX_TONEMAP_REINHARD],
    [abi.PHX_CURVE_LINEAR, abi.PHX_CURVE_SRGB, abi.PHX_CURVE_GAMMA22])))
def test_display_transform_matches_oracle(gpu_ctx, tonemap, curve):
    import torch
    from photon_amd.backend import display_transform
    x = _hdr(100003, tonemap * 3 + curve)
    for channels, exposure, clamp in ((4, 0.0, 1), (3, 1.5, 0), (3, -2.0, 1)):
        d = abi.display(tonemap=tonemap, curve=curve, exposure=exposure, clamp=clamp)
        src = x[:, :channels].copy()
        g, g8 = display_transform(gpu_ctx, torch.from_numpy(src).cuda(), d)
        g, g8 = g.cpu().numpy(), g8.cpu().numpy()
        r, r8 = oracle.display(src, d)
        both_nan = np.isnan(g) & np.isnan(r)
        tol = 1e-6 * np.abs(r) + 1e-30
        ok = both_nan | (np.abs(g - r) <= tol) | (g == r)
        if curve == abi.PHX_CURVE_LINEAR:
            ok = both_nan | (g == r)  # no powf: bit-exact
        bad = np.argwhere(~ok)[:5]
        assert ok.all(), (channels, exposure, clamp, [(tuple(b), g[tuple(b)], r[tuple(b)]) for b in bad])
        diff8 = np.abs(g8.astype(int) - r8.astype(int))
        assert diff8.max() <= 1 and np.mean(diff8 > 0) < 1e-3


def test_last_frame_display(gpu_ctx):
    from photon_amd import scenes
    from photon_amd.backend import HipBackend, PathTracer, make_camera
    s = scenes.cornell_box()
    pt = PathTracer(HipBackend(context=gpu_ctx))
    pt.set_scene(s)
    pt.camera = make_camera(s.camera["look_from"], s.camera["look_at"], s.camera["vup"], s.camera["vfov"], 1.0)
    pt.params = abi.default_render_params(64, 48, 4, 6)
    res = pt.render()
    g, g8 = pt.display()
    r, r8 = oracle.display(res.color, abi.display())
    assert np.allclose(g.cpu().numpy(), r, rtol=1e-6, atol=0)
    assert np.abs(g8.cpu().numpy().astype(int) - r8.astype(int)).max() <= 1
