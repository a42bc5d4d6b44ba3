"""Display transforms: the oracle's restatement of tone_mapping.h /
color.h / photon_render's postprocess pinned by closed forms, and the
library's host-only denoiser seam.  The GPU transform is compared with the
oracle in test_gpu_display.py."""
import ctypes as C

import numpy as np
import pytest

import oracle
from photon_amd import abi


def _one(x, **kw):
    out, out8 = oracle.display(np.array([[x, x, x]], np.float32), abi.display(**kw))
    return float(out[0, 0]), int(out8[0, 0])


RAW = dict(tonemap=abi.PHX_TONEMAP_NONE, curve=abi.PHX_CURVE_LINEAR, clamp=0)


def test_aces_known_values():
    aces = dict(RAW, tonemap=abi.PHX_TONEMAP_ACES)
    assert _one(0.0, **aces)[0] == 0.0                      # 0 / 0.14
    assert abs(_one(1.0, **aces)[0] - 2.54 / 3.16) < 1e-6   # (2.51 + 0.03) / (2.43 + 0.59 + 0.14)
    assert abs(_one(1e6, **aces)[0] - 2.51 / 2.43) < 1e-5   # the curve's asymptote
    # the raw curve crosses 1 at the root of 0.08 x^2 - 0.56 x - 0.14, x =
    # 7.2426...: the reference's tonemap_test (10, 2, 0.5) lies in [0, 1] only
    # after photon_render's clamp01
    x1 = (0.56 + np.sqrt(0.56 ** 2 + 4 * 0.08 * 0.14)) / (2 * 0.08)
    assert _one(x1 * 0.999, **aces)[0] < 1.0 < _one(x1 * 1.001, **aces)[0]
    assert _one(10.0, **aces)[0] > 1.0


def test_reinhard_exposure_and_curves():
    rh = dict(RAW, tonemap=abi.PHX_TONEMAP_REINHARD)
    assert _one(1.0, **rh)[0] == 0.5
    assert _one(3.0, **rh)[0] == 0.75
    assert _one(0.75, **dict(RAW, exposure=1.0))[0] == 1.5   # c * 2^1, exactly
    assert _one(0.75, **dict(RAW, exposure=-2.0))[0] == 0.1875
    srgb = dict(RAW, curve=abi.PHX_CURVE_SRGB)
    assert _one(0.0, **srgb)[0] == 0.0
    assert abs(_one(1.0, **srgb)[0] - 1.0) < 1e-6
    assert abs(_one(0.5, **srgb)[0] - (1.055 * 0.5 ** (1 / 2.4) - 0.055)) < 1e-6
    knee = 0.0031308
    lo = _one(knee, **srgb)[0]
    assert abs(lo - 12.92 * knee) < 1e-8
    # the two branches meet at the knee (to the constants' precision)
    assert abs(lo - (1.055 * knee ** (1 / 2.4) - 0.055)) < 1e-4
    g22 = dict(RAW, curve=abi.PHX_CURVE_GAMMA22)
    assert abs(_one(0.5, **g22)[0] - 0.5 ** (1 / 2.2)) < 1e-6
    # linear_to_srgb then srgb_to_linear (pow 2.2) returns the value (math_test.cpp:116-119)
    for x in (0.25, 0.5, 0.75):
        assert abs(_one(x, **g22)[0] ** 2.2 - x) < 2e-4


def test_postprocess_and_u8():
    # photon_render's postprocess: ACES, clamp, sRGB curve, clamp; to_u8
    hdr = np.array([[10.0, 2.0, 0.5], [-1.0, 0.0, 1e9]], np.float32)
    out, out8 = oracle.display(hdr, abi.display())
    assert np.all(np.isfinite(out)) and np.all((out >= 0) & (out <= 1))
    # saturated channels: the sRGB curve of 1 is 1.055f - 0.055f = 1 - 2^-24 in float
    assert out[0, 0] == out[1, 2] == np.float32(1.055) * np.float32(1.0) - np.float32(0.055)
    # ACES of -1 is 2.48 / 1.98 > 1 (the curve's numerator and denominator
    # both change sign), so -1 saturates too; 0 stays 0
    assert abs(_one(-1.0, **dict(RAW, tonemap=abi.PHX_TONEMAP_ACES))[0] - 2.48 / 1.98) < 1e-6
    assert out[1, 0] == out[0, 0] and out[1, 1] == 0.0 and out8[0, 0] == 255
    assert list(out8[0]) == [int(out[0, k] * 255.999) for k in range(3)]
    assert _one(0.5, **RAW)[1] == 127 and _one(1.0, **RAW)[1] == 255
    assert _one(-3.0, **RAW)[1] == 0 and _one(7.0, **RAW)[1] == 255
    # four input channels: the fourth is skipped
    rgba = np.array([[0.2, 0.4, 0.6, 9.0]], np.float32)
    np.testing.assert_array_equal(oracle.display(rgba, abi.display())[0], oracle.display(rgba[:, :3], abi.display())[0])


def test_denoiser_seam_without_oidn():
    from photon_amd._lib import lib
    assert lib().phx_denoiser_is_available() == 0
    c = np.random.default_rng(0).random((4, 5, 3)).astype(np.float32)
    keep = c.copy()
    a = np.ones_like(c)
    n = np.zeros_like(c)
    p = lambda x: C.c_void_p(x.ctypes.data)
    assert lib().phx_denoise_buffer(p(c), p(a), p(n), 5, 4) == 0
    np.testing.assert_array_equal(c, keep)
    assert lib().phx_denoise_buffer(None, None, None, 5, 4) != 0


def test_display_rejects_bad_arguments_without_a_gpu():
    from photon_amd._lib import lib
    d = abi.display()
    assert lib().phx_display_transform(None, None, 3, 0, C.byref(d), None, None, None) != 0
