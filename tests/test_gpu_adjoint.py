"""The path-replay adjoint: the Disney derivatives against double central
differences, the full gradient (materials and texels) against FP64 central
differences of the oracle's detached estimator (the same sampling decisions,
weights in double), and the torch.autograd.Function harness."""
import ctypes as C

import numpy as np
import pytest

import oracle
from photon_amd import abi, scenes
from tests.helpers import random_dirs, random_material

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _dev(a, dtype=None):
    torch = _torch()
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def test_bsdf_derivatives_match_double_differences(gpu_ctx):
    torch = _torch()
    from photon_amd._lib import lib
    rng = np.random.default_rng(5)
    n = 4096
    mats = []
    for _ in range(n):
        m = random_material(rng)
        m.roughness = float(rng.uniform(0.05, 0.95))
        m.metallic = float(rng.uniform(0.05, 0.95))
        m.transmission = 0.0
        mats.append(m)
    ng = random_dirs(n, rng)
    sn = ng + 0.15 * rng.normal(size=(n, 3))
    sn /= np.linalg.norm(sn, axis=1, keepdims=True)
    wo = random_dirs(n, rng)
    wi = random_dirs(n, rng)
    wo[np.sum(wo * ng, 1) < 0] *= -1
    wi[np.sum(wi * ng, 1) < 0] *= -1
    arr = (abi.Material * n)(*mats)
    dm = torch.tensor(np.frombuffer(bytes(arr), np.uint8).copy(), device="cuda")
    tw = [_dev(x.astype(np.float32)) for x in (wo, wi, ng, sn)]
    f = torch.zeros(n * 3, device="cuda")
    df = torch.zeros(n * 15, device="cuda")
    p = [C.c_void_p(t.data_ptr()) for t in [dm] + tw + [f, df]]
    assert lib().phx_bsdf_eval_grad(gpu_ctx.h, p[0], p[1], p[2], p[3], p[4], n, p[5], p[6], None) == 0
    gpu_ctx.synchronize()
    g = df.cpu().numpy().reshape(n, 5, 3)
    ref = oracle.bsdf_eval_grad_fd(mats, wo, wi, ng, sn, h=1e-4)
    scale = np.abs(ref).max(axis=(1, 2), keepdims=True) + 1e-6
    err = np.abs(g - ref) / scale
    print(f"bsdf derivative: max scaled error {err.max():.2e}, 99.9th pct {np.percentile(err, 99.9):.2e}", flush=True)
    assert np.percentile(err, 99.9) < 1e-3
    assert np.mean(err < 1e-4) > 0.99


def _c3_small(tex_size=32):
    tex = scenes.colored_noise_texture(tex_size, 11)
    s = scenes.textured_cornell(n_materials=16, tex_size=tex_size, texture=tex, seed=3)
    return s


def test_adjoint_matches_fp64_central_differences(gpu_ctx):
    torch = _torch()
    from photon_amd.backend import make_camera
    from photon_amd.diff import DiffRenderer
    s = _c3_small()
    w = h = 48
    cam = make_camera(s.camera["look_from"], s.camera["look_at"], s.camera["vup"], s.camera["vfov"], 1.0)
    dr = DiffRenderer(s, cam, w, h, spp=4, max_depth=6, adjoint_spp=4, n_train=16)
    rng = np.random.default_rng(21)
    delta = rng.normal(size=(h, w, 3)).astype(np.float32) * 1e-2
    gp, gt = dr.adjoint(torch.tensor(delta, device="cuda"))
    gp = gp.cpu().numpy().astype(np.float64)
    gt = gt.cpu().numpy().astype(np.float64)
    o = oracle.OracleScene(s)
    mp = np.array([[m.base_color[0], m.base_color[1], m.base_color[2], m.roughness, m.metallic]
                   for m in s.materials], np.float64)
    tex0 = np.asarray(s.texture_pixels, np.float64).reshape(-1, 3)
    hstep = 1e-4
    # 48 material parameters (the largest gradients of materials 1..15) and
    # 16 texels with non-zero gradient
    # parameters at least 1e-3 inside (0, 1), where the resolve clamps them
    cand = [(m, k) for m in range(1, 16) for k in range(5) if 1e-3 < mp[m, k] < 1 - 1e-3]
    cand.sort(key=lambda mk: -abs(gp[mk]))
    picks = [("mat", mk) for mk in cand[:48]]
    tex_nz = np.argwhere(np.abs(gt.reshape(-1, 3)) > 0)
    sel = tex_nz[rng.choice(len(tex_nz), 16, replace=False)]
    picks += [("tex", (int(a), int(b))) for a, b in sel]
    bad = []
    gmax = max(np.abs(gp).max(), np.abs(gt).max())
    for kind, key in picks:
        if kind == "mat":
            plus, minus = mp.copy(), mp.copy()
            plus[key] += hstep
            minus[key] -= hstep
            lp, _ = o.render_detached(cam, dr.adj_params, mat_params=plus, delta=delta)
            lm, _ = o.render_detached(cam, dr.adj_params, mat_params=minus, delta=delta)
            g = gp[key]
        else:
            plus, minus = tex0.copy(), tex0.copy()
            plus[key] += hstep
            minus[key] -= hstep
            lp, _ = o.render_detached(cam, dr.adj_params, texels=plus, delta=delta)
            lm, _ = o.render_detached(cam, dr.adj_params, texels=minus, delta=delta)
            g = gt.reshape(-1, 3)[key]
        fd = (lp - lm) / (2 * hstep)
        ok = abs(g - fd) <= 1e-3 * abs(fd) + 1e-7 * gmax
        if not ok:
            bad.append((kind, key, g, fd))
    print(f"adjoint vs FP64 central differences: {len(picks) - len(bad)} of {len(picks)} within 1e-3", flush=True)
    for b in bad[:10]:
        print("  mismatch", b, flush=True)
    assert not bad


def test_torch_autograd_function_descends():
    torch = _torch()
    from photon_amd.backend import make_camera
    from photon_amd.diff import DiffRenderer
    target_scene = _c3_small(16)
    start_scene = scenes.textured_cornell(n_materials=16, tex_size=16, seed=4)
    cam = make_camera(start_scene.camera["look_from"], start_scene.camera["look_at"], start_scene.camera["vup"],
                      start_scene.camera["vfov"], 1.0)
    tgt = DiffRenderer(target_scene, cam, 32, 32, spp=16, max_depth=6, n_train=16)
    p_t, tex_t = tgt.initial_params(target_scene)
    target = tgt.render(p_t, tex_t).detach()
    dr = DiffRenderer(start_scene, cam, 32, 32, spp=16, max_depth=6, adjoint_spp=16, n_train=16)
    params, tex = dr.initial_params(start_scene)
    params.requires_grad_(True)
    tex.requires_grad_(True)
    img = dr.render(params, tex)
    loss0 = ((img - target) ** 2).mean()
    loss0.backward()
    assert torch.isfinite(params.grad).all() and torch.isfinite(tex.grad).all()
    assert params.grad.abs().sum() > 0 and tex.grad.abs().sum() > 0
    # the harness passes dLoss/dImage to the adjoint unchanged
    direct_p, direct_t = dr.adjoint(2 * (img.detach() - target) / img.numel())
    torch.testing.assert_close(params.grad, direct_p, rtol=1e-6, atol=1e-9)
    with torch.no_grad():
        params -= 0.5 * params.grad / params.grad.abs().max() * 0.02
        tex -= 0.5 * tex.grad / tex.grad.abs().max() * 0.05
        params.clamp_(0.01, 0.99)
        tex.clamp_(0.0, 1.0)
    loss1 = ((dr.render(params, tex) - target) ** 2).mean()
    print(f"loss {loss0.item():.5f} -> {loss1.item():.5f}", flush=True)
    assert loss1 < loss0
