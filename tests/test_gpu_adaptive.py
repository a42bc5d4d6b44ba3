"""GPU parity, through the C ABI, of adaptive sampling: phx_render with
adaptive_sampling = 1 (pathtracer.cpp:676-723, :1129-1139).  A pixel with 16
or more samples stops before sample s >= 16 once its largest per-channel
sample variance over its squared luminance is below 0.01.  Checked against
the oracle's per-pixel sample counts and image on the same sample streams,
and against plain GPU renders: a pixel that stopped after n samples equals,
bit for bit, the plain render of n samples (the same samples summed in the
same order).  The later samples run in groups of sample indices with the
rule applied in order inside a group: groups of 1 (the reference's barrier
per sample), 8 and all at once (the default) give the same counts.  One
case splits the first 16 samples over several runs (a small run budget) so
the running sums carry across runs.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from photon_amd import abi, scenes
from photon_amd._lib import PhxError, lib
from tests.helpers import rel_mse

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("depth,budget_spp,group", [(1, 0, 0), (1, 0, 1), (1, 0, 8), (4, 5, 0)])
def test_adaptive_render_matches_oracle_and_plain_renders(gpu_ctx, monkeypatch, depth, budget_spp, group):
    from photon_amd.backend import HipBackend, PathTracer, make_camera
    w, h, spp = 48, 48, 48
    if budget_spp:
        monkeypatch.setenv("PHX_SAMPLE_BUDGET", str(w * h * budget_spp))
    if group:
        monkeypatch.setenv("PHX_ADAPTIVE_GROUP", str(group))
    # the Cornell box through a 70 degree field of view: the room fills the
    # middle, the rest sees the black background (zero variance)
    s = scenes.cornell_box()
    c = s.camera
    cam = make_camera(c["look_from"], c["look_at"], c["vup"], 70.0, w / h)
    pt = PathTracer(HipBackend(context=gpu_ctx))
    pt.set_scene(s)
    pt.camera = cam
    p = abi.default_render_params(w, h, spp, depth)
    p.adaptive_sampling = 1
    pt.params = p
    print(f"adaptive cornell {w}x{h} {spp} spp depth {depth} budget {budget_spp} group {group}", flush=True)
    res = pt.render()
    ns = pt.sample_counts().astype(np.int64)
    ref, ex, _ = oracle.OracleScene(s).render(cam, p, aovs=True, nsamp=True)
    ons = ex["nsamp"].astype(np.int64)
    agree = float(np.mean(ns == ons))
    err = rel_mse(res.color, ref)
    m = ns == ons
    same = float(np.mean(np.abs(res.color[m] - ref[m]) <= 1e-4 * (1 + np.abs(ref[m]))))
    hist = dict(zip(*[a.tolist() for a in np.unique(ns, return_counts=True)]))
    print(f"counts equal {agree:.4f}, relMSE {err:.3e}, pixels equal to 1e-4 {same:.4f}, counts {hist}", flush=True)
    assert ns.min() == 16 and ns.max() <= spp
    assert np.all(ns[ex["depth"] >= 1e29] == 16)
    assert (ns == spp).any()
    if depth == 1:
        assert ((ns > 16) & (ns < spp)).any()  # stops inside the groups
    assert agree > 0.98
    assert err < 1e-3
    assert same > 0.9

    for n in np.unique(ns):
        pt.params = abi.default_render_params(w, h, int(n), depth)
        plain = pt.render()
        sel = ns == n
        assert np.array_equal(res.color[sel], plain.color[sel]), n
        assert np.array_equal(res.alpha[sel], plain.alpha[sel]), n
    with pytest.raises(PhxError, match="not adaptively sampled"):
        pt.sample_counts()


def test_adaptive_is_offered_by_phx_render_only(gpu_ctx):
    # phx_render_batch and phx_adjoint refuse it (the batch shares its runs
    # across views; the adjoint's replay needs fixed sample counts)
    from photon_amd.backend import BatchRenderer, Context
    from photon_amd.backend import HipBackend, PathTracer, make_camera
    ctx = Context(0)
    s = scenes.cornell_box()
    pt = PathTracer(HipBackend(context=ctx))
    pt.set_scene(s)
    c = s.camera
    cam = make_camera(c["look_from"], c["look_at"], c["vup"], 40.0, 1.0)
    p = abi.default_render_params(8, 8, 4, 2)
    p.adaptive_sampling = 1
    br = BatchRenderer(ctx)
    with pytest.raises(PhxError, match="adaptive"):
        br.render([(ctx, cam, p)])
    # phx_adjoint refuses before it reads or writes the buffers
    g = torch.zeros(8 * 8 * 3, dtype=torch.float32, device="cuda:0")
    gm = torch.zeros(1024, dtype=torch.float32, device="cuda:0")
    with pytest.raises(PhxError, match="adaptive"):
        ctx.check(lib().phx_adjoint(ctx.h, C.byref(cam), C.byref(p), C.c_void_p(g.data_ptr()),
                                    C.c_void_p(gm.data_ptr()), None))
    ctx.close()


def test_adaptive_fog_box_matches_oracle_and_plain_renders(gpu_ctx):
    # a volume scene: the stopped pixels' empty queue entries also pass
    # through the volume-event pass (their empty interval skips the tracking)
    from photon_amd.backend import HipBackend, PathTracer, make_camera
    w, h, spp, depth = 32, 32, 40, 4
    s = scenes.fog_box(n=32, macro=8, env_size=(64, 32))
    c = s.camera
    cam = make_camera(c["look_from"], c["look_at"], c["vup"], 70.0, w / h)
    pt = PathTracer(HipBackend(context=gpu_ctx))
    pt.set_scene(s)
    pt.camera = cam
    p = abi.default_render_params(w, h, spp, depth)
    p.adaptive_sampling = 1
    pt.params = p
    print(f"adaptive fog box {w}x{h} {spp} spp depth {depth}", flush=True)
    res = pt.render()
    ns = pt.sample_counts().astype(np.int64)
    ref, ex, _ = oracle.OracleScene(s).render(cam, p, nsamp=True)
    agree = float(np.mean(ns == ex["nsamp"]))
    err = rel_mse(res.color, ref)
    hist = dict(zip(*[a.tolist() for a in np.unique(ns, return_counts=True)]))
    print(f"counts equal {agree:.4f}, relMSE {err:.3e}, counts {hist}", flush=True)
    assert ns.min() == 16 and ns.max() <= spp
    assert (ns == 16).mean() > 0.2 and (ns == spp).mean() > 0.2
    assert agree > 0.98
    assert err < 1e-3
    for n in np.unique(ns):
        pt.params = abi.default_render_params(w, h, int(n), depth)
        plain = pt.render()
        sel = ns == n
        assert np.array_equal(res.color[sel], plain.color[sel]), n
