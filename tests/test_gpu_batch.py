"""phx_render_batch: many views of several scenes in one wavefront.  Each view's
result must equal phx_render of that view bit for bit (same sample streams,
same summation order), with or without the captured graph, however the views
are split into runs, and whatever the launch shape."""
import os

import numpy as np
import pytest

from photon_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _tracer(ctx_factory, s):
    from photon_amd.backend import HipBackend, PathTracer
    pt = PathTracer(HipBackend(context=ctx_factory()))
    pt.set_scene(s)
    return pt


def _view(pt, s, w, h, spp, depth, offset=0, orbit=0.0):
    from photon_amd.backend import make_camera
    eye = np.asarray(s.camera["look_from"], np.float64)
    at = np.asarray(s.camera["look_at"], np.float64)
    d = eye - at
    c, sn = np.cos(orbit), np.sin(orbit)
    eye = at + np.array([c * d[0] + sn * d[2], d[1], -sn * d[0] + c * d[2]])
    cam = make_camera(eye, at, s.camera["vup"], s.camera["vfov"], w / h)
    p = abi.default_render_params(w, h, spp, depth)
    p.sample_offset = offset
    return (pt, cam, p)


@pytest.fixture(scope="module")
def batch_setup():
    from photon_amd.backend import Context
    ctxs = []

    def make():
        ctxs.append(Context(0))
        return ctxs[-1]
    a = scenes.cornell_box()
    b = scenes.cornell_box(wall_colors=((0.2, 0.6, 0.7), (0.7, 0.6, 0.1), (0.5, 0.5, 0.5)))
    c = scenes.sphere_grid(seed=7, grid=2, level=2, env_size=(64, 32))
    pa, pb, pc = _tracer(make, a), _tracer(make, b), _tracer(make, c)
    views = [
        _view(pa, a, 32, 24, 4, 8),
        _view(pc, c, 40, 40, 3, 6, offset=5, orbit=0.7),
        _view(pb, b, 16, 16, 8, 8),
        _view(pc, c, 24, 8, 1, 8, orbit=-1.1),
        _view(pa, a, 8, 8, 16, 4, offset=3, orbit=0.3),
        _view(pb, b, 33, 17, 2, 8, orbit=-0.4),
    ]
    plain = []
    for pt, cam, p in views:
        pt.camera = cam
        pt.params = p
        plain.append(pt.render())
    yield views, plain, make


def _check_equal(res, ref, tag):
    for f in ("color", "alpha", "depth", "normal", "albedo"):
        np.testing.assert_array_equal(getattr(res, f), getattr(ref, f), err_msg=f"{tag}: {f}")


@pytest.mark.parametrize("use_graph", [1, 0])
def test_batch_equals_plain_renders(batch_setup, use_graph):
    from photon_amd.backend import BatchRenderer
    views, plain, make = batch_setup
    br = BatchRenderer(make())
    br.ctx.set_tuning(use_graph=use_graph)
    for rep in range(2):  # the second call replays the captured graph
        br.render(views)
        for v in range(len(views)):
            _check_equal(br.read(v), plain[v], f"graph={use_graph} rep={rep} view {v}")
    st = br.ctx.stats()
    assert st.samples == sum(p.width * p.height * p.samples_per_pixel for _, _, p in views)


@pytest.mark.parametrize("overlap", ["1", "0"])
def test_batch_split_into_runs(batch_setup, monkeypatch, overlap):
    # a small run budget splits views across runs and spp ranges; the runs
    # stream on two lanes, each starting beside the previous one's drain
    # (PHX_OVERLAP=0: one lane), and reduce in run order
    from photon_amd.backend import BatchRenderer
    views, plain, make = batch_setup
    monkeypatch.setenv("PHX_SAMPLE_BUDGET", "3000")
    monkeypatch.setenv("PHX_OVERLAP", overlap)
    br = BatchRenderer(make())
    for rep in range(2):  # the second call reuses both lanes' buffers and graphs
        br.render(views)
        for v in range(len(views)):
            _check_equal(br.read(v), plain[v], f"split overlap={overlap} rep={rep} view {v}")
    st = br.ctx.stats()
    assert st.samples == sum(p.width * p.height * p.samples_per_pixel for _, _, p in views)
    # the plain path splits the same way and still gives the same frame
    pt, cam, p = views[1]
    pt.camera, pt.params = cam, p
    _check_equal(pt.render(), plain[1], "plain split")


def test_batch_of_many_small_views(batch_setup):
    # 40 views of three scenes: item indices well past the first few travel in
    # the queue entries, and each view still equals its own render
    from photon_amd.backend import BatchRenderer
    views, plain, make = batch_setup
    scenes_ = [views[0][0], views[1][0], views[2][0]]
    many = []
    for k in range(40):
        pt, cam, _ = views[k % 3]
        p = abi.default_render_params(8 + (k % 5) * 3, 8 + (k % 4) * 2, 1 + k % 3, 6)
        p.sample_offset = k
        many.append((pt, cam, p))
    br = BatchRenderer(make())
    br.render(many)
    for k in (0, 1, 2, 17, 38, 39):
        pt, cam, p = many[k]
        pt.camera, pt.params = cam, p
        _check_equal(br.read(k), pt.render(), f"view {k} of 40")


def test_launch_shape_does_not_change_results(batch_setup):
    views, plain, make = batch_setup
    pt, cam, p = views[1]
    pt.camera, pt.params = cam, p
    for kw in (dict(trace_block=128), dict(queue_chunk=64, trace_groups_per_cu=3), dict(paths_in_flight=777),
               dict(refill_lanes=1), dict(refill_lanes=64), dict(sort_by_material=4), dict(sort_by_material=15),
               dict(trace_block=256, queue_chunk=1024, trace_groups_per_cu=16, paths_in_flight=0, refill_lanes=0,
                    sort_by_material=0)):
        pt.ctx.set_tuning(**kw)
        _check_equal(pt.render(), plain[1], f"tuning {kw}")
    pt.ctx.set_tuning(trace_block=0, queue_chunk=0, trace_groups_per_cu=0, paths_in_flight=0, refill_lanes=0,
                      sort_by_material=0)


def test_autotune_keeps_a_valid_shape(batch_setup):
    from photon_amd.backend import BatchRenderer
    views, plain, make = batch_setup
    br = BatchRenderer(make())
    t, ms = br.autotune(views[:3], repeats=1)
    assert ms > 0
    assert t.trace_block in (128, 256) and t.queue_chunk % 64 == 0
    br.render(views)
    for v in range(len(views)):
        _check_equal(br.read(v), plain[v], f"after autotune view {v}")


def test_batch_errors(batch_setup):
    from photon_amd._lib import PhxError
    from photon_amd.backend import BatchRenderer, Context
    views, _, make = batch_setup
    br = BatchRenderer(make())
    with pytest.raises(PhxError):
        br.render([])
    empty = Context(0)  # no scene set
    with pytest.raises(PhxError):
        br.render([(empty, views[0][1], views[0][2])])
    bad = abi.default_render_params(8, 8, 1, 4)
    bad.adaptive_sampling = 1
    with pytest.raises(PhxError):
        br.render([(views[0][0], views[0][1], bad)])


@pytest.mark.parametrize("use_graph", [1, 0])
def test_batch_with_a_volume_scene_equals_plain_renders(use_graph):
    # A fog box (volume events sampled ahead of shading) batched with a scene
    # without volumes, and two views of the fog box alone: each view equals
    # its own render.
    from photon_amd.backend import BatchRenderer, Context
    fog = scenes.fog_box(n=32, macro=8, env_size=(64, 32))
    cb = scenes.cornell_box()
    pf, pc = _tracer(lambda: Context(0), fog), _tracer(lambda: Context(0), cb)
    mixed = [_view(pf, fog, 24, 16, 4, 12), _view(pc, cb, 16, 16, 4, 8), _view(pf, fog, 16, 24, 2, 12, offset=3,
                                                                             orbit=0.5)]
    plain = []
    for pt, cam, p in mixed:
        pt.camera, pt.params = cam, p
        plain.append(pt.render())
    for views, refs, tag in ((mixed, plain, "fog + cornell"), ([mixed[0], mixed[2]], [plain[0], plain[2]], "fog")):
        br = BatchRenderer(Context(0))
        br.ctx.set_tuning(use_graph=use_graph)
        for rep in range(2):
            br.render(views)
            for v in range(len(views)):
                _check_equal(br.read(v), refs[v], f"{tag} graph={use_graph} rep={rep} view {v}")


def _overlapping_soup(n, seed):
    # Large, heavily overlapping triangles: most children of a node are entered,
    # so traversal stacks run past the LDS part into the spill area; about one
    # triangle in twelve emits, so the any-hit pass traces as deep.
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-1, 1, (n, 1, 3))
    pos = (centers + rng.normal(scale=0.6, size=(n, 3, 3))).reshape(-1, 3)
    mids = (rng.random(n) < 1 / 12).astype(np.uint32)
    return scenes.Scene(pos, np.arange(3 * n, dtype=np.uint32).reshape(-1, 3),
                        [scenes.diffuse((0.6, 0.5, 0.4)), scenes.emitter((4.0, 4.0, 4.0), front_only=False)],
                        material_ids=mids)


@pytest.mark.parametrize("which", ["soup", "fog"])
def test_forked_bounce_loop_equals_serial(monkeypatch, which):
    # The any-hit pass runs on a second stream beside the closest-hit and
    # shading passes; on one stream (PHX_SERIAL_BOUNCE=1) the same launches run
    # in turn.  In the soup both traversal passes spill deep stacks, each to
    # its own area; in the fog box the second stream runs the shadow rays'
    # ratio tracking (k_shadow_transmit), which scales the shadow
    # contributions in place.  The images must agree bit for bit.
    from photon_amd.backend import Context, HipBackend, PathTracer, make_camera
    pt = PathTracer(HipBackend(context=Context(0)))
    if which == "soup":
        s = _overlapping_soup(6000, seed=11)
        pt.set_scene(s)
        pt.camera = make_camera((0.3, 0.2, 4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, 1.0)
        pt.params = abi.default_render_params(96, 96, 8, 8)
    else:
        s = scenes.fog_box(n=32, macro=8)
        pt.set_scene(s)
        pt.camera = make_camera(s.camera["look_from"], s.camera["look_at"], s.camera["vup"], s.camera["vfov"], 1.0)
        pt.params = abi.default_render_params(96, 96, 8, 12)
    forked = pt.render()
    assert pt.ctx.stats().shadow_rays > 0
    monkeypatch.setenv("PHX_SERIAL_BOUNCE", "1")
    serial = pt.render()
    monkeypatch.delenv("PHX_SERIAL_BOUNCE")
    _check_equal(forked, serial, "forked vs serial")
    assert float(np.abs(forked.color).sum()) > 0.0
    _check_equal(pt.render(), forked, "forked again")
