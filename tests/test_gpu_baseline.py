"""The reference-structure baseline (phx_render_reference_structure) renders the
same frames as the wavefront backend and the oracle: it is the stand-in for the
reference's Kokkos-HIP build, so its images must match before its time means
anything."""
import numpy as np
import pytest

import oracle
from photon_amd import abi, scenes
from tests.helpers import rel_mse

pytestmark = pytest.mark.gpu


def _tracer(ctx, s, w, h, spp, depth):
    from photon_amd.backend import HipBackend, PathTracer, make_camera
    pt = PathTracer(HipBackend(context=ctx))
    pt.set_scene(s)
    pt.camera = make_camera(s.camera["look_from"], s.camera["look_at"], s.camera["vup"], s.camera["vfov"], w / h)
    pt.params = abi.default_render_params(w, h, spp, depth)
    return pt


@pytest.mark.parametrize("name", ["cornell", "spheres"])
def test_reference_structure_matches_oracle_and_wavefront(gpu_ctx, name):
    if name == "cornell":
        s, w, h, spp = scenes.cornell_box(), 64, 64, 16
    else:
        s, w, h, spp = scenes.sphere_grid(seed=1234, grid=2, level=3), 48, 32, 4
    pt = _tracer(gpu_ctx, s, w, h, spp, 8)
    base = pt.render(structure="reference")
    wave = pt.render()
    ref, ex, _ = oracle.OracleScene(s).render(pt.camera, pt.params, aovs=True)
    e_ref = rel_mse(base.color, ref)
    e_wave = rel_mse(base.color, wave.color)
    same = np.mean(np.abs(base.color - ref) <= 1e-4 * (1 + np.abs(ref)))
    print(f"{name}: baseline vs oracle relMSE {e_ref:.3e} ({same:.4f} of pixels equal to 1e-4), "
          f"vs wavefront {e_wave:.3e}", flush=True)
    assert e_ref < 1e-3 and e_wave < 1e-3
    assert same > 0.9
    np.testing.assert_allclose(base.alpha, ex["alpha"], atol=1e-6)
    st = pt.ctx.stats()
    assert st.samples == w * h * spp


def test_reference_structure_sample_offset_and_repeat(gpu_ctx):
    s = scenes.cornell_box()
    pt = _tracer(gpu_ctx, s, 40, 24, 3, 6)
    a = pt.render(structure="reference").color
    b = pt.render(structure="reference").color
    np.testing.assert_array_equal(a, b)
    pt.params.sample_offset = 7
    c = pt.render(structure="reference").color
    ref, _, _ = oracle.OracleScene(s).render(pt.camera, pt.params)
    assert not np.array_equal(a, c)
    assert rel_mse(c, ref) < 1e-3
