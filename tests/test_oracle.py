"""Pins the oracle (CPU restatement of the reference) before it is trusted.

Every check here is independent of the oracle's own formula: known answers of
the reference's tests (tests/*.cpp of the reference), closed forms, textbook
microfacet models, brute force and invariants.  A plausible slip in the
oracle (dropped term, wrong sign or index, swapped operand) fails one of them.
"""
import math

import numpy as np
import pytest

import oracle
from photon_amd import abi, scenes
from tests.helpers import golden, golden_bsdf_inputs, random_dirs, single_triangle, triangle_soup, two_quads

M32 = 0xFFFFFFFF


# ---- RNG and stream hashing (rng.h:15-29, pathtracer.cpp:20-34) ---------------
def _xorshift(x, n):
    out = []
    for _ in range(n):
        x ^= (x << 13) & M32
        x ^= x >> 17
        x ^= (x << 5) & M32
        out.append(x)
    return out


def _hash(x):
    x ^= x >> 16
    x = (x * 0x7FEB352D) & M32
    x ^= x >> 15
    x = (x * 0x846CA68B) & M32
    x ^= x >> 16
    return x


def test_rng_known_answers():
    # integer recurrences evaluated by hand-written Python big-int arithmetic
    assert oracle.hash_u32(0) == 0
    for x in (1, 2, 12345, 0x9E3779B9, M32):
        assert oracle.hash_u32(x) == _hash(x)
    for (p, s, d) in ((0, 0, 0), (7, 3, 1), (1023 * 1024 + 5, 63, 0x10000 + 4)):
        seed = _hash((p + _hash((s + _hash((d + 0x9E3779B9) & M32)) & M32)) & M32)
        seed = seed or 1
        u, f = oracle.rng_stream(p, s, d, 16)
        assert list(u) == _xorshift(seed, 16)
        np.testing.assert_array_equal(f, np.array([(v >> 8) / 16777216.0 for v in u], np.float32))


def test_rng_matches_golden_streams():
    # tests/golden/rng_streams.json: seeds and streams from an independent
    # big-integer restatement (make_golden.py)
    for r in golden("rng_streams.json"):
        u, f = oracle.rng_stream(r["pixel"], r["sample"], r["dim"], 16)
        assert list(u) == r["u32"], r
        np.testing.assert_array_equal(f, np.array(r["f32"], np.float32))


def test_xorshift_zero_seed_maps_to_one():
    # Rng(0) starts at 1 (rng.h:13): the stream of 1 is 270369, 67634689, ...
    assert _xorshift(1, 2) == [270369, 67634689]


# ---- sampling (sampling_test.cpp:38-107) ---------------------------------------
def test_power_heuristic_known_answers():
    assert oracle.power_heuristic(0.5, 0.5) == pytest.approx(0.5)
    assert oracle.power_heuristic(1.0, 0.0) == pytest.approx(1.0)
    assert oracle.power_heuristic(0.0, 0.0) == 0.0
    assert oracle.power_heuristic(3.0, 1.0) == pytest.approx(0.9, rel=1e-6)


def test_onb_orthonormal():
    rng = np.random.default_rng(0)
    for n in list(random_dirs(64, rng, upper=False)) + [np.array([0, 0, -1.0]), np.array([0, 0, 1.0])]:
        n = (n / np.linalg.norm(n)).astype(np.float32)
        t, b = oracle.onb(n)
        m = np.stack([t, b, n]).astype(np.float64)
        np.testing.assert_allclose(m @ m.T, np.eye(3), atol=1e-4)  # float ONB near n.z = -1


# ---- triangle and mesh intersection (mesh_test.cpp, backend_test.cpp) ----------
def test_mesh_hit_closed_form():
    o = oracle.OracleScene(single_triangle())
    h = o.trace_closest_full(np.array([[0.25, 0.25, 1.0]]), np.array([[0.0, 0.0, -1.0]]), [1e-3], [1e30])[0]
    assert h.hit == 1
    np.testing.assert_allclose(list(h.position), [0.25, 0.25, 0.0], atol=1e-6)
    np.testing.assert_allclose(list(h.uv), [0.25, 0.25], atol=1e-6)
    np.testing.assert_allclose(list(h.bary), [0.25, 0.25], atol=1e-6)
    assert h.material_id == 7
    sn = np.array([0, 0, 1.0]) * 0.5 + np.array([0, 1.0, 0]) * 0.25 + np.array([1.0, 0, 0]) * 0.25
    np.testing.assert_allclose(list(h.shading_normal), sn / np.linalg.norm(sn), atol=1e-6)
    np.testing.assert_allclose(list(h.normal), [0, 0, 1], atol=1e-7)
    assert h.t == pytest.approx(1.0)
    # without normals / texcoords: sn = n, uv = (0, 0)
    o2 = oracle.OracleScene(single_triangle(with_attrs=False))
    h2 = o2.trace_closest_full(np.array([[0.25, 0.25, 1.0]]), np.array([[0.0, 0.0, -1.0]]), [1e-3], [1e30])[0]
    assert h2.hit == 1
    np.testing.assert_allclose(list(h2.shading_normal), list(h2.normal))
    assert list(h2.uv) == [0.0, 0.0]


def test_two_quads_backend_semantics():
    o = oracle.OracleScene(two_quads())
    org = np.array([[0, 0, 2], [0, 0, 2], [0, 0, 2], [0, 0, 2]], np.float32)
    d = np.array([[0, 0, -1], [0, 0, 1], [0, 0, -1], [0, 0, -1]], np.float32)
    tmin = np.array([1e-3, 1e-3, 1.0, 2.5], np.float32)
    tmax = np.array([1e30, 1e30, 0.5, 3.0], np.float32)   # ray 2 inactive (tmax < tmin)
    t, prim, _, _ = o.trace_closest(org, d, tmin, tmax)
    assert prim[0] != 0xFFFFFFFF and t[0] == pytest.approx(2.0)
    assert prim[1] == 0xFFFFFFFF
    assert prim[2] == 0xFFFFFFFF
    assert prim[3] != 0xFFFFFFFF and t[3] == pytest.approx(3.0)  # interval is inclusive: t = tmax hits
    occ = o.trace_occluded(org, d, tmin, tmax)
    assert list(occ) == [1, 0, 0, 1]


@pytest.mark.parametrize("scene_fn", [lambda: scenes.cornell_box(), lambda: triangle_soup(3000, seed=3)])
def test_reference_bvh_equals_brute_force(scene_fn):
    s = scene_fn()
    o = oracle.OracleScene(s)
    org, d, tmin, tmax = scenes.random_rays(s, 4000, seed=11)
    t1, p1, u1, v1 = o.trace_closest(org, d, tmin, tmax)
    t2, p2, u2, v2 = o.trace_closest(org, d, tmin, tmax, brute=True)
    same = p1 == p2
    # identical except exact-t ties between triangles, which the reference
    # resolves by visit order and brute force by index
    assert same.mean() > 0.999
    np.testing.assert_array_equal(t1[same], t2[same])
    np.testing.assert_array_equal(u1[same], u2[same])
    np.testing.assert_array_equal(t1, t2)
    o1 = o.trace_occluded(org, d, tmin, tmax)
    o2 = o.trace_occluded(org, d, tmin, tmax, brute=True)
    np.testing.assert_array_equal(o1, o2)


def test_hit_distance_closed_form_plane():
    # a ray against a large axis-aligned quad: t = (z0 - oz) / dz
    v, t = scenes._quad((-10, -10, 0.5), (20, 0, 0), (0, 20, 0))
    s = scenes.Scene(v.astype(np.float32), t, [abi.default_material()])
    o = oracle.OracleScene(s, brute=True)
    rng = np.random.default_rng(5)
    org = rng.uniform(-1, 1, (500, 3)).astype(np.float32)
    org[:, 2] = rng.uniform(2, 4, 500)
    d = random_dirs(500, rng).astype(np.float32)
    d[:, 2] = -np.abs(d[:, 2]) - 0.2
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    tt, prim, _, _ = o.trace_closest(org, d, np.zeros(500), np.full(500, 1e30))
    expect = (0.5 - org[:, 2].astype(np.float64)) / d[:, 2]
    p = org.astype(np.float64) + d * expect[:, None]
    inside = np.all(np.abs(p[:, :2]) < 9.999, axis=1)
    outside = np.any(np.abs(p[:, :2]) > 10.001, axis=1)
    assert inside.sum() > 400
    np.testing.assert_allclose(tt[inside], expect[inside], rtol=2e-6)
    assert np.all(prim[outside] == 0xFFFFFFFF)


# ---- Disney BSDF ---------------------------------------------------------------
def _frame_tuple(n, rng):
    wo = random_dirs(n, rng).astype(np.float32)
    wi = random_dirs(n, rng).astype(np.float32)
    ng = np.tile(np.array([0, 0, 1], np.float32), (n, 1))
    return wo, wi, ng


def _ggx_D(noh, a):
    a2 = a * a
    d = noh * noh * (a2 - 1) + 1
    return a2 / (math.pi * d * d)


def _smith_G1(nov, a):
    a2 = a * a
    return 2 * nov / (nov + np.sqrt(a2 + (1 - a2) * nov * nov))


def test_bsdf_lambert_special_case():
    # specular = metallic = transmission = sheen = clearcoat = subsurface = 0:
    # f*cos = base * NoL / pi, pdf = NoL / pi (disney_bsdf.h:104-137,148-175)
    rng = np.random.default_rng(1)
    n = 400
    wo, wi, ng = _frame_tuple(n, rng)
    mats = []
    for _ in range(n):
        m = abi.default_material()
        m.base_color[:] = tuple(rng.uniform(0, 1, 3))
        m.specular = 0.0
        m.roughness = float(rng.uniform(0.01, 1))
        mats.append(m)
    f, pdf = oracle.bsdf_eval(mats, wo, wi, ng, ng)
    base = np.array([list(m.base_color) for m in mats], np.float64)
    nol = wi[:, 2].astype(np.float64)
    np.testing.assert_allclose(f, base * nol[:, None] / math.pi, rtol=2e-6, atol=1e-9)
    np.testing.assert_allclose(pdf, nol / math.pi, rtol=2e-6, atol=1e-9)


def test_bsdf_subsurface_at_normal_incidence():
    # specular = metallic = transmission = sheen = clearcoat = 0, so the
    # diffuse weight is 1 and no other lobe contributes.  At NoL = 1 the
    # subsurface factor's Fresnel term is 1 (its (1 - NoL)^5 vanishes) and
    # eval_diffuse (disney_bsdf.h:177-193) reduces to Lambert scaled by
    # 1.25 / (1 + 0.001): subsurface = 1 gives the subsurface colour times that,
    # subsurface = 0.5 the mean of Lambert in the base colour and the scaled
    # lobe in the mean colour, whatever the outgoing direction.
    rng = np.random.default_rng(11)
    n = 300
    wo, _, ng = _frame_tuple(n, rng)
    wi = ng.copy()
    k = 1.25 / 1.001 / math.pi
    for s_amt in (1.0, 0.5):
        mats = []
        for _ in range(n):
            m = abi.default_material()
            m.base_color[:] = tuple(rng.uniform(0, 1, 3))
            m.subsurface_color[:] = tuple(rng.uniform(0, 1, 3))
            m.specular = 0.0
            m.subsurface = s_amt
            m.roughness = float(rng.uniform(0.01, 1))
            mats.append(m)
        f, _ = oracle.bsdf_eval(mats, wo, wi, ng, ng, use_double=True)
        base = np.array([list(m.base_color) for m in mats], np.float64)
        ssc = np.array([list(m.subsurface_color) for m in mats], np.float64)
        if s_amt == 1.0:
            want = ssc * k
        else:
            want = 0.5 * base / math.pi + 0.5 * (0.5 * (base + ssc)) * k
        np.testing.assert_allclose(f, want, rtol=2e-6, atol=1e-9, err_msg=f"subsurface {s_amt}")


def test_bsdf_matches_golden_closed_forms():
    # tests/golden/bsdf_closed_forms.json: Lambert and F = 1 metal GGX in double
    mats, wo, wi, ng, f_ref, pdf_ref = golden_bsdf_inputs()
    f, pdf = oracle.bsdf_eval(mats, wo, wi, ng, ng, use_double=True)
    # the fixture's directions are rounded to float32 here: a few 1e-7 of slack
    np.testing.assert_allclose(f, f_ref, rtol=2e-6)
    np.testing.assert_allclose(pdf, pdf_ref, rtol=2e-6)


def test_bsdf_metal_ggx_special_case():
    # metallic = 1, base = 1, no clearcoat: F = 1, f*cos = D G / (4 NoV),
    # pdf = D NoH / (4 VoH) (disney_bsdf.h:95-100,111-125,200-234)
    rng = np.random.default_rng(2)
    n = 400
    wo, wi, ng = _frame_tuple(n, rng)
    mats, alphas = [], []
    for _ in range(n):
        m = abi.default_material()
        m.base_color[:] = (1.0, 1.0, 1.0)
        m.metallic = 1.0
        m.roughness = float(rng.uniform(0.05, 1.0))
        alphas.append(np.float32(m.roughness))
        mats.append(m)
    f, pdf = oracle.bsdf_eval(mats, wo, wi, ng, ng, use_double=True)
    a = np.array(alphas, np.float64)
    wo64, wi64 = wo.astype(np.float64), wi.astype(np.float64)
    h = wo64 + wi64
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    nov, nol, noh = wo64[:, 2], wi64[:, 2], h[:, 2]
    voh = np.sum(wo64 * h, axis=1)
    D = _ggx_D(noh, a)
    G = _smith_G1(nov, a) * _smith_G1(nol, a)
    np.testing.assert_allclose(f[:, 0], D * G / (4 * nov), rtol=1e-5)
    np.testing.assert_allclose(f[:, 1], f[:, 0], rtol=0)
    np.testing.assert_allclose(pdf, D * noh / (4 * voh), rtol=1e-5)


def test_bsdf_sheen_and_clearcoat_textbook():
    # sheen only: (1-HoL)^5 * sheen * NoL (Burley 2012), tint 0 -> white;
    # clearcoat only: 0.25 * cc * F(0.04) D G / (4 NoV) with the GGX D, G.
    rng = np.random.default_rng(3)
    n = 200
    wo, wi, ng = _frame_tuple(n, rng)
    wo64, wi64 = wo.astype(np.float64), wi.astype(np.float64)
    h = wo64 + wi64
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    sheen_m, cc_m = [], []
    for _ in range(n):
        m = abi.default_material()
        m.base_color[:] = (0.0, 0.0, 0.0)  # no diffuse energy
        m.specular = 0.0
        m.sheen = 0.7
        m.sheen_tint = 0.0
        sheen_m.append(m)
        c = abi.default_material()
        c.base_color[:] = (0.0, 0.0, 0.0)
        c.specular = 0.0
        c.clearcoat = 0.8
        c.clearcoat_roughness = 0.2
        cc_m.append(c)
    f, _ = oracle.bsdf_eval(sheen_m, wo, wi, ng, ng, use_double=True)
    hol = np.abs(np.sum(h * wi64, axis=1))
    np.testing.assert_allclose(f[:, 0], 0.7 * (1 - hol) ** 5 * wi64[:, 2], rtol=1e-5, atol=1e-12)
    f, _ = oracle.bsdf_eval(cc_m, wo, wi, ng, ng, use_double=True)
    voh = np.sum(wo64 * h, axis=1)
    F = 0.04 + 0.96 * (1 - voh) ** 5
    a = 0.2
    D = _ggx_D(h[:, 2], a)
    G = _smith_G1(wo64[:, 2], a) * _smith_G1(wi64[:, 2], a)
    np.testing.assert_allclose(f[:, 0], 0.25 * 0.8 * F * D * G / (4 * wo64[:, 2]), rtol=1e-5)


def test_bsdf_anisotropy_limit_and_D_normalization():
    rng = np.random.default_rng(4)
    n = 200
    wo, wi, ng = _frame_tuple(n, rng)
    iso, an = [], []
    for _ in range(n):
        m = abi.default_material()
        m.metallic = 1.0
        m.base_color[:] = (1, 1, 1)
        m.roughness = 0.3
        iso.append(m)
        m2 = abi.default_material()
        m2.metallic = 1.0
        m2.base_color[:] = (1, 1, 1)
        m2.roughness = 0.3
        m2.anisotropic = 1e-7
        an.append(m2)
    f0, _ = oracle.bsdf_eval(iso, wo, wi, ng, ng, use_double=True)
    f1, _ = oracle.bsdf_eval(an, wo, wi, ng, ng, use_double=True)
    np.testing.assert_allclose(f1, f0, rtol=1e-5)


def test_bsdf_reciprocity():
    # f(wo, wi)/NoL is symmetric for the lobes that do not use the
    # NoV-dependent diffuse split (specular = 0 or metallic = 1)
    rng = np.random.default_rng(5)
    n = 300
    wo, wi, ng = _frame_tuple(n, rng)
    mats = []
    for k in range(n):
        m = abi.default_material()
        m.base_color[:] = tuple(rng.uniform(0, 1, 3))
        m.roughness = float(rng.uniform(0.05, 1))
        if k % 2:
            m.metallic = 1.0
        else:
            m.specular = 0.0
            m.sheen = 0.5
        m.clearcoat = float(rng.uniform())
        mats.append(m)
    fa, _ = oracle.bsdf_eval(mats, wo, wi, ng, ng, use_double=True)
    fb, _ = oracle.bsdf_eval(mats, wi, wo, ng, ng, use_double=True)
    np.testing.assert_allclose(fa / wi[:, 2:3], fb / wo[:, 2:3], rtol=1e-5, atol=1e-12)


def test_bsdf_sample_pdf_consistency_and_furnace():
    # disney_bsdf_test.cpp:85-121: pdf(sample.wi) equals the sample's pdf,
    # and a white Lambert-like material has E[f/pdf] = 1.
    rng = np.random.default_rng(6)
    n = 4000
    wo = np.tile(np.array([0.2, 0.1, 0.97], np.float32) / np.linalg.norm([0.2, 0.1, 0.97]), (n, 1)).astype(np.float32)
    ng = np.tile(np.array([0, 0, 1], np.float32), (n, 1))
    mats = [abi.default_material() for _ in range(n)]
    wi, f, pdf, flags = oracle.bsdf_sample(mats, wo, ng, ng, np.arange(1, n + 1))
    ok = (pdf > 0) & ((flags & 1) == 0)
    f2, pdf2 = oracle.bsdf_eval([mats[i] for i in np.nonzero(ok)[0]], wo[ok], wi[ok], ng[ok], ng[ok])
    np.testing.assert_allclose(pdf2, pdf[ok], rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(f2, f[ok], rtol=1e-5, atol=1e-7)
    white = []
    for _ in range(n):
        m = abi.default_material()
        m.base_color[:] = (1, 1, 1)
        m.roughness = 1.0
        m.specular = 0.0
        white.append(m)
    wi, f, pdf, flags = oracle.bsdf_sample(white, wo, ng, ng, np.arange(1, n + 1))
    est = np.where(pdf[:, None] > 0, f / np.maximum(pdf, 1e-30)[:, None], 0).mean(0)
    np.testing.assert_allclose(est, 1.0, rtol=2e-3)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_bsdf_pdf_integrates_to_sampler_acceptance(seed):
    # The integral of pdf over the sphere equals the probability that
    # bsdf_sample returns a valid non-delta direction: estimated by uniform
    # sphere quadrature in FP64 and by sampling.  A dropped or misweighted
    # pdf term breaks the equality.
    rng = np.random.default_rng(100 + seed)
    from tests.helpers import random_material
    m = random_material(rng)
    m.transmission = 0.0
    wo = random_dirs(1, rng)[0].astype(np.float32)
    wo[2] = max(wo[2], 0.2)
    wo /= np.linalg.norm(wo)
    nq = 200000
    dirs = random_dirs(nq, rng, upper=False).astype(np.float32)
    ng = np.tile(np.array([0, 0, 1], np.float32), (nq, 1))
    _, pdf = oracle.bsdf_eval([m] * nq, np.tile(wo, (nq, 1)), dirs, ng, ng, use_double=True)
    integral = 4 * math.pi * pdf.astype(np.float64).mean()
    ns = 100000
    _, _, spdf, flags = oracle.bsdf_sample([m] * ns, np.tile(wo, (ns, 1)), ng[:ns], ng[:ns],
                                           np.arange(1, ns + 1) * 7919)
    accept = np.mean((spdf > 0) & ((flags & 1) == 0))
    assert integral == pytest.approx(accept, abs=0.02, rel=0.03)


# ---- environment map (light_test.cpp:100-171) -----------------------------------
def _env_scene(pix):
    s = two_quads()
    s.env_pixels = pix
    return s


def test_env_bright_texel_and_pdf():
    pix = np.zeros((2, 4, 3), np.float32)
    pix[0, 2] = 10.0
    o = oracle.OracleScene(_env_scene(pix))
    d, pdf = o.env_sample(100 + np.arange(200))
    _, rgb = o.env_eval(d)
    assert np.sum((pdf > 0) & (rgb[:, 0] > 1.0)) > 120
    # the sampled pdf equals pdf(dir); the uv -> dir -> uv round trip moves
    # sin(theta) by a few ulps near the poles
    pdf2, _ = o.env_eval(d)
    np.testing.assert_allclose(pdf2, pdf, rtol=1e-2)
    assert np.mean(np.abs(pdf2 / pdf - 1) < 1e-4) > 0.97


def test_env_pdf_normalized_and_uv_roundtrip():
    rng = np.random.default_rng(7)
    pix = rng.uniform(0, 2, (16, 32, 3)).astype(np.float32)
    o = oracle.OracleScene(_env_scene(pix))
    dirs = random_dirs(400000, rng, upper=False).astype(np.float32)
    pdf, _ = o.env_eval(dirs)
    assert 4 * math.pi * pdf.astype(np.float64).mean() == pytest.approx(1.0, abs=0.02)
    m, c = o.env_cdf()
    assert m[0] == 0 and m[-1] == 1 and np.all(np.diff(m) >= 0)
    assert np.all(c[:, 0] == 0) and np.all(c[:, -1] == 1)
    # the marginal follows luminance * sin(theta) of each row
    lum = (0.2126 * pix[..., 0] + 0.7152 * pix[..., 1] + 0.0722 * pix[..., 2]).sum(1)
    w = lum * np.sin((np.arange(16) + 0.5) * np.pi / 16)
    np.testing.assert_allclose(np.diff(m), w / w.sum(), rtol=1e-4)


# ---- camera (camera_test.cpp:26-74) ------------------------------------------
def test_camera_pinhole_basis():
    cam = oracle.camera_perspective((0, 0, 0), (0, 0, -1), (0, 1, 0), 60.0, 1.0)
    np.testing.assert_allclose(list(cam.w), [0, 0, 1], atol=1e-7)
    np.testing.assert_allclose(list(cam.u), [1, 0, 0], atol=1e-7)
    np.testing.assert_allclose(list(cam.v), [0, 1, 0], atol=1e-7)
    h = math.tan(math.radians(30))
    np.testing.assert_allclose(list(cam.horizontal), [2 * h, 0, 0], rtol=1e-6)
    np.testing.assert_allclose(list(cam.lower_left), [-h, -h, -1], rtol=1e-6)


# ---- volumes (tf_volume.h) -----------------------------------------------------
def _slab_volume(sigma, size=8):
    vol = abi.Volume()
    vol.world_to_local[:] = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]
    vol.origin[:] = (-1.0, -1.0, -1.0)
    vol.spacing[:] = (2.0 / (size - 1),) * 3
    vol.dims[:] = (size,) * 3
    vol.tf_size = 2
    vol.value_lo, vol.value_hi = 0.0, 1.0
    vol.density_scale = sigma
    vol.linear = 1
    values = np.ones(size ** 3, np.float32)
    tf = np.array([[1, 1, 1, 1], [1, 1, 1, 1]], np.float32)
    return vol, values, tf


def _volume_scene(vol, values, tf, maj):
    s = two_quads()
    s.volumes = [vol]
    s.volume_values = values
    s.volume_tf = tf.reshape(-1)
    s.volume_majorants = maj
    return s


def test_ratio_tracking_constant_slab_closed_form():
    sigma = 0.7
    vol, values, tf = _slab_volume(sigma)
    maj = oracle.volume_majorants(vol, values, tf)
    assert vol.majorant == pytest.approx(sigma)
    np.testing.assert_allclose(maj, sigma * 1.01, rtol=1e-6)
    o = oracle.OracleScene(_volume_scene(vol, values, tf, maj))
    seeds = np.arange(1, 40001)
    tr = o.volume_transmittance(np.array([-3.0, 0.1, 0.2], np.float32), np.array([1.0, 0, 0], np.float32), 0.0,
                                1e30, seeds)
    assert tr.mean() == pytest.approx(math.exp(-sigma * 2.0), rel=0.02)
    t = o.volume_scatter(np.array([-3.0, 0.1, 0.2], np.float32), np.array([1.0, 0, 0], np.float32), 0.0, 1e30, seeds)
    frac = np.mean(t >= 0)
    assert frac == pytest.approx(1 - math.exp(-sigma * 2.0), abs=0.01)
    # scatter distances follow the exponential law inside the slab
    inside = t[t >= 0] - 2.0
    assert np.all((inside >= -1e-5) & (inside <= 2.0 + 1e-5))
    assert np.mean(inside < 1.0) == pytest.approx((1 - math.exp(-sigma)) / (1 - math.exp(-2 * sigma)), abs=0.015)


def test_majorants_bound_the_extinction():
    rng = np.random.default_rng(8)
    size = 20
    vol, _, _ = _slab_volume(3.0, size)
    values = rng.uniform(0, 1, size ** 3).astype(np.float32)
    tf = np.stack([np.linspace(0, 1, 16)] * 4, 1).astype(np.float32)
    tf[:, 3] = np.clip(np.linspace(-0.5, 1.0, 16), 0, 1)
    maj = oracle.volume_majorants(vol, values, tf)
    assert maj.size == np.prod(list(vol.macro_dims))
    # brute force: the trilinear field at many points stays below its cell bound
    g = rng.uniform(0, size - 1, (20000, 3))
    i = np.minimum(g.astype(int), size - 2)
    f = g - i
    v = values.reshape(size, size, size)  # z, y, x

    def at(dx, dy, dz):
        return v[i[:, 2] + dz, i[:, 1] + dy, i[:, 0] + dx]
    val = 0
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                val = val + at(dx, dy, dz) * (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * \
                    (f[:, 2] if dz else 1 - f[:, 2])
    alpha = np.interp(val, np.linspace(0, 1, 16), tf[:, 3]) * 3.0
    md = np.array(list(vol.macro_dims))
    cell = np.minimum((g / (size - 1) * md).astype(int), md - 1)
    bound = maj[(cell[:, 2] * md[1] + cell[:, 1]) * md[0] + cell[:, 0]]
    assert np.all(alpha <= bound * (1 + 1e-5))


# ---- integrator -------------------------------------------------------------
def test_cornell_render_smoke():
    # pt_smoke_test.cpp:17-36: finite, mostly lit
    s = scenes.cornell_box()
    o = oracle.OracleScene(s)
    cam = oracle.camera_perspective(s.camera["look_from"], s.camera["look_at"], s.camera["vup"], 40.0, 1.0)
    p = abi.default_render_params(32, 32, 8, 8)
    img, ex, _ = o.render(cam, p, aovs=True)
    assert np.all(np.isfinite(img))
    assert np.mean(img.sum(-1) > 0) > 0.25
    assert np.all(ex["alpha"] == 1.0)
    nn = np.linalg.norm(ex["normal"], axis=-1)
    np.testing.assert_allclose(nn[ex["depth"] < 1e29], 1.0, atol=1e-5)


def test_env_lit_lambert_plane_closed_form():
    # A white Lambertian plane of albedo a under a constant environment of
    # radiance 1 reflects exactly a (no other geometry): the estimator with NEE
    # over the env, MIS and BSDF sampling must converge to a.
    v, t = scenes._quad((-50, 0, 50), (100, 0, 0), (0, 0, -100))
    s = scenes.Scene(v.astype(np.float32), t, [scenes.diffuse((0.5, 0.5, 0.5))], material_ids=np.zeros(2, np.uint32))
    s.env_pixels = np.ones((8, 16, 3), np.float32)
    o = oracle.OracleScene(s)
    cam = oracle.camera_perspective((0, 2, 0), (0, 0, -1), (0, 1, 0), 30.0, 1.0)
    p = abi.default_render_params(16, 16, 64, 4)
    img, _, _ = o.render(cam, p)
    assert img.mean() == pytest.approx(0.5, rel=0.01)


def test_float_and_double_engines_agree():
    s = scenes.cornell_box()
    o = oracle.OracleScene(s)
    cam = oracle.camera_perspective(s.camera["look_from"], s.camera["look_at"], s.camera["vup"], 40.0, 1.0)
    p = abi.default_render_params(24, 24, 4, 6)
    a, _, _ = o.render(cam, p)
    b, _, _ = o.render(cam, p, use_double=True)
    # same sample streams: paths agree except rare rounding-driven divergences
    assert np.mean(np.abs(a - b) < 1e-3 * (1 + np.abs(b))) > 0.95


def test_emissive_quad_irradiance_closed_form():
    # NEE over an emissive mesh (emissive CDF, area sampling, area-to-solid-angle
    # pdf, one-sided emission) combined by MIS with BSDF hits on the emitter:
    # a white Lambertian floor of albedo a under a downward-facing square
    # emitter of radiance Le (black, so nothing bounces back) reflects
    # a * Le * F at the point below the centre, F the differential-area to
    # parallel-rectangle form factor (sum of four corner rectangles).
    a, le, h, half = 0.5, 3.0, 1.0, 0.5
    fv, ft = scenes._quad((-50, 0, 50), (100, 0, 0), (0, 0, -100))
    ev, et = scenes._quad((-half, h, -half), (2 * half, 0, 0), (0, 0, 2 * half))
    v = np.concatenate([fv, ev]).astype(np.float32)
    t = np.concatenate([ft, et + len(fv)]).astype(np.uint32)
    mats = [scenes.diffuse((a, a, a)), scenes.emitter((le, le, le))]
    s = scenes.Scene(v, t, mats, material_ids=np.array([0, 0, 1, 1], np.uint32))
    o = oracle.OracleScene(s)
    cam = oracle.camera_perspective((0, 0.5, 3.0), (0, 0, 0), (0, 1, 0), 0.5, 1.0)
    p = abi.default_render_params(4, 4, 4096, 3)
    img, _, _ = o.render(cam, p)

    def corner(x, y):
        return (x / math.sqrt(1 + x * x) * math.atan(y / math.sqrt(1 + x * x)) +
                y / math.sqrt(1 + y * y) * math.atan(x / math.sqrt(1 + y * y))) / (2 * math.pi)
    f = 4 * corner(half / h, half / h)
    assert img.mean() == pytest.approx(a * le * f, rel=0.01)


# ---- detached estimator (the FD engine of the adjoint checks) -----------------
def test_detached_engine_equals_the_render_at_the_scene_parameters():
    s = scenes.textured_cornell(n_materials=8, tex_size=16, texture=scenes.colored_noise_texture(16, 3))
    o = oracle.OracleScene(s)
    cam = oracle.camera_perspective(s.camera["look_from"], s.camera["look_at"], s.camera["vup"], 40.0, 1.0)
    p = abi.default_render_params(24, 24, 4, 6)
    img, _, _ = o.render(cam, p)
    _, det = o.render_detached(cam, p, image=True)
    close = np.abs(det - img) <= 1e-4 * (1 + np.abs(img))
    assert close.mean() > 0.995
    assert abs(det.sum() - img.sum()) <= 1e-4 * img.sum()


def test_detached_engine_is_linear_in_the_albedo_of_a_lone_plane():
    # A Lambertian plane under a constant environment: every weight of every
    # path is linear in the albedo channel it carries, so dI_c/da_c = I_c / a_c
    # for the detached estimator (sampling decisions fixed).
    v, t = scenes._quad((-50, 0, 50), (100, 0, 0), (0, 0, -100))
    a0 = (0.4, 0.55, 0.7)
    s = scenes.Scene(v.astype(np.float32), t, [scenes.diffuse(a0)], material_ids=np.zeros(2, np.uint32))
    s.env_pixels = np.ones((8, 16, 3), np.float32)
    o = oracle.OracleScene(s)
    cam = oracle.camera_perspective((0, 2, 0), (0, 0, -1), (0, 1, 0), 30.0, 1.0)
    p = abi.default_render_params(8, 8, 16, 4)
    m = s.materials[0]
    base = np.array([[m.base_color[0], m.base_color[1], m.base_color[2], m.roughness, m.metallic]], np.float64)
    _, img = o.render_detached(cam, p, mat_params=base, image=True)
    h = 1e-4
    for c in range(3):
        plus, minus = base.copy(), base.copy()
        plus[0, c] += h
        minus[0, c] -= h
        delta = np.zeros((8, 8, 3), np.float32)
        delta[..., c] = 1.0
        lp, _ = o.render_detached(cam, p, mat_params=plus, delta=delta)
        lm, _ = o.render_detached(cam, p, mat_params=minus, delta=delta)
        fd = (lp - lm) / (2 * h)
        assert fd == pytest.approx(img[..., c].sum() / a0[c], rel=1e-7)
