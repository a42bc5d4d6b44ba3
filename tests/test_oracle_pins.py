"""Pins of the oracle functions that the reference's tests leave open.

Each check fixes the oracle by something other than its own formula: values
derived by hand from the reference's definitions (written out in the
comments), textbook physics (Fresnel, Snell), the reference's own test cases
(light_test.cpp, texture_test.cpp), closed forms of linear functions under
multilinear interpolation, and exact replays of the xorshift stream (the RNG
is pinned by big-integer known answers in test_oracle.py).  Functions pinned:
disney_lobe_weights, the subsurface term away from normal incidence, the
dielectric lobe of disney_bsdf_sample, sample_light, TextureAtlas::sample,
SamplerTable::sample, the material parameter sources and normal maps,
the background image and the exact random-number use of volume tracking.
"""
import math

import numpy as np
import pytest

import oracle
from photon_amd import abi, scenes
from tests.helpers import single_triangle, two_quads

M32 = 0xFFFFFFFF


class XorShift:
    """rng.h:9-30: xorshift32, 0 seeded as 1, floats from the top 24 bits."""

    def __init__(self, seed):
        self.x = seed if seed else 1

    def u32(self):
        x = self.x
        x ^= (x << 13) & M32
        x ^= x >> 17
        x ^= (x << 5) & M32
        self.x = x
        return x

    def f(self):
        return (self.u32() >> 8) / 16777216.0


def _seeds(n, key):
    """n seeds spread over 32 bits (xorshift's first draws from small seeds are small)."""
    return np.random.default_rng(key).integers(1, 1 << 32, n, dtype=np.uint64).astype(np.uint32)


def _mat(**kw):
    m = abi.default_material()
    for k, v in kw.items():
        if k in ("base_color", "emission", "subsurface_color"):
            getattr(m, k)[:] = v
        else:
            setattr(m, k, v)
    return m


def _lum(c):
    return 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]


# ---- disney_lobe_weights (disney_bsdf.h:148-175) ------------------------------
# Hand derivation for a dielectric with specular s and ior 1.5, transmission t,
# metallic m, base colour b:
#   opaque = 1 - t, dielectric = 1 - m, scale = 2 s
#   f0d = min(0.04 scale, 1), f90d = min(scale, 1)
#   diffuse scale(NoV) = 1 - (f0d + (f90d - f0d) (1 - NoV)^5)
#   wd = dielectric opaque scale(NoV) max(lum(b), 0.05) (+ 0.1 sheen)
#   F0 = lerp(f0d opaque, b, m), F90 = lerp(f90d opaque, 1, m)
#   ws = max(lum(F0 + (F90 - F0)(1 - NoV)^5), 0.05) if that lum > 0
#   wt = t dielectric,  wc = max(0.25 cc (0.04 + 0.96 (1 - NoV)^5), 0.02)
LOBE_CASES = [
    # (material fields, NoV, expected (wd, ws, wt, wc))
    (dict(base_color=(0.5, 0.5, 0.5), specular=0.5), 1.0, (0.96 * 0.5, 0.05, 0.0, 0.0)),
    (dict(base_color=(0.5, 0.5, 0.5), specular=0.5), 0.0, (0.0, 1.0, 0.0, 0.0)),
    (dict(base_color=(0.9, 0.6, 0.3), metallic=1.0), 1.0, (0.0, _lum((0.9, 0.6, 0.3)), 0.0, 0.0)),
    (dict(base_color=(0.9, 0.6, 0.3), metallic=1.0), 0.0, (0.0, 1.0, 0.0, 0.0)),
    (dict(base_color=(0.2, 0.2, 0.2), specular=0.0, sheen=0.4), 1.0, (0.2 + 0.04, 0.0, 0.0, 0.0)),
    (dict(base_color=(0.02, 0.02, 0.02), specular=0.0), 1.0, (0.05, 0.0, 0.0, 0.0)),  # lum floor 0.05
    (dict(base_color=(0.5, 0.5, 0.5), specular=0.0, clearcoat=0.8), 1.0, (0.5, 0.0, 0.0, 0.02)),
    (dict(base_color=(0.5, 0.5, 0.5), specular=0.0, clearcoat=0.8), 0.0, (0.5, 0.0, 0.0, 0.2)),
    # t = 0.5, s = 0.5: opaque 0.5, f0d 0.04 -> F0 0.02, F90 0.5; NoV 1: F = F0
    (dict(base_color=(0.5, 0.5, 0.5), specular=0.5, transmission=0.5), 1.0, (0.5 * 0.96 * 0.5, 0.05, 0.5, 0.0)),
    # NoV 0.5: (1 - NoV)^5 = 1/32; scale(NoV) = 1 - (0.04 + 0.96/32) = 0.93
    (dict(base_color=(0.5, 0.5, 0.5), specular=0.5, transmission=0.5), 0.5,
     (0.5 * 0.93 * 0.5, max(0.02 + 0.48 / 32, 0.05), 0.5, 0.0)),
    (dict(base_color=(0.5, 0.5, 0.5), specular=0.5, transmission=0.6, metallic=0.25), 1.0,
     (0.75 * 0.4 * 0.96 * 0.5, _lum([0.04 * 0.4 * 0.75 + 0.5 * 0.25] * 3), 0.45, 0.0)),
]


@pytest.mark.parametrize("case", range(len(LOBE_CASES)))
def test_lobe_weights_known_answers(case):
    fields, nov, want = LOBE_CASES[case]
    w = oracle.lobe_weights(_mat(**fields), nov, use_double=True)
    np.testing.assert_allclose(w, want, rtol=1e-6, atol=1e-7, err_msg=str(fields))


def test_lobe_selection_follows_the_weights():
    # disney_bsdf_sample's first draw picks the lobe by the normalized weights
    # (disney_bsdf.h:405-412).  At NoV = 1 this material has wd = 0.24,
    # ws = 0.05, wt = 0.5 (LOBE_CASES): the draw x of Rng(seed), replayed
    # here, must give a diffuse sample iff x < 0.24 / 0.79, a transmission
    # sample (flagged specular) iff x >= 0.29 / 0.79, a glossy one between.
    m = _mat(base_color=(0.5, 0.5, 0.5), specular=0.5, transmission=0.5, roughness=0.3,
             flags=abi.PHX_MATERIAL_FLAG_REFRACTIVE)
    n = 4000
    seeds = _seeds(n, 1)
    up = np.tile(np.array([0, 0, 1], np.float32), (n, 1))
    wi, f, pdf, flags = oracle.bsdf_sample([m] * n, up, up, up, seeds)
    pd, ps = 0.24 / 0.79, 0.05 / 0.79
    counts = [0, 0, 0]
    for i, seed in enumerate(seeds):
        x = XorShift(int(seed)).f()
        if abs(x - pd) < 1e-6 or abs(x - pd - ps) < 1e-6:
            continue
        if x < pd:
            assert flags[i] == 2, (seed, x)
            counts[0] += 1
        elif x < pd + ps:
            assert flags[i] == 0, (seed, x)
            counts[1] += 1
        else:
            assert flags[i] == 1 or pdf[i] == 0, (seed, x)
            counts[2] += 1
    assert min(counts) > 100


# ---- subsurface term (disney_bsdf.h:177-193) away from NoL = 1 ----------------
@pytest.mark.parametrize("nol", [0.5, 0.25])
def test_subsurface_term_off_normal(nol):
    # specular = metallic = transmission = 0: diffuse weight w = 1.  With
    # FL = (1 - NoL)^5 (1/32 at 0.5, 243/1024 at 0.25, exact in binary):
    #   Fss = 1 + (0.5 - 1) FL,  ss = 1.25 (Fss (1 / (NoL + 0.001) - 0.5) + 0.5)
    #   f cos = lerp(b NoL / pi, lerp(b, c, s) NoL ss / pi, s)
    fl = (1 - nol) ** 5
    fss = 1 + (0.5 - 1) * fl
    ss = 1.25 * (fss * (1 / (nol + 0.001) - 0.5) + 0.5)
    rng = np.random.default_rng(3)
    wi = np.array([[math.sqrt(1 - nol * nol), 0.0, nol]], np.float64)
    for s_amt in (1.0, 0.3):
        b = rng.uniform(0.1, 0.9, 3)
        c = rng.uniform(0.1, 0.9, 3)
        m = _mat(base_color=tuple(b), subsurface_color=tuple(c), subsurface=s_amt, specular=0.0, roughness=0.5)
        wo = np.array([[0.3, -0.2, math.sqrt(1 - 0.13)]], np.float64)
        ng = np.array([[0.0, 0.0, 1.0]])
        f, _ = oracle.bsdf_eval([m], wo.astype(np.float32), wi.astype(np.float32), ng.astype(np.float32),
                                ng.astype(np.float32), use_double=True)
        nol_f = float(np.float32(wi[0, 2]))  # the direction is handed over in float
        assert nol_f == nol
        lam = b * nol / math.pi
        sub = (b + (c - b) * s_amt) * nol * ss / math.pi
        want = lam + (sub - lam) * s_amt
        np.testing.assert_allclose(f[0], want, rtol=1e-6, err_msg=f"NoL {nol}, subsurface {s_amt}")


# ---- dielectric lobe of disney_bsdf_sample (disney_bsdf.h:455-505) ------------
def _fresnel_textbook(cos_i, eta):
    """Unpolarized Fresnel reflectance of a smooth boundary, eta = n_i / n_t."""
    sin2_t = eta * eta * (1 - cos_i * cos_i)
    if sin2_t >= 1:
        return 1.0
    cos_t = math.sqrt(1 - sin2_t)
    rs = (eta * cos_i - cos_t) / (eta * cos_i + cos_t)
    rp = (cos_i - eta * cos_t) / (cos_i + eta * cos_t)
    return 0.5 * (rs * rs + rp * rp)


def _glass(**kw):
    # transmission 1, metallic 0: wd = ws = wc = 0 and wt = 1, so every sample
    # takes the dielectric lobe; roughness below 0.001 makes it smooth (no
    # half-vector draws): the draws are the lobe pick, then the Fresnel coin
    f = dict(base_color=(0.9, 0.7, 0.5), transmission=1.0, roughness=0.0005, ior=1.5,
             flags=abi.PHX_MATERIAL_FLAG_REFRACTIVE)
    f.update(kw)
    return _mat(**f)


def _sample_dir(m, wo, n_seeds, seed0=1):
    n = n_seeds
    seeds = _seeds(n, seed0)
    wo_a = np.tile(np.asarray(wo, np.float32), (n, 1))
    up = np.tile(np.array([0, 0, 1], np.float32), (n, 1))
    wi, f, pdf, flags = oracle.bsdf_sample([m] * n, wo_a, up, up, seeds)
    coins = []
    for s in seeds:
        r = XorShift(int(s))
        r.f()
        coins.append(r.f())
    return wi, f, pdf, flags, np.array(coins)


@pytest.mark.parametrize("theta_deg", [0.0, 35.0, 70.0])
def test_dielectric_coin_fresnel_and_snell(theta_deg):
    th = math.radians(theta_deg)
    wo = (math.sin(th), 0.0, math.cos(th))
    m = _glass()
    wi, f, pdf, flags, coin = _sample_dir(m, wo, 600)
    wo32 = np.asarray(wo, np.float32).astype(np.float64)
    F = _fresnel_textbook(wo32[2], 1 / 1.5)
    if theta_deg == 0.0:
        assert F == pytest.approx(((1 - 1 / 1.5) / (1 + 1 / 1.5)) ** 2, rel=1e-12)  # 0.04
    refl = wi[:, 2] > 0
    for i in range(len(coin)):
        if abs(coin[i] - F) < 1e-5:
            continue
        assert refl[i] == (coin[i] < F), (i, coin[i], F)
    assert np.all(pdf == 1.0) and np.all(flags == 1)  # a delta lobe, flagged specular
    # reflection: the mirror direction, weight 1
    np.testing.assert_allclose(wi[refl], np.tile([-wo32[0], 0, wo32[2]], (refl.sum(), 1)), atol=2e-6)
    np.testing.assert_allclose(f[refl], 1.0, rtol=1e-6)
    # refraction: Snell's law sin t = sin i / 1.5 in the plane of incidence,
    # below the surface, weight the base colour
    t = ~refl
    assert t.sum() > 300
    np.testing.assert_allclose(np.hypot(wi[t, 0], wi[t, 1]), math.sin(th) / 1.5, atol=2e-6)
    assert np.all(wi[t, 2] < 0) and np.all(np.abs(wi[t, 1]) < 1e-6) and np.all(wi[t, 0] * wo32[0] <= 1e-7)
    np.testing.assert_allclose(f[t], np.tile([0.9, 0.7, 0.5], (t.sum(), 1)), rtol=1e-6)


def test_dielectric_total_internal_reflection():
    # from inside the glass (wo below the geometric normal): eta = 1.5, and
    # beyond the critical angle asin(1 / 1.5) = 41.8 deg Fresnel is 1: every
    # sample reflects, back into the glass
    for theta_deg in (50.0, 75.0):
        th = math.radians(theta_deg)
        wo = (math.sin(th), 0.0, -math.cos(th))
        wi, f, pdf, flags, coin = _sample_dir(_glass(), wo, 300)
        assert np.all(wi[:, 2] < 0)
        np.testing.assert_allclose(wi, np.tile([-math.sin(th), 0, -math.cos(th)], (300, 1)), atol=2e-6)
    # below the critical angle the coin decides again, against eta = 1.5
    th = math.radians(20.0)
    wo = (math.sin(th), 0.0, -math.cos(th))
    wi, f, pdf, flags, coin = _sample_dir(_glass(), wo, 600)
    F = _fresnel_textbook(float(np.float32(math.cos(th))), 1.5)
    back_in = wi[:, 2] < 0
    for i in range(600):
        if abs(coin[i] - F) > 1e-5:
            assert back_in[i] == (coin[i] < F)
    out = ~back_in
    np.testing.assert_allclose(np.hypot(wi[out, 0], wi[out, 1]), math.sin(th) * 1.5, atol=3e-6)


@pytest.mark.parametrize("variant", ["not refractive", "thin"])
def test_dielectric_pass_through_with_schlick(variant):
    # without refraction (no REFRACTIVE flag, or a thin surface) the coin is
    # Schlick's F = f0 + (1 - f0)(1 - cos)^5 with f0 = ((1 - ior)/(1 + ior))^2,
    # and a transmitted sample passes straight through: wi = -wo
    m = _glass(flags=0) if variant == "not refractive" else _glass(thin=1)
    th = math.radians(60.0)
    wo = (math.sin(th), 0.0, math.cos(th))
    wi, f, pdf, flags, coin = _sample_dir(m, wo, 600)
    c = float(np.float32(math.cos(th)))
    f0 = ((1 - 1.5) / (1 + 1.5)) ** 2
    F = f0 + (1 - f0) * (1 - c) ** 5
    refl = wi[:, 2] > 0
    for i in range(600):
        if abs(coin[i] - F) > 1e-5:
            assert refl[i] == (coin[i] < F)
    t = ~refl
    np.testing.assert_allclose(wi[t], -np.tile(np.asarray(wo, np.float32), (t.sum(), 1)), atol=1e-6)


# ---- sample_light (light.h:72-148) ---------------------------------------------
def _light(**kw):
    l = abi.Light()
    l.direction[:] = (0.0, -1.0, 0.0)
    l.color[:] = (1.0, 1.0, 1.0)
    l.intensity = 1.0
    l.spot_angle, l.spot_falloff = 0.5, 0.1
    for k, v in kw.items():
        if k in ("position", "direction", "color", "edge1", "edge2"):
            getattr(l, k)[:] = v
        else:
            setattr(l, k, v)
    return l


def test_sample_light_reference_cases():
    # light_test.cpp:37-98 of the reference, as known answers
    wi, Li, pdf, dist = oracle.sample_light(_light(type=0, position=(0, 5, 0), intensity=10.0), (0, 0, 0), 1)
    np.testing.assert_allclose(wi, [0, 1, 0], atol=1e-4)
    assert np.all(Li > 0) and pdf > 0 and dist == pytest.approx(5.0, abs=1e-4)
    np.testing.assert_allclose(Li, 10.0 / 25.0, rtol=1e-6)  # intensity / distance^2
    d = np.array([1.0, -2.0, 3.0]) / math.sqrt(14)
    wi, Li, pdf, dist = oracle.sample_light(_light(type=1, direction=tuple(d), color=(0.5, 0.25, 0.75), intensity=2.0),
                                            (1, 2, 3), 2)
    np.testing.assert_allclose(wi, -d, atol=1e-6)
    np.testing.assert_allclose(Li, [1.0, 0.5, 1.5], rtol=1e-6)
    assert pdf > 0 and dist > 1e20
    wi, Li, pdf, dist = oracle.sample_light(_light(type=2, position=(0, 5, 0), intensity=10.0), (0, 0, 0), 3)
    assert Li[0] > 0
    wi, Li, pdf, dist = oracle.sample_light(_light(type=2, position=(0, 5, 0), intensity=10.0, spot_angle=0.2,
                                                   spot_falloff=0.05), (10, 0, 0), 4)
    assert np.all(Li < 1e-3)


def test_spot_falloff_smoothstep():
    # a point on the rim between the inner (angle - falloff) and outer cones:
    # falloff = smoothstep(cos outer, cos inner, cos angle) = s^2 (3 - 2 s)
    ang = 0.45  # spot_angle 0.5, falloff 0.1: inner 0.4, outer 0.5
    p = (5 * math.tan(ang), 0.0, 0.0)
    wi, Li, pdf, dist = oracle.sample_light(_light(type=2, position=(0, 5, 0), intensity=10.0), p, 5)
    ca = math.cos(math.atan2(p[0], 5.0))
    s = (ca - math.cos(0.5)) / (math.cos(0.4) - math.cos(0.5))
    want = 10.0 / (25.0 + p[0] ** 2) * s * s * (3 - 2 * s)
    np.testing.assert_allclose(Li, want, rtol=2e-4)


def test_sphere_and_cone_lights():
    # a point light of radius r: the sampled point lies on the disk of radius r
    # through the centre facing the shading point; Li stays intensity / d^2 of
    # the centre
    for seed in _seeds(50, 2):
        wi, Li, pdf, dist = oracle.sample_light(_light(type=0, position=(0, 4, 0), intensity=8.0, radius=0.5),
                                                (0, 0, 0), int(seed))
        q = wi * dist
        assert abs(q[1] - 4.0) < 1e-5 and math.hypot(q[0], q[2]) <= 0.5 + 1e-5
        np.testing.assert_allclose(Li, 8.0 / 16.0, rtol=1e-6)
    # a directional light of angular radius r: directions uniform in the cone,
    # so cos(angle to the axis) is uniform on [cos r, 1]
    r = 0.3
    cosines = []
    for seed in _seeds(4000, 3):
        wi, Li, pdf, dist = oracle.sample_light(_light(type=1, direction=(0, -1, 0), radius=r), (0, 0, 0), int(seed))
        cosines.append(wi[1])
    cosines = np.array(cosines)
    assert cosines.min() >= math.cos(r) - 1e-6
    assert cosines.mean() == pytest.approx((1 + math.cos(r)) / 2, abs=2e-3)
    assert np.std(cosines) == pytest.approx((1 - math.cos(r)) / math.sqrt(12), rel=0.05)


def test_area_light_pdf_and_sides():
    # parallelogram p + u e1 + v e2 (u, v from the first two draws), one
    # sided: pdf = d^2 / (area cos_light), distance shortened by 0.999
    l = _light(type=3, position=(-0.5, 2.0, -0.5), edge1=(1.0, 0, 0), edge2=(0, 0, 1.0), intensity=3.0,
               color=(1.0, 0.5, 0.25))
    for seed in _seeds(40, 4):
        seed = int(seed)
        r = XorShift(seed)
        u, v = r.f(), r.f()
        sp = np.array([-0.5 + u, 2.0, -0.5 + v])
        wi, Li, pdf, dist = oracle.sample_light(l, (0.1, 0.0, 0.2), seed)
        to = sp - np.array([0.1, 0.0, 0.2])
        d = np.linalg.norm(to)
        # the light's normal is e1 x e2 = (0, 0, 1) x ... = (0, -1, 0): it faces down
        cos_l = to[1] / d
        np.testing.assert_allclose(wi, to / d, atol=1e-6)
        assert pdf == pytest.approx(d * d / (1.0 * cos_l), rel=1e-5)
        assert dist == pytest.approx(0.999 * d, rel=1e-6)
        np.testing.assert_allclose(Li, [3.0, 1.5, 0.75], rtol=1e-6)
    # from above, the back side: no light
    wi, Li, pdf, dist = oracle.sample_light(l, (0.0, 5.0, 0.0), 7)
    assert pdf == 0.0 and np.all(Li == 0)


# ---- TextureAtlas::sample (texture.h:25-64) -----------------------------------
def test_texture_reference_case_and_invalid_ids():
    # texture_test.cpp:14-44: a 2x2 texture looked up at its centre blends all
    # four texels equally; an unknown id gives white
    px = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], np.float32)
    infos = [(0, 2, 2)]
    np.testing.assert_allclose(oracle.texture_sample(px, infos, 0, 0.5, 0.5), [0.5, 0.5, 0.5], atol=1e-12)
    for bad in (-1, 99):
        np.testing.assert_array_equal(oracle.texture_sample(px, infos, bad, 0.5, 0.5), [1, 1, 1])
    np.testing.assert_array_equal(oracle.texture_sample(px, [(0, 0, 2)], 0, 0.5, 0.5), [1, 1, 1])


def test_texture_bilinear_ramp_vflip_and_repeat():
    # texel (x, y) holds (x, y, 7): bilinear interpolation reproduces a linear
    # function exactly, so away from the wrap seam the lookup at (u, v) is
    # (u W - 0.5, (1 - v) H - 0.5, 7): v is flipped (row 0 at v = 1)
    W, H = 8, 4
    px = np.array([[x, y, 7] for y in range(H) for x in range(W)], np.float32)
    infos = [(3, W, H)]
    px = np.concatenate([np.zeros((3, 3), np.float32), px])  # the texture starts at atlas offset 3
    for u, v in ((0.3, 0.6), (0.5, 0.5), (0.81, 0.2), (0.2, 0.85)):
        want = [u * W - 0.5, (1 - v) * H - 0.5, 7]
        np.testing.assert_allclose(oracle.texture_sample(px, infos, 0, u, v), want, atol=1e-9)
        # repeat wrap: whole periods do not change the lookup
        np.testing.assert_allclose(oracle.texture_sample(px, infos, 0, u + 2.0, v - 3.0), want, atol=1e-9)
    # at the seam the columns W - 1 and 0 blend: u = 0.5 / W / 2 is a quarter
    # texel inside column 0, so x = 0.75 * 0 + 0.25 * (W - 1)
    got = oracle.texture_sample(px, infos, 0, 0.25 / W, 0.5)
    assert got[0] == pytest.approx(0.25 * (W - 1), abs=1e-9)


# ---- SamplerTable::sample (sampler.h:73-138) -----------------------------------
def _sampler(**kw):
    return abi.sampler_info(**kw)


def test_transform_sampler_is_affine():
    # a transform sampler is out_M (in_M a + in_o) + out_o
    rng = np.random.default_rng(4)
    mi, mo = rng.normal(size=(4, 4)), rng.normal(size=(4, 4))
    oi, oo = rng.normal(size=4), rng.normal(size=4)
    si = _sampler(type_=abi.PHX_SAMPLER_TRANSFORM, in_transform=tuple(mi.reshape(-1)), in_offset=tuple(oi),
                  out_transform=tuple(mo.reshape(-1)), out_offset=tuple(oo))
    a = rng.normal(size=4)
    want = mo @ (mi.astype(np.float32).astype(np.float64) @ a + oi.astype(np.float32)) + oo.astype(np.float32)
    mi32, mo32 = mi.astype(np.float32).astype(np.float64), mo.astype(np.float32).astype(np.float64)
    want = mo32 @ (mi32 @ a + oi.astype(np.float32)) + oo.astype(np.float32)
    np.testing.assert_allclose(oracle.sampler_sample([si], None, 0, a), want, rtol=1e-12, atol=1e-12)
    # unknown samplers give (0, 0, 0, 1)
    np.testing.assert_array_equal(oracle.sampler_sample([si], None, 3, a), [0, 0, 0, 1])
    np.testing.assert_array_equal(oracle.sampler_sample([si], None, -1, a), [0, 0, 0, 1])


def test_image_samplers_reproduce_linear_functions():
    # texel (x, y, z) of an image holds (x, y, z, x + 10 y + 100 z): linear
    # lookups of interior coordinates return the coordinate times the size
    # minus one half, exactly (multilinear interpolation of a linear function)
    size = (5, 4, 3)
    tex = np.array([[x, y, z, x + 10 * y + 100 * z] for z in range(3) for y in range(4) for x in range(5)],
                   np.float32)
    for dims, typ in ((1, abi.PHX_SAMPLER_IMAGE1D), (2, abi.PHX_SAMPLER_IMAGE2D), (3, abi.PHX_SAMPLER_IMAGE3D)):
        sz = size[:dims] + (1,) * (3 - dims)
        n = sz[0] * sz[1] * sz[2]
        # the texel block of this image: a sub-array of the 3-D one
        sub = np.array([[x, y, z, x + 10 * y + 100 * z] for z in range(sz[2]) for y in range(sz[1])
                        for x in range(sz[0])], np.float32)
        texels = np.concatenate([np.full((2, 4), -5, np.float32), sub])  # starts at texel 2
        si = _sampler(type_=typ, size=sz, offset=2)
        for c in ((0.35, 0.5, 0.6), (0.62, 0.3, 0.45)):
            g = [c[d] * sz[d] - 0.5 if d < dims else 0.0 for d in range(3)]
            want = [g[0], g[1], g[2], g[0] + 10 * g[1] + 100 * g[2]]
            got = oracle.sampler_sample([si], texels, 0, (c[0], c[1], c[2], 1.0))
            np.testing.assert_allclose(got, want, atol=1e-9, err_msg=f"{dims}-D")
        assert n == sub.shape[0]
    del tex


@pytest.mark.parametrize("wrap,want", [
    # nearest lookups of texel value x on 4 texels at x = floor(4 c)
    (abi.PHX_WRAP_CLAMP, {-0.3: 0, 0.1: 0, 0.6: 2, 1.3: 3}),
    (abi.PHX_WRAP_REPEAT, {-0.3: 2, 0.1: 0, 0.6: 2, 1.3: 1}),     # floor(-1.2) = -2 -> 2; 5 -> 1
    (abi.PHX_WRAP_MIRROR, {-0.3: 1, 0.1: 0, 0.6: 2, 1.3: 2}),     # -2 -> 1; 5 -> 8 - 1 - 5 = 2
])
def test_sampler_wrap_modes(wrap, want):
    texels = np.array([[x, 0, 0, 1] for x in range(4)], np.float32)
    si = _sampler(type_=abi.PHX_SAMPLER_IMAGE1D, size=(4, 1, 1), linear=0, wrap=(wrap, 0, 0))
    for c, x in want.items():
        assert oracle.sampler_sample([si], texels, 0, (c, 0, 0, 1))[0] == x, (wrap, c)


# ---- material parameter sources and normal maps (pathtracer.cpp:111-247) ------
def _source_triangle(mat, attrs=None, samplers=(), texels=None):
    s = single_triangle()
    s.materials = [mat] * 8
    s.attributes = attrs or {}
    s.samplers = list(samplers)
    s.sampler_texels = texels
    return oracle.OracleScene(s)


def _src(m, name, kind, index):
    getattr(m, name).kind, getattr(m, name).index = kind, index


def test_attribute_sources_interpolate_barycentrically():
    # vertex attributes a0, a1, a2 at barycentrics (b1, b2): a0 (1 - b1 - b2) +
    # a1 b1 + a2 b2.  Slot 0 drives the base colour (its alpha multiplies alpha
    # with base_color_alpha), slot 1 roughness (squared: perceptual), metallic,
    # transmission and alpha; negative emission is cut to 0; values clamp to [0, 1]
    a0 = np.array([[1.0, 0.0, 0.0, 1.0], [0.6, 1.7, -0.5, 0.4], [0.0, 0.0, 0.0, 0.0]])
    a1 = np.array([[0.0, 1.0, 0.0, 0.5], [0.6, 1.7, -0.5, 0.4], [0.0, 0.0, 0.0, 0.0]])
    a2 = np.array([[0.0, 0.0, 1.0, 0.25], [0.6, 1.7, -0.5, 0.4], [0.0, 0.0, 0.0, 0.0]])
    attrs = {0: np.array([a0[0], a1[0], a2[0]]), 1: np.array([a0[1], a1[1], a2[1]])}
    m = abi.default_material()
    m.alpha = 0.8
    m.alpha_mode = abi.PHX_ALPHA_BLEND
    _src(m, "base_color_src", abi.PHX_SOURCE_ATTRIBUTE, 0)
    m.base_color_alpha = 1
    _src(m, "roughness_src", abi.PHX_SOURCE_ATTRIBUTE, 1)
    m.roughness_src_squared = 1
    o = _source_triangle(m, attrs)
    b1, b2 = 0.25, 0.35
    w0 = 1 - b1 - b2
    mat, alpha, n = o.surface_resolve(0, b1, b2, (0.25, 0.35, 0.0), (0, 0, 1))
    col = w0 * a0[0] + b1 * a1[0] + b2 * a2[0]
    np.testing.assert_allclose(list(mat.base_color), col[:3], rtol=1e-6)
    assert alpha == pytest.approx(0.8 * col[3], rel=1e-6)
    assert mat.roughness == pytest.approx(0.6 * 0.6, rel=1e-6)
    m.roughness_src_squared = 0
    _src(m, "metallic_src", abi.PHX_SOURCE_ATTRIBUTE, 1)       # .x = 0.6
    _src(m, "transmission_src", abi.PHX_SOURCE_ATTRIBUTE, 0)   # .x interpolated
    _src(m, "emission_src", abi.PHX_SOURCE_ATTRIBUTE, 1)       # (0.6, 1.7, -0.5) -> (0.6, 1.7, 0)
    _src(m, "alpha_src", abi.PHX_SOURCE_ATTRIBUTE, 1)          # .x = 0.6 multiplies alpha
    mat, alpha, n = _source_triangle(m, attrs).surface_resolve(0, b1, b2, (0.25, 0.35, 0.0), (0, 0, 1))
    assert mat.roughness == pytest.approx(0.6, rel=1e-6)
    assert mat.metallic == pytest.approx(0.6, rel=1e-6)
    assert mat.transmission == pytest.approx(col[0], rel=1e-6)
    np.testing.assert_allclose(list(mat.emission), [0.6, 1.7, 0.0], rtol=1e-6)
    assert alpha == pytest.approx(0.8 * col[3] * 0.6, rel=1e-6)
    # an absent slot gives (0, 0, 0, 1)
    _src(m, "base_color_src", abi.PHX_SOURCE_ATTRIBUTE, 3)
    mat, alpha, n = _source_triangle(m, attrs).surface_resolve(0, b1, b2, (0.25, 0.35, 0.0), (0, 0, 1))
    assert list(mat.base_color) == [0, 0, 0]


def test_sampler_source_of_the_world_position_and_normal():
    # a transform sampler reading the world position: base colour =
    # max(0.5 p + 0.1, 0); one reading the shading normal slot returns it
    m = abi.default_material()
    _src(m, "base_color_src", abi.PHX_SOURCE_SAMPLER, 0)
    _src(m, "metallic_src", abi.PHX_SOURCE_SAMPLER, 1)
    s0 = _sampler(type_=abi.PHX_SAMPLER_TRANSFORM, in_attribute=abi.PHX_ATTRIBUTE_WORLD_POSITION,
                  in_transform=(0.5, 0, 0, 0, 0, 0.5, 0, 0, 0, 0, 0.5, 0, 0, 0, 0, 1), in_offset=(0.1, 0.1, 0.1, 0))
    s1 = _sampler(type_=abi.PHX_SAMPLER_TRANSFORM, in_attribute=abi.PHX_ATTRIBUTE_WORLD_NORMAL,
                  in_transform=(0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0), in_offset=(0, 0, 0, 0),
                  out_transform=(0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0))
    s1.in_transform[:] = (1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1)
    o = _source_triangle(m, samplers=[s0, s1])
    p = (0.3, -0.6, 0.2)
    mat, alpha, n = o.surface_resolve(0, 0.3, 0.2, p, (0.0, 0.6, 0.8))
    np.testing.assert_allclose(list(mat.base_color), [0.25, 0.0, 0.2], rtol=1e-6, atol=1e-7)
    assert mat.metallic == pytest.approx(0.8, rel=1e-6)  # the normal's z


def _uv_quad_scene(mat, uv, texel):
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    s = scenes.Scene(pos, np.array([[0, 1, 2]], np.uint32), [mat], material_ids=np.zeros(1, np.uint32))
    s.attributes = {0: np.concatenate([np.asarray(uv, np.float32), np.zeros((3, 2), np.float32)], axis=1)}
    s.samplers = [_sampler(type_=abi.PHX_SAMPLER_IMAGE2D, in_attribute=0, size=(1, 1, 1))]
    s.sampler_texels = np.array([texel], np.float32)
    return oracle.OracleScene(s)


@pytest.mark.parametrize("uv,texel,want", [
    # uv = (x, y): tangent +x, bitangent +y; the map's value c encodes 2 c - 1
    ([[0, 0], [1, 0], [0, 1]], (1.0, 0.5, 0.5, 1), (1, 0, 0)),
    ([[0, 0], [1, 0], [0, 1]], (0.5, 0.5, 1.0, 1), (0, 0, 1)),
    ([[0, 0], [1, 0], [0, 1]], (0.75, 0.5, 0.75, 1), (math.sqrt(0.5), 0, math.sqrt(0.5))),
    ([[0, 0], [1, 0], [0, 1]], (0.5, 1.0, 0.5, 1), (0, 1, 0)),
    # uv = (y, -x): u grows along +y, v along -x
    ([[0, 0], [0, -1], [1, 0]], (1.0, 0.5, 0.5, 1), (0, 1, 0)),
    ([[0, 0], [0, -1], [1, 0]], (0.5, 1.0, 0.5, 1), (-1, 0, 0)),
    # degenerate coordinates: the shading normal stays
    ([[0.2, 0.2], [0.2, 0.2], [0.2, 0.2]], (1.0, 0.5, 0.5, 1), (0, 0, 1)),
])
def test_sourced_normal_map_frame(uv, texel, want):
    m = abi.default_material()
    _src(m, "normal_src", abi.PHX_SOURCE_SAMPLER, 0)
    o = _uv_quad_scene(m, uv, texel)
    mat, alpha, n = o.surface_resolve(0, 0.3, 0.3, (0.3, 0.3, 0.0), (0, 0, 1))
    np.testing.assert_allclose(n, want, atol=1e-7)


# ---- background image (pathtracer.cpp:370-389, 862-877) ------------------------
def test_background_image_at_texel_centres_and_in_between():
    # a scene the camera rays miss; an image of W x H RGBA texels rendered at
    # W x H puts every pixel centre on a texel centre (bilinear weight 0), so
    # each pixel is that texel's rgb * a and alpha a; rows of the image start
    # at the bottom, the frame's rows at the top unless bottom_up
    import ctypes as C
    W, H = 4, 3
    rng = np.random.default_rng(9)
    img = rng.uniform(0.1, 1.0, (H, W, 4)).astype(np.float32)
    s = scenes.Scene(np.array([[0, 0, 50], [1, 0, 50], [0, 1, 50]], np.float32), np.array([[0, 1, 2]], np.uint32),
                     [scenes.diffuse((0.5, 0.5, 0.5))])
    o = oracle.OracleScene(s)
    cam = oracle.camera_perspective((0, 0, 0), (0, 0, -1), (0, 1, 0), 40.0, W / H)
    for bottom_up in (0, 1):
        p = abi.default_render_params(W, H, 1, 2)
        p.bottom_up = bottom_up
        p.background_image = img.ctypes.data
        p.background_width, p.background_height = W, H
        color, extras, _ = o.render(cam, p, aovs=True)
        rows = img if bottom_up else img[::-1]
        np.testing.assert_allclose(color, rows[..., :3] * rows[..., 3:], rtol=1e-6)
        np.testing.assert_allclose(extras["alpha"], rows[..., 3], rtol=1e-6)
    # at twice the resolution the pixel centres fall a quarter texel from the
    # texel centres: weights 3/4 and 1/4 along each axis, the edges clamped
    p = abi.default_render_params(2 * W, 2 * H, 1, 2)
    p.bottom_up = 1
    p.background_image = img.ctypes.data
    p.background_width, p.background_height = W, H
    color, extras, _ = o.render(cam, p, aovs=True)

    def axis_weights(i, n):
        f = (i + 0.5) / (2 * n) * n - 0.5
        lo = math.floor(f)
        t = f - lo
        return [(min(max(lo, 0), n - 1), 1 - t), (min(max(lo + 1, 0), n - 1), t)]

    want = np.zeros((2 * H, 2 * W, 4))
    for y in range(2 * H):
        for x in range(2 * W):
            for yy, wy in axis_weights(y, H):
                for xx, wx in axis_weights(x, W):
                    want[y, x] += wy * wx * img[yy, xx]
    np.testing.assert_allclose(color, want[..., :3] * want[..., 3:], rtol=1e-5)
    np.testing.assert_allclose(extras["alpha"], want[..., 3], rtol=1e-5)
    del C


# ---- exact random-number use of volume tracking (tf_volume.h:150-283) ----------
def _homogeneous_slab(sigma, majorant):
    # a 2x2x2 grid of value 1 over [-1, 1]^3 with a two-entry table of opacity
    # 1: extinction sigma everywhere; one macro cell bounding it by `majorant`
    vol = abi.Volume()
    vol.world_to_local[:] = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]
    vol.origin[:] = (-1.0, -1.0, -1.0)
    vol.spacing[:] = (2.0, 2.0, 2.0)
    vol.dims[:] = (2, 2, 2)
    vol.tf_size = 2
    vol.value_lo, vol.value_hi = 0.0, 1.0
    vol.density_scale = sigma
    vol.majorant = majorant
    vol.linear = 1
    vol.macro_dims[:] = (1, 1, 1)
    s = two_quads()
    s.volumes = [vol]
    s.volume_values = np.ones(8, np.float32)
    s.volume_tf = np.ones(8, np.float32)
    s.volume_majorants = np.array([majorant], np.float32)
    return oracle.OracleScene(s)


def test_delta_tracking_draws_replayed():
    # a ray entering at t0 = 2 and leaving at t1 = 4 (origin x = -3, +x).
    # Delta tracking against the bound 1 (sigma 0.5): per step a draw u gives
    # t += -ln(1 - u); past t1 there is no event; otherwise a second draw
    # accepts the collision as a scattering event when it is below 0.5
    o = _homogeneous_slab(0.5, 1.0)
    seeds = _seeds(3000, 5)
    got = o.volume_scatter(np.array([-3.0, 0.1, 0.2], np.float32), np.array([1.0, 0, 0], np.float32), 0.0, 1e30,
                           seeds)
    events = 0
    for i, seed in enumerate(seeds):
        r = XorShift(int(seed))
        t, want = 2.0, -1.0
        while True:
            t = t - math.log(1 - r.f())
            if t >= 4.0:
                break
            if r.f() * 1.0 < 0.5:
                want = t
                break
        if want < 0:
            assert got[i] == -1.0, seed
        else:
            events += 1
            assert got[i] == pytest.approx(want, rel=2e-6), seed
    assert 500 < events < 2800


def test_ratio_tracking_draws_replayed():
    # ratio tracking against the bound 1 (sigma 0.5) halves the transmittance
    # at every collision (exactly, in float); below 0.05 Russian roulette
    # draws u and keeps tr = 0.05 when u 0.05 < tr, else ends at 0
    o = _homogeneous_slab(0.5, 1.0)
    seeds = _seeds(3000, 6)
    got = o.volume_transmittance(np.array([-3.0, 0.1, 0.2], np.float32), np.array([1.0, 0, 0], np.float32), 0.0,
                                 1e30, seeds)
    for i, seed in enumerate(seeds):
        r = XorShift(int(seed))
        t, tr = 2.0, 1.0
        while True:
            t = t - math.log(1 - r.f())
            if t >= 4.0:
                break
            tr *= 0.5
            if tr < 0.05:
                if r.f() * 0.05 >= tr:
                    tr = 0.0
                    break
                tr = 0.05
        assert got[i] == pytest.approx(tr, rel=1e-6), seed
    # the mean is exp(-sigma * 2) (ratio tracking is unbiased)
    assert got.mean() == pytest.approx(math.exp(-1.0), rel=0.03)


def _adaptive_cornell(w, h, spp, depth):
    # the Cornell box seen through a 70 degree field of view: the room fills
    # the middle of the image, the rest sees the black background
    s = scenes.cornell_box()
    c = s.camera
    cam = oracle.camera_perspective(c["look_from"], c["look_at"], c["vup"], 70.0, w / h)
    p = abi.default_render_params(w, h, spp, depth)
    return oracle.OracleScene(s), cam, p


def test_adaptive_sampling_stops_where_the_variance_rule_says():
    # pathtracer.cpp:676-723 and :1129-1139.  Three checks that do not use the
    # oracle's own loop: (1) a pixel that sees only the constant background has
    # zero variance, so it stops at exactly the 16 minimum samples; (2) the
    # stopping sample of every pixel, re-derived in float64 from its sample
    # values (one-sample renders at sample_offset k, independent of the
    # adaptive loop), agrees for all but marginal pixels; (3) a pixel that
    # stopped after n samples equals, bit for bit, a plain render of n samples
    # (the same sample streams summed in the same order).
    # (depth 1: direct light only, where a few pixels sit near the threshold)
    w, h, spp, depth = 48, 48, 48, 1
    o, cam, p = _adaptive_cornell(w, h, spp, depth)
    p.adaptive_sampling = 1
    col, ex, _ = o.render(cam, p, aovs=True, nsamp=True)
    ns = ex["nsamp"].astype(np.int64)
    miss = ex["depth"] >= 1e29
    assert 0.1 < miss.mean() < 0.9
    assert np.all(ns[miss] == 16)
    assert ns.min() == 16 and ns.max() <= spp
    assert (ns == spp).any()  # noisy pixels take every sample
    assert ((ns > 16) & (ns < spp)).any()  # the check runs before every later sample

    vals = []
    for k in range(spp):
        q = abi.default_render_params(w, h, 1, depth)
        q.sample_offset = k
        vals.append(o.render(cam, q)[0].astype(np.float64))
    vals = np.stack(vals)  # (spp, h, w, 3): each sample's radiance
    want = np.full((h, w), spp, np.int64)
    stopped = np.zeros((h, w), bool)
    ratio = np.zeros((spp, h, w))
    for s in range(16, spp):
        mean = vals[:s].sum(0) / s
        sq = (vals[:s] ** 2).sum(0) / s
        max_var = (sq - mean * mean).max(-1)
        lum = np.maximum(0.2126 * mean[..., 0] + 0.7152 * mean[..., 1] + 0.0722 * mean[..., 2], 1e-4)
        ratio[s] = max_var / (lum * lum)
        stop = ~stopped & (ratio[s] < 0.01)
        want[stop] = s
        stopped |= stop
    assert np.mean(ns == want) > 0.99
    # a pixel may decide otherwise only where its ratio sits at the threshold
    # (float32 sums against float64 ones) at the first sample they part at
    yy, xx = np.nonzero(ns != want)
    first = np.minimum(ns, want)[yy, xx]
    assert np.all(np.abs(ratio[first, yy, xx] / 0.01 - 1) < 0.02), ratio[first, yy, xx]
    assert ((want > 16) & (want < spp)).any()  # some pixels stop in between

    for n in np.unique(ns):
        q = abi.default_render_params(w, h, int(n), depth)
        plain = o.render(cam, q)[0]
        m = ns == n
        assert np.array_equal(col[m], plain[m]), n
