"""Shared test scenes and metrics."""
import os

import numpy as np

from photon_amd import abi, scenes


def single_triangle(with_attrs=True):
    """mesh_test.cpp:48-71: one triangle, vertex normals, texcoords, material 7."""
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    mats = [abi.default_material() for _ in range(8)]
    kw = {}
    if with_attrs:
        kw = dict(normals=np.array([[0, 0, 1], [0, 1, 0], [1, 0, 0]], np.float32),
                  texcoords=np.array([[0, 0], [1, 0], [0, 1]], np.float32),
                  material_ids=np.array([7], np.uint32))
    return scenes.Scene(pos, idx, mats, **kw)


def two_quads():
    """SceneBuilder::make_two_quads-like scene (builder.cpp:33-105): two
    parallel unit quads at z = 0 and z = -1 facing +z."""
    v0, t0 = scenes._quad((-0.5, -0.5, 0.0), (1, 0, 0), (0, 1, 0))
    v1, t1 = scenes._quad((-0.5, -0.5, -1.0), (1, 0, 0), (0, 1, 0))
    pos = np.concatenate([v0, v1]).astype(np.float32)
    idx = np.concatenate([t0, t1 + 4]).astype(np.uint32)
    return scenes.Scene(pos, idx, [scenes.diffuse((0.8, 0.8, 0.8))], material_ids=np.zeros(4, np.uint32))


def triangle_soup(n, seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-1, 1, (n, 1, 3))
    pos = (centers + rng.normal(scale=0.15 * scale, size=(n, 3, 3))).reshape(-1, 3).astype(np.float32)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)
    return scenes.Scene(pos, idx, [scenes.diffuse((0.5, 0.5, 0.5))], material_ids=np.zeros(n, np.uint32))


def rel_mse(x, ref):
    """Appendix-A metric: mean((x - ref)^2 / (ref^2 + 1e-2)) over channels."""
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))


def random_dirs(n, rng, upper=True):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    if upper:
        d[:, 2] = np.abs(d[:, 2])
    return d


def random_material(rng, delta=False):
    m = abi.default_material()
    m.base_color[:] = tuple(rng.uniform(0, 1, 3))
    m.metallic = float(rng.uniform() < 0.3) if rng.uniform() < 0.5 else float(rng.uniform())
    m.roughness = 0.0005 if delta else float(rng.uniform(0.02, 1.0))
    m.specular = float(rng.choice([0.0, 0.5, rng.uniform(0, 1)]))
    m.transmission = float(rng.uniform() < 0.2) * float(rng.uniform(0.5, 1.0))
    m.clearcoat = float(rng.uniform() < 0.5) * float(rng.uniform())
    m.clearcoat_roughness = float(rng.uniform(0.01, 0.5))
    m.sheen = float(rng.uniform() < 0.5) * float(rng.uniform(0, 0.5))
    m.sheen_tint = float(rng.uniform())
    m.specular_tint = float(rng.uniform() < 0.3) * float(rng.uniform())
    m.anisotropic = float(rng.uniform() < 0.5) * float(rng.uniform(0, 0.8))
    m.subsurface = float(rng.uniform() < 0.2) * float(rng.uniform())
    m.subsurface_color[:] = tuple(rng.uniform(0, 1, 3))
    m.ior = float(rng.uniform(1.2, 2.0))
    m.flags = abi.PHX_MATERIAL_FLAG_REFRACTIVE
    return m


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    """A known-answer fixture of tests/golden/ (written by make_golden.py)."""
    import json
    with open(os.path.join(GOLDEN, name)) as fh:
        return json.load(fh)


def golden_bsdf_inputs():
    """Materials and float32 directions (normal +z) of bsdf_closed_forms.json,
    with the expected f*cos and pdf."""
    from photon_amd import abi
    rows = golden("bsdf_closed_forms.json")
    mats = []
    for r in rows:
        m = abi.default_material()
        m.base_color[:] = tuple(r["base"])
        m.roughness = r["roughness"]
        if r["kind"] == "lambert":
            m.specular = 0.0
        else:
            m.metallic = 1.0
        mats.append(m)
    wo = np.array([r["wo"] for r in rows], np.float32)
    wi = np.array([r["wi"] for r in rows], np.float32)
    ng = np.tile(np.array([0.0, 0.0, 1.0], np.float32), (len(rows), 1))
    f = np.array([r["f"] for r in rows], np.float64)
    pdf = np.array([r["pdf"] for r in rows], np.float64)
    return mats, wo, wi, ng, f, pdf
