"""A pbrt-v4 scene loaded by phx_pbrt_load and rendered on the GPU, against
the oracle's render of the oracle's own reading of the same file
(oracle/pbrt.py): walls with an image texture, a gold and a glass sphere,
an emitting quad under the ceiling and a rotated environment image; same
bar as the other renders (relMSE < 1e-3 on the same sample streams, >= 90 %
of pixels equal to 1e-4 relative)."""
import numpy as np
import pytest

import oracle
from oracle import pbrt as op
from tests.helpers import rel_mse
from tests.test_pbrt import _pfm

pytestmark = pytest.mark.gpu

SCENE = """LookAt 0 1 3.6  0 1 0  0 1 0
Camera "perspective" "float fov" [ 45 ]
Film "rgb" "integer xresolution" [ 48 ] "integer yresolution" [ 40 ]
Sampler "halton" "integer pixelsamples" [ 8 ]
Integrator "volpath" "integer maxdepth" [ 6 ]
WorldBegin
AttributeBegin
  Rotate 60 0 1 0
  LightSource "infinite" "string filename" "sky.pfm" "rgb L" [ 0.8 0.8 0.8 ]
AttributeEnd
Texture "grid" "spectrum" "imagemap" "string filename" "grid.pfm"
MakeNamedMaterial "floor" "string type" "diffuse" "texture reflectance" "grid"
MakeNamedMaterial "wall" "string type" "diffuse" "rgb reflectance" [ 0.6 0.55 0.5 ]
MakeNamedMaterial "gold" "string type" "conductor" "spectrum eta" [ "metal-Au-eta" ] "spectrum k" [ "metal-Au-k" ] "float roughness" 0.15
MakeNamedMaterial "glass" "string type" "dielectric" "float eta" 1.5
NamedMaterial "floor"
Shape "trianglemesh" "point3 P" [ -1.5 0 -1.5  1.5 0 -1.5  1.5 0 1.5  -1.5 0 1.5 ] "integer indices" [ 0 2 1 0 3 2 ]
  "point2 uv" [ 0 0  2 0  2 2  0 2 ]
NamedMaterial "wall"
Shape "trianglemesh" "point3 P" [ -1.5 0 -1.5  1.5 0 -1.5  1.5 2.5 -1.5  -1.5 2.5 -1.5 ] "integer indices" [ 0 1 2 0 2 3 ]
AttributeBegin
  NamedMaterial "gold"
  Transform [ 1 0 0 0  0 1 0 0  0 0 1 0  -0.5 0.45 -0.3 1 ]
  Scale 0.45 0.45 0.45
  Shape "sphere" "float radius" 1
AttributeEnd
AttributeBegin
  NamedMaterial "glass"
  Transform [ 1 0 0 0  0 1 0 0  0 0 1 0  0.55 0.35 0.3 1 ]
  Shape "sphere" "float radius" 0.35
AttributeEnd
AttributeBegin
  AreaLightSource "diffuse" "rgb L" [ 12 11 9 ]
  Shape "trianglemesh" "point3 P" [ -0.4 2.2 -0.4  0.4 2.2 -0.4  0.4 2.2 0.4  -0.4 2.2 0.4 ] "integer indices" [ 0 1 2 0 2 3 ]
AttributeEnd
"""


def test_pbrt_scene_renders_like_the_oracle(gpu_ctx, tmp_path):
    from photon_amd.backend import HipBackend, PathTracer
    from photon_amd.pbrt import load_pbrt
    rng = np.random.default_rng(8)
    grid = np.where((np.add.outer(np.arange(8), np.arange(8)) % 2)[..., None] > 0, 0.8, 0.15) * np.ones(3)
    _pfm(tmp_path / "grid.pfm", grid.astype(np.float32))
    sky = np.linspace(0.2, 1.5, 16)[:, None, None] * np.ones((16, 32, 3)) + rng.uniform(0, 0.2, (16, 32, 3))
    _pfm(tmp_path / "sky.pfm", sky.astype(np.float32))
    path = tmp_path / "scene.pbrt"
    path.write_text(SCENE)
    sc = load_pbrt(str(path))
    assert sc.info.textures == 1 and sc.desc().env_width == 32
    pt = PathTracer(HipBackend(context=gpu_ctx))
    pt.set_scene(sc)
    pt.camera, pt.params = sc.camera, sc.render_params()
    res = pt.render()
    conv = op.convert(op.parse(str(path)))
    s = op.scene_of(conv, "pbrt_scene")
    ref, ex, _ = oracle.OracleScene(s).render(conv["camera"], sc.render_params(), aovs=True)
    err = rel_mse(res.color, ref)
    same = np.mean(np.abs(res.color - ref) <= 1e-4 * (1 + np.abs(ref)))
    print(f"pbrt scene: relMSE {err:.3e}, pixels equal {same:.4f}, mean {ref.mean():.4f}", flush=True)
    assert ref.mean() > 0.05
    assert err < 1e-3 and same >= 0.9
    np.testing.assert_allclose(res.alpha, ex["alpha"], atol=1e-5)
