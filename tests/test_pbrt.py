"""pbrt-v4 scene files: phx_pbrt_load (the library's reader and conversion)
against the oracle's restatement of parse_pbrt_file / convert_pbrt_scene
(oracle/pbrt.py) on hand-written scenes, bit for bit, and both pinned by
what the files fix: translations, scales and rotations of known vertices,
the tessellated sphere's radius and counts, the conductor's Fresnel F0 in
closed form, the LookAt camera's frame, ASCII and binary PLY giving the same
mesh, the environment's scale, the instancing limits.  Host code only; the
GPU render of a loaded scene is in test_gpu_pbrt.py."""
import struct

import numpy as np
import pytest

from oracle import pbrt as op
from photon_amd.pbrt import load_pbrt

FIELDS = ("positions", "indices", "normals", "texcoords", "material_ids", "emissive_ids", "emissive_areas",
          "texture_pixels", "env_pixels", "env_world_to_tex")


def load_both(path, use_instancing=False):
    s = load_pbrt(str(path), use_instancing)
    got = s.arrays()
    ref = op.convert(op.parse(str(path), use_instancing))
    for k in FIELDS:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    assert tuple(got["env_size"]) == tuple(ref["env_size"])
    assert len(got["materials"]) == len(ref["materials"])
    for i, (a, b) in enumerate(zip(got["materials"], ref["materials"])):
        assert bytes(a) == bytes(b), f"material {i}"
    assert len(got["lights"]) == len(ref["lights"])
    for a, b in zip(got["lights"], ref["lights"]):
        assert bytes(a) == bytes(b)
    assert got["texture_infos"] == ref["texture_infos"]
    assert bytes(s.camera) == bytes(ref["camera"])
    assert (s.info.width, s.info.height, s.info.spp, s.info.max_depth) == ref["info"]
    return s, got


QUAD = '"point3 P" [ -1 0 -1  1 0 -1  1 0 1  -1 0 1 ] "integer indices" [ 0 1 2 0 2 3 ]'


def write(path, text):
    path.write_text(text)
    return path


def test_basic_scene_and_settings(tmp_path):
    f = write(tmp_path / "a.pbrt", f"""# comment
LookAt 0 1 4  0 1 0  0 1 0
Camera "perspective" "float fov" [ 40 ]
Film "rgb" "integer xresolution" [ 64 ] "integer yresolution" [ 48 ]
Sampler "halton" "integer pixelsamples" 16
Integrator "path" "integer maxdepth" [ 5 ]
WorldBegin
MakeNamedMaterial "white" "string type" [ "diffuse" ] "rgb reflectance" [ 0.7 0.7 0.7 ]
AttributeBegin
  NamedMaterial "white"
  Shape "trianglemesh" {QUAD}
AttributeEnd
AttributeBegin
  AreaLightSource "diffuse" "rgb L" [ 10 10 10 ]
  Transform [ 1 0 0 0  0 1 0 0  0 0 1 0  0 1.9 0 1 ]
  Shape "trianglemesh" "point3 P" [ -0.3 0 -0.3  0.3 0 -0.3  0.3 0 0.3  -0.3 0 0.3 ] "integer indices" [ 0 2 1 0 3 2 ]
AttributeEnd
""")
    s, a = load_both(f)
    assert (s.info.width, s.info.height, s.info.spp, s.info.max_depth) == (64, 48, 16, 5)
    # every triangle has vertices of its own; the light quad moved up by 1.9
    assert a["positions"].shape == (12, 3) and a["indices"].tolist() == [[3 * t, 3 * t + 1, 3 * t + 2] for t in range(4)]
    np.testing.assert_array_equal(a["positions"][6:, 1], np.float32(1.9))
    # the emitter's own material (black base, emission 10, strength 1)
    m = a["materials"][a["material_ids"][2]]
    assert list(m.emission) == [10, 10, 10] and m.emission_strength == 1 and list(m.base_color) == [0, 0, 0]
    assert a["emissive_ids"].tolist() == [2, 3]
    np.testing.assert_allclose(a["emissive_areas"], 0.18, rtol=1e-6)  # half of a 0.6 x 0.6 quad
    # the LookAt camera: eye at (0,1,4), w = normalize(eye - at)
    assert list(s.camera.origin) == [0, 1, 4] and list(s.camera.w) == [0, 0, 1]
    assert len(a["lights"]) == 1 and a["lights"][0].mesh_prim_count == 2


def test_transforms_pin(tmp_path):
    # Transform is a column-major matrix (the translation in its last four
    # values); Scale multiplies it on the left in matrix terms, as the
    # reference's index product does; Translate is not a directive the
    # reference reads and changes nothing; Rotate builds its matrix row by
    # row while the conversion reads the 16 values as columns, so Rotate 90
    # about z turns by -90 degrees: x goes to -y, y to x
    tri = '"point3 P" [ 1 0 0  0 1 0  0 0 1 ] "integer indices" [ 0 1 2 ]'
    f = write(tmp_path / "t.pbrt", f"""WorldBegin
AttributeBegin
  Transform [ 1 0 0 0  0 1 0 0  0 0 1 0  1 2 3 1 ]
  Scale 2 2 2
  Translate 5 5 5
  Shape "trianglemesh" {tri}
AttributeEnd
AttributeBegin
  Rotate 90 0 0 1
  Shape "trianglemesh" {tri}
AttributeEnd
AttributeBegin
  ConcatTransform [ 1 0 0 0  0 1 0 0  0 0 1 0  0 0 -4 1 ]
  ConcatTransform [ 3 0 0 0  0 3 0 0  0 0 3 0  0 0 0 1 ]
  Shape "trianglemesh" {tri}
AttributeEnd
""")
    _, a = load_both(f)
    p = a["positions"].astype(np.float64)
    base = np.eye(3)
    np.testing.assert_array_equal(p[0:3], 2 * (base + [1, 2, 3]))
    np.testing.assert_allclose(p[3:6], [[0, -1, 0], [1, 0, 0], [0, 0, 1]], atol=1e-6)
    # ConcatTransform post-multiplies: the scale acts first, then the translation
    np.testing.assert_array_equal(p[6:9], 3 * base + [0, 0, -4])


def test_sphere_tessellation(tmp_path):
    f = write(tmp_path / "s.pbrt", 'WorldBegin\nShape "sphere" "float radius" 0.75\n')
    _, a = load_both(f)
    assert a["indices"].shape == (16 * 32 * 2, 3)
    r = np.linalg.norm(a["positions"].astype(np.float64), axis=1)
    np.testing.assert_allclose(r, 0.75, rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(np.linalg.norm(a["normals"].astype(np.float64), axis=1), 1.0, rtol=1e-6)
    # no material: the reference's fallback (grey 0.5, roughness 1)
    assert len(a["materials"]) == 1 and list(a["materials"][0].base_color) == [0.5] * 3


def test_materials_of_every_type(tmp_path):
    f = write(tmp_path / "m.pbrt", f"""WorldBegin
MakeNamedMaterial "au" "string type" "conductor" "spectrum eta" [ "metal-Au-eta" ] "spectrum k" [ "metal-Au-k" ] "float uroughness" 0.1 "float vroughness" 0.3
MakeNamedMaterial "cu_unbracketed" "string type" "conductor" "spectrum eta" "metal-Cu-eta" "float roughness" 0.2
MakeNamedMaterial "glass" "string type" "dielectric" "float eta" 1.33 "float roughness" 0.05
MakeNamedMaterial "coat" "string type" "coateddiffuse" "rgb reflectance" [ 0.2 0.3 0.4 ] "float roughness" 0.4
MakeNamedMaterial "leaf" "string type" "diffusetransmission" "rgb reflectance" [ 0.1 0.6 0.1 ]
MakeNamedMaterial "other" "string type" "measured" "float roughness" 0.7
Material "diffuse" "rgb reflectance" [ 1 0 0 ]
NamedMaterial "au"
Shape "trianglemesh" {QUAD}
NamedMaterial "nosuch"
Shape "trianglemesh" {QUAD}
""")
    _, a = load_both(f)
    names = sorted(["au", "cu_unbracketed", "glass", "coat", "leaf", "other"])
    m = {n: a["materials"][i] for i, n in enumerate(names)}
    eta, k = np.array([0.143, 0.374, 1.442]), np.array([3.983, 2.387, 1.603])
    f0 = ((eta - 1) ** 2 + k ** 2) / ((eta + 1) ** 2 + k ** 2)
    np.testing.assert_allclose(list(m["au"].base_color), f0, rtol=1e-6)
    assert m["au"].metallic == 1 and abs(m["au"].roughness - 0.2) < 1e-7  # (0.1 + 0.3) / 2
    # a spectrum name outside brackets is not taken (the reference reads it
    # as a parameter of its own): eta and k stay 0, so F0 = 1
    assert list(m["cu_unbracketed"].base_color) == [1, 1, 1]
    assert abs(m["glass"].ior - 1.33) < 1e-7 and m["glass"].transmission == 1
    assert m["coat"].clearcoat == 1 and abs(m["coat"].clearcoat_roughness - 0.2) < 1e-7
    assert m["leaf"].transmission == 0.5 and m["leaf"].roughness == 1
    assert abs(m["other"].roughness - 0.7) < 1e-7 and list(m["other"].base_color) == [0.5] * 3
    # the anonymous Material is skipped; an unknown name falls back to material 0
    assert len(a["materials"]) == 6
    assert a["material_ids"].tolist() == [names.index("au")] * 2 + [0] * 2


def _ply(path, binary, quads=True):
    v = [(-1, 0, -1, 0, 1, 0, 0, 0), (1, 0, -1, 0, 1, 0, 1, 0), (1, 0, 1, 0, 1, 0, 1, 1), (-1, 0, 1, 0, 1, 0, 0, 1),
         (0, 1, 0, 0, 1, 0, 0.5, 0.5)]
    faces = [[0, 1, 2, 3], [0, 4, 1]]
    head = ["ply", f"format {'binary_little_endian' if binary else 'ascii'} 1.0", "comment test",
            f"element vertex {len(v)}"] + [f"property float {n}" for n in ("x", "y", "z", "nx", "ny", "nz", "u", "v")] + \
        [f"element face {len(faces)}", "property list uchar int vertex_indices", "end_header"]
    with open(path, "wb") as fh:
        fh.write(("\n".join(head) + "\n").encode())
        if binary:
            for r in v:
                fh.write(struct.pack("<8f", *r))
            for fc in faces:
                fh.write(struct.pack("<B", len(fc)) + struct.pack(f"<{len(fc)}i", *fc))
        else:
            for r in v:
                fh.write((" ".join(str(x) for x in r) + "\n").encode())
            for fc in faces:
                fh.write((f"{len(fc)} " + " ".join(str(x) for x in fc) + "\n").encode())


def test_ply_ascii_and_binary(tmp_path):
    _ply(tmp_path / "a.ply", False)
    _ply(tmp_path / "b.ply", True)
    res = []
    for name in ("a.ply", "b.ply"):
        f = write(tmp_path / f"{name}.pbrt", f'WorldBegin\nScale 1 2 1\nShape "plymesh" "string filename" "{name}"\n')
        res.append(load_both(f)[1])
    for k in ("positions", "normals", "texcoords", "indices"):
        np.testing.assert_array_equal(res[0][k], res[1][k])
    # the quad fans into two triangles, the triangle stays one
    assert res[0]["indices"].shape == (3, 3)
    assert res[0]["texcoords"][:3].tolist() == [[0, 0], [1, 0], [1, 1]]
    np.testing.assert_array_equal(res[0]["positions"][3 * 2 + 1], [0, 2, 0])  # y scaled by 2


def _pfm(path, img):
    h, w = img.shape[:2]
    with open(path, "wb") as fh:
        fh.write(f"PF\n{w} {h}\n-1.0\n".encode())
        fh.write(img[::-1].astype("<f4").tobytes())  # rows bottom first


def test_textures_alpha_and_environment(tmp_path):
    rng = np.random.default_rng(3)
    tex = rng.uniform(0, 1, (4, 5, 3)).astype(np.float32)
    env = rng.uniform(0, 2, (8, 16, 3)).astype(np.float32)
    _pfm(tmp_path / "t.pfm", tex)
    _pfm(tmp_path / "env.pfm", env)
    _ply(tmp_path / "leaf.ply", False)
    f = write(tmp_path / "x.pbrt", f"""WorldBegin
Texture "wood" "spectrum" "imagemap" "string filename" "t.pfm"
Texture "cut" "float" "imagemap" "string filename" "t.pfm"
Texture "gone" "spectrum" "imagemap" "string filename" "missing.png"
MakeNamedMaterial "m" "string type" "diffuse" "texture reflectance" "wood"
AttributeBegin
  Rotate 30 0 1 0
  LightSource "infinite" "string filename" "env.pfm" "rgb L" [ 1 2 3 ]
AttributeEnd
NamedMaterial "m"
Shape "plymesh" "string filename" "leaf.ply" "texture alpha" "cut"
""")
    _, a = load_both(f)
    # textures by name: "cut" then "wood" (the missing file is left out)
    assert a["texture_infos"] == [(0, 5, 4), (20, 5, 4)]
    np.testing.assert_array_equal(a["texture_pixels"][:20], tex.reshape(-1, 3))
    # the environment times the mean of L, rows from the top
    np.testing.assert_array_equal(a["env_pixels"].reshape(8, 16, 3), env * np.float32(2))
    w2t = a["env_world_to_tex"].reshape(3, 3).astype(np.float64)
    # the light's transform turns by -30 degrees about y (Rotate's matrix read
    # as columns), so world-to-map, its inverse, turns by +30
    c, s = np.cos(np.radians(30)), np.sin(np.radians(30))
    np.testing.assert_allclose(w2t, np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]]), atol=1e-6)
    # the alpha-textured mesh gets a copy of its material with alpha_tex
    mid = a["material_ids"][0]
    assert a["materials"][mid].alpha_tex == 0 and a["materials"][mid].base_color_tex == 1


def test_instances_include_and_camera_transform(tmp_path):
    write(tmp_path / "geo.pbrt", 'ObjectBegin "tri"\nShape "trianglemesh" "point3 P" [ 0 0 0 1 0 0 0 1 0 ] '
                                 '"integer indices" [ 0 1 2 ]\nObjectEnd\n')
    insts = "\n".join(f"AttributeBegin\nTransform [ 1 0 0 0 0 1 0 0 0 0 1 0 {i} 0 0 1 ]\nObjectInstance \"tri\"\n"
                      f"AttributeEnd" for i in range(502))
    f = write(tmp_path / "i.pbrt", f"""Transform [ 1 0 0 0  0 1 0 0  0 0 1 0  0 -1 -5 1 ]
Camera "perspective" "float fov" 50 "float lensradius" 0.1 "float focaldistance" 4.5
WorldBegin
Include "geo.pbrt"
ObjectInstance "none"
{insts}
""")
    s, a = load_both(f)
    # at most 500 instances of one object are placed
    assert s.info.instances == 502 and s.info.objects == 1 and s.info.meshes == 500
    assert a["indices"].shape == (500, 3)
    np.testing.assert_array_equal(a["positions"][3 * 499], [499, 0, 0])
    # the camera from the inverse of the world-to-camera transform
    assert list(s.camera.origin) == [0, 1, 5]
    assert abs(s.camera.lens_radius - 0.1) < 1e-7 and abs(s.camera.focus_dist - 4.5) < 1e-6
    # with use_instancing the instances are counted, not placed
    s2, a2 = load_both(f, use_instancing=True)
    assert s2.info.meshes == 0 and a2["indices"].shape == (1, 3)


def test_missing_file_is_an_empty_scene(tmp_path):
    s, a = load_both(tmp_path / "nope.pbrt")
    # one degenerate triangle stands in for the empty mesh
    assert a["indices"].shape == (1, 3) and not a["positions"].any()
    assert (s.info.width, s.info.height) == (512, 512)


def test_oracle_catches_a_planted_mistake(tmp_path, monkeypatch):
    # the comparison fails when the oracle's ConcatTransform order is flipped
    f = write(tmp_path / "c.pbrt", 'WorldBegin\nConcatTransform [ 1 0 0 0 0 1 0 0 0 0 1 0 0 0 -4 1 ]\n'
                                   'ConcatTransform [ 3 0 0 0 0 3 0 0 0 0 3 0 0 0 0 1 ]\n'
                                   'Shape "trianglemesh" "point3 P" [ 1 0 0 0 1 0 0 0 1 ] "integer indices" [ 0 1 2 ]\n')
    real = op.mul16
    monkeypatch.setattr(op, "mul16", lambda a, b: real(b, a))
    with pytest.raises(AssertionError):
        load_both(f)
