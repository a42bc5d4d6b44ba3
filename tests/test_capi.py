"""The C ABI library: loads, exports every entry point of include/phx.h, its
struct layouts match the header, and its host-only helpers (camera set-up,
volume majorants) agree with the oracle.  No kernel is launched here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from photon_amd import abi
from photon_amd._lib import LIB_PATH, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "phx.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(phx_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib()
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (phx_[a-z_]+)\b", out))
    declared = declared_functions()
    assert len(declared) >= 25
    missing = [f for f in declared if f not in exported]
    assert not missing, f"declared but not exported: {missing}"
    extra = sorted(exported - set(declared))
    assert not extra, f"exported but not declared: {extra}"


def test_struct_layouts_match_header():
    abi.check_layout()
    # offsets of pt::Material / pt::TfVolume (material.h:35-82, tf_volume.h:14-33)
    assert abi.Material.emission.offset == 56
    assert abi.Material.base_color_src.offset == 124
    assert abi.Material.flags.offset == 188
    assert abi.Volume.macro_dims.offset == 120
    assert C.sizeof(abi.SceneDesc) > 0


def test_abi_types_equal_the_reference_declarations():
    # tests/layout/reference_types.cpp restates the reference's boundary types
    # member by member and static_asserts every offset and size against
    # include/phx.h: compiling it is the check
    src = os.path.join(ROOT, "tests", "layout", "reference_types.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_layout_check_catches_a_planted_mistake(tmp_path):
    # the same check against a header whose HitResult bools are widened to
    # 32 bits must fail to compile
    bad = tmp_path / "include"
    bad.mkdir()
    text = open(HEADER).read().replace("  uint8_t hit;\n  uint8_t has_interpolated_color;\n  uint8_t pad_[2];",
                                        "  uint32_t hit;\n  uint32_t has_interpolated_color;\n  uint32_t pad_;")
    assert "uint32_t has_interpolated_color" in text
    (bad / "phx.h").write_text(text)
    layout = tmp_path / "tests" / "layout"
    layout.mkdir(parents=True)
    src = open(os.path.join(ROOT, "tests", "layout", "reference_types.cpp")).read()
    (layout / "reference_types.cpp").write_text(src)
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", str(layout / "reference_types.cpp")],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "phx_hit" in r.stderr


def test_backend_name_and_abi_version():
    assert lib().phx_backend_name() == b"hip"
    assert lib().phx_abi_version() == 3


def test_errors_without_gpu_or_arguments():
    # no context: invalid argument, not a crash
    assert lib().phx_create(None, 0) == -1
    assert lib().phx_set_tuning(None, None) == -1


def test_camera_matches_oracle_bitwise():
    from photon_amd.backend import make_camera
    for args in (((0, 1, 3.7), (0, 1, 0), (0, 1, 0), 40.0, 1.0), ((1, 2, 3), (-1, 0.5, 0.2), (0, 1, 0), 63.0, 1.5)):
        a = make_camera(*args)
        b = oracle.camera_perspective(*args)
        assert bytes(a) == bytes(b)


def test_orthographic_camera_matches_oracle_bitwise():
    from oracle.scene_data import _orthographic
    from photon_amd.backend import make_orthographic_camera
    for f, at, up, h, asp in (((0, 1, 3.7), (0, 1, 0), (0, 1, 0), 2.4, 1.2), ((1, 2, 3), (-1, 0.5, 0.2), (0, 1, 0), 0.7, 0.5)):
        a = make_orthographic_camera(f, at, up, h, asp)
        b = _orthographic(np.array(f, np.float32), np.array(at, np.float32), np.array(up, np.float32), h, asp)
        assert bytes(a) == bytes(b)
        # closed form: the view plane is centred on the eye, h tall, h*aspect wide
        assert a.is_ortho == 1 and a.lens_radius == 0
        assert np.isclose(np.linalg.norm(a.vertical), h, rtol=1e-6)
        assert np.isclose(np.linalg.norm(a.horizontal), h * asp, rtol=1e-6)
        centre = np.array(a.lower_left) + 0.5 * np.array(a.horizontal) + 0.5 * np.array(a.vertical)
        assert np.allclose(centre, f, atol=1e-6)


def test_volume_majorants_match_oracle():
    rng = np.random.default_rng(0)
    for size, scale in ((9, 2.0), (40, 20.0)):
        v1 = abi.Volume()
        v1.world_to_local[:] = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]
        v1.origin[:] = (0, 0, 0)
        v1.spacing[:] = (1, 1, 1)
        v1.dims[:] = (size, size, size)
        v1.value_lo, v1.value_hi = 0.0, 1.0
        v1.density_scale = scale
        v1.linear = 1
        v2 = abi.Volume.from_buffer_copy(v1)
        values = rng.uniform(0, 1, size ** 3).astype(np.float32)
        tf = rng.uniform(0, 1, (256, 4)).astype(np.float32)
        ref = oracle.volume_majorants(v1, values, tf)
        n = C.c_uint32()
        vp = values.ctypes.data_as(C.POINTER(C.c_float))
        tp = tf.ctypes.data_as(C.POINTER(C.c_float))
        assert lib().phx_volume_majorants(C.byref(v2), vp, tp, 256, None, 0, C.byref(n)) == 0
        out = np.zeros(n.value, np.float32)
        assert lib().phx_volume_majorants(C.byref(v2), vp, tp, 256, out.ctypes.data_as(C.POINTER(C.c_float)), n.value,
                                          C.byref(n)) == 0
        np.testing.assert_array_equal(out, ref)
        assert list(v2.macro_dims) == list(v1.macro_dims)
        assert v2.majorant == v1.majorant
    # too small an output buffer is an error, not an overrun
    assert lib().phx_volume_majorants(C.byref(v2), vp, tp, 256, out.ctypes.data_as(C.POINTER(C.c_float)), 1,
                                      C.byref(n)) == -4


def test_backend_factory_names():
    from photon_amd.backend import create_best_backend
    os.environ["PHOTON_BACKEND"] = "embree"
    try:
        with pytest.raises(RuntimeError, match="not part of this package"):
            create_best_backend()
        os.environ["PHOTON_BACKEND"] = "nonsense"
        with pytest.raises(RuntimeError, match="unknown backend"):
            create_best_backend()
    finally:
        del os.environ["PHOTON_BACKEND"]


def test_scene_generators():
    from photon_amd import scenes
    s = scenes.cornell_box()
    assert s.triangle_count == 36
    g = scenes.sphere_grid(level=2, grid=2, env_size=(32, 16))
    assert g.triangle_count == 4 * 320 + 2 + 8
    assert g.env_pixels.shape == (16, 32, 3)


def test_ctypes_layouts_equal_the_c_header(tmp_path):
    # sizeof / offsetof from the C compiler against the ctypes mirror
    structs = {"phx_material": abi.Material, "phx_light": abi.Light, "phx_texture_info": abi.TextureInfo,
               "phx_volume": abi.Volume, "phx_scene_desc": abi.SceneDesc, "phx_camera": abi.Camera,
               "phx_render_params": abi.RenderParams, "phx_hit": abi.Hit, "phx_hit_compact": abi.HitCompact,
               "phx_tuning": abi.Tuning, "phx_stats": abi.Stats, "phx_view": abi.View,
               "phx_camera_data": abi.CameraData, "phx_render_settings": abi.RenderSettings,
               "phx_bridge_material": abi.BridgeMaterial, "phx_sampler_info": abi.SamplerInfo,
               "phx_param_source": abi.ParamSource, "phx_light_data": abi.LightData,
               "phx_sampler_data": abi.SamplerData, "phx_environment_data": abi.EnvironmentData,
               "phx_volume_data": abi.VolumeData, "phx_scene_data": abi.SceneData, "phx_display": abi.Display,
               "phx_pbrt_info": abi.PbrtInfo}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "phx.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'  printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("  return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(l.rsplit(" ", 1) for l in out if l)
    for cname, cls in structs.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(cls, fname).offset, f"{cname}.{fname}"


def _fields(m):
    out = {}
    for name, _ in m._fields_:
        v = getattr(m, name)
        if isinstance(v, C.Structure):
            v = tuple(_fields(v).items())
        out[name] = tuple(v) if hasattr(v, "__len__") else v
    return out


def test_convert_material_follows_the_bridge():
    # convert_material (RenderBridge.cpp:324-370) on a material that sets every
    # field the bridge reads
    from photon_amd.renderer import convert_material
    f32 = np.float32
    bm = abi.bridge_material(base_color=(0.2, 0.3, 0.4), opacity=0.35, metallic=0.25, roughness=0.6, ior=1.33,
                             transmission=0.7, specular=0.8, emission=(1.0, 2.0, 3.0), clearcoat=0.4,
                             clearcoat_roughness=0.1, sheen=0.3, alpha_mode=2, alpha_cutoff=0.25, thin_walled=1,
                             emitter_only=1, hidden_to_camera=1, emission_side=1)
    m = convert_material(bm)
    assert m.roughness == f32(0.6) * f32(0.6)          # the BSDF takes the GGX alpha
    assert m.specular == f32(0.5) * f32(0.8)           # 0.5 is the full dielectric reflectance
    assert tuple(m.base_color) == tuple(f32([0.2, 0.3, 0.4]))
    assert tuple(m.emission) == (1.0, 2.0, 3.0) and m.emission_strength == 1.0
    assert (m.metallic, m.ior, m.transmission) == (f32(0.25), f32(1.33), f32(0.7))
    assert (m.clearcoat, m.clearcoat_roughness, m.sheen) == (f32(0.4), f32(0.1), f32(0.3))
    assert m.alpha == f32(0.35) and m.alpha_mode == abi.PHX_ALPHA_MASK and m.alpha_cutoff == 0.25
    assert m.thin == 1
    assert m.flags == (abi.PHX_MATERIAL_FLAG_REFRACTIVE | abi.PHX_MATERIAL_FLAG_EMITTER_ONLY
                       | abi.PHX_MATERIAL_FLAG_HIDDEN_TO_CAMERA | abi.PHX_MATERIAL_FLAG_EMIT_BACK_ONLY)
    # fields the bridge leaves alone keep pt::Material's defaults (material.h:35-82)
    d = abi.default_material()
    for name in ("sheen_tint", "specular_tint", "anisotropic", "subsurface", "subsurface_radius"):
        assert getattr(m, name) == getattr(d, name), name
    assert tuple(m.subsurface_color) == (1.0, 1.0, 1.0)
    assert (m.base_color_tex, m.normal_tex, m.roughness_tex, m.metallic_tex, m.emission_tex, m.alpha_tex) == (-1,) * 6
    # alpha-mode and emission-side codings
    modes = {0: abi.PHX_ALPHA_BLEND, 1: abi.PHX_ALPHA_OPAQUE, 2: abi.PHX_ALPHA_MASK, 7: abi.PHX_ALPHA_OPAQUE}
    for code, want in modes.items():
        assert convert_material(abi.bridge_material(alpha_mode=code)).alpha_mode == want, code
    sides = {0: 0, 1: abi.PHX_MATERIAL_FLAG_EMIT_BACK_ONLY, 2: abi.PHX_MATERIAL_FLAG_EMIT_BOTH_SIDES}
    for side, bit in sides.items():
        assert convert_material(abi.bridge_material(emission_side=side)).flags == abi.PHX_MATERIAL_FLAG_REFRACTIVE | bit


def test_convert_material_copies_parameter_sources():
    # convert_source (RenderBridge.cpp:314-322): kind 1 attribute, 2 sampler,
    # anything else constant; the index is kept.  Sourced roughness is
    # perceptual (roughness_src_squared = 1); a sourced opacity scales a
    # neutral constant (alpha = 1).
    from photon_amd.renderer import convert_material
    names = {"base_color_src": "base_color_src", "opacity_src": "alpha_src", "metallic_src": "metallic_src",
             "roughness_src": "roughness_src", "emission_src": "emission_src", "transmission_src": "transmission_src",
             "normal_src": "normal_src"}
    for src, dst in names.items():
        for kind, want in ((1, abi.PHX_SOURCE_ATTRIBUTE), (2, abi.PHX_SOURCE_SAMPLER), (5, abi.PHX_SOURCE_CONSTANT)):
            m = convert_material(abi.bridge_material(opacity=0.4, **{src: (kind, 3)}))
            assert (getattr(m, dst).kind, getattr(m, dst).index) == (want, 3), (src, kind)
            assert m.alpha == (1.0 if (src == "opacity_src" and want != 0) else np.float32(0.4)), (src, kind)
            assert m.roughness_src_squared == 1
    assert convert_material(abi.bridge_material(base_color_alpha=1)).base_color_alpha == 1
    assert convert_material(abi.bridge_material()).base_color_alpha == 0


def test_scene_materials_are_converted_bridge_materials():
    # the scenes' diffuse and emitter materials are MaterialData after convert_material
    from photon_amd import scenes
    from photon_amd.renderer import convert_material
    rgb = (0.73, 0.2, 0.05)
    assert _fields(scenes.diffuse(rgb)) == _fields(convert_material(abi.bridge_material(base_color=rgb,
                                                                                          emission_side=0)))
    e = convert_material(abi.bridge_material(base_color=(0.0, 0.0, 0.0), emission=(17.0, 12.0, 4.0), emitter_only=1,
                                             emission_side=0))
    assert _fields(scenes.emitter((17.0, 12.0, 4.0))) == _fields(e)
