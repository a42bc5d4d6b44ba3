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
    # offsets that the kernels rely on
    assert abi.Material.emission.offset == 64
    assert abi.Material.flags.offset == 100
    assert abi.Volume.macro_dims.offset == 120
    assert C.sizeof(abi.SceneDesc) > 0


def test_backend_name_and_abi_version():
    assert lib().phx_backend_name() == b"hip"
    assert lib().phx_abi_version() == 2


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
               "phx_camera_data": abi.CameraData, "phx_render_settings": abi.RenderSettings}
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
