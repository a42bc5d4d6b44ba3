"""Loader of the HIP library libphx.so (built by build_lib.py, in-tree).

There is no fallback: a missing library or GPU raises, so nothing can silently
run a CPU path in place of the kernels.
"""
import ctypes as C
import os

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libphx.so")

_lib = None


class PhxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"phx error {code}: {msg}")
        self.code = code


def _sig(lib):
    P = C.c_void_p
    PF = C.POINTER(C.c_float)
    PU = C.POINTER(C.c_uint32)
    sigs = {
        "phx_create": (C.c_int, [C.POINTER(P), C.c_int]),
        "phx_destroy": (None, [P]),
        "phx_last_error": (C.c_char_p, [P]),
        "phx_backend_name": (C.c_char_p, []),
        "phx_abi_version": (C.c_int, []),
        "phx_set_tuning": (C.c_int, [P, C.POINTER(abi.Tuning)]),
        "phx_get_tuning": (C.c_int, [P, C.POINTER(abi.Tuning)]),
        "phx_get_stats": (C.c_int, [P, C.POINTER(abi.Stats)]),
        "phx_build_accel": (C.c_int, [P, C.POINTER(abi.SceneDesc)]),
        "phx_set_scene": (C.c_int, [P, C.POINTER(abi.SceneDesc)]),
        "phx_trace_closest": (C.c_int, [P, P, P, P, P, C.c_uint32, P, P]),
        "phx_trace_closest_compact": (C.c_int, [P, P, P, P, P, C.c_uint32, P, P]),
        "phx_trace_occluded": (C.c_int, [P, P, P, P, P, C.c_uint32, P, P]),
        "phx_trace_closest_noalloc": (C.c_int, [P, P, P, P, P, C.c_uint32, P, P]),
        "phx_display_transform": (C.c_int, [P, P, C.c_uint32, C.c_uint64, C.POINTER(abi.Display), P, P, P]),
        "phx_denoiser_is_available": (C.c_int, []),
        "phx_denoise_buffer": (C.c_int, [P, P, P, C.c_uint32, C.c_uint32]),
        "phx_pbrt_load": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(P)]),
        "phx_pbrt_scene_desc": (C.c_int, [P, C.POINTER(abi.SceneDesc), C.POINTER(abi.Camera),
                                          C.POINTER(abi.PbrtInfo)]),
        "phx_pbrt_free": (None, [P]),
        "phx_synchronize": (C.c_int, [P]),
        "phx_bsdf_eval": (C.c_int, [P, P, P, P, P, P, C.c_uint32, P, P, P]),
        "phx_bsdf_sample": (C.c_int, [P, P, P, P, P, P, C.c_uint32, P, P, P, P, P]),
        "phx_bsdf_eval_grad": (C.c_int, [P, P, P, P, P, P, C.c_uint32, P, P, P]),
        "phx_adjoint": (C.c_int, [P, C.POINTER(abi.Camera), C.POINTER(abi.RenderParams), P, P, P]),
        "phx_set_material_params": (C.c_int, [P, P, C.c_uint32]),
        "phx_set_texture_pixels": (C.c_int, [P, P, C.c_uint32, C.c_uint32]),
        "phx_render": (C.c_int, [P, C.POINTER(abi.Camera), C.POINTER(abi.RenderParams)]),
        "phx_render_reference_structure": (C.c_int, [P, C.POINTER(abi.Camera), C.POINTER(abi.RenderParams)]),
        "phx_read_result": (C.c_int, [P, PF, PF, PF, PF, PF]),
        "phx_render_batch": (C.c_int, [P, C.POINTER(abi.View), C.c_uint32]),
        "phx_read_batch_result": (C.c_int, [P, C.c_uint32, PF, PF, PF, PF, PF]),
        "phx_autotune": (C.c_int, [P, C.POINTER(abi.View), C.c_uint32, C.c_uint32, C.POINTER(abi.Tuning),
                                   C.POINTER(C.c_double)]),
        "phx_result_device_color": (C.c_int, [P, C.POINTER(P)]),
        "phx_read_sample_counts": (C.c_int, [P, P]),
        "phx_camera_perspective": (C.c_int, [C.POINTER(abi.Camera), PF, PF, PF, C.c_float, C.c_float, C.c_float,
                                             C.c_float]),
        "phx_camera_orthographic": (C.c_int, [C.POINTER(abi.Camera), PF, PF, PF, C.c_float, C.c_float]),
        "phx_volume_majorants": (C.c_int, [C.POINTER(abi.Volume), PF, PF, C.c_uint32, PF, C.c_uint32, PU]),
        "phx_convert_material": (C.c_int, [C.POINTER(abi.BridgeMaterial), C.POINTER(abi.Material)]),
        "phx_bridge_render": (C.c_int, [P, C.POINTER(abi.CameraData), C.POINTER(abi.RenderSettings), C.c_int]),
        "phx_bridge_accumulated_samples": (C.c_uint32, [P]),
        "phx_set_scene_data": (C.c_int, [P, C.POINTER(abi.SceneData)]),
        "phx_bridge_read_color": (C.c_int, [P, PF]),
        "phx_bridge_read_depth": (C.c_int, [P, PF]),
        "phx_bridge_read_normal": (C.c_int, [P, PF]),
        "phx_bridge_read_albedo": (C.c_int, [P, PF]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return sigs


def lib():
    """The loaded library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with photon_amd.build_lib.build()")
        abi.check_layout()
        # PyTorch ships its own copy of the HIP runtime under the same soname
        # (libamdhip64.so.7) but is linked against it by another name, so
        # loading this library first would put the system runtime in the
        # process beside torch's, and torch then finds no GPU.  With torch
        # loaded first the library binds to torch's runtime: one runtime, and
        # device pointers pass between the two.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        handle = C.CDLL(LIB_PATH)
        _sig(handle)
        _lib = handle
    return _lib


def check(ctx, code):
    if code != 0:
        msg = lib().phx_last_error(ctx)
        raise PhxError(code, msg.decode() if msg else "")
    return code


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def volume_majorants(vol, values, tf):
    """phx_volume_majorants: fills vol.macro_dims / vol.majorant, returns the cells."""
    import numpy as np
    values = np.ascontiguousarray(values, np.float32)
    tf = np.ascontiguousarray(tf, np.float32)
    n = C.c_uint32()
    L = lib()
    code = L.phx_volume_majorants(C.byref(vol), fptr(values), fptr(tf), tf.size // 4, None, 0, C.byref(n))
    if code != 0:
        raise PhxError(code, "phx_volume_majorants")
    out = np.zeros(n.value, np.float32)
    code = L.phx_volume_majorants(C.byref(vol), fptr(values), fptr(tf), tf.size // 4, fptr(out), n.value, C.byref(n))
    if code != 0:
        raise PhxError(code, "phx_volume_majorants")
    return out
