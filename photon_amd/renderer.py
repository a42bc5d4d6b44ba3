"""Mirror of photon::bridge::Renderer (reference src/photon/anari/RenderBridge.h:175-214).

Same method set and meaning: ``initialize`` / ``set_scene`` / ``render(camera,
settings, reset)`` / ``accumulated_samples`` / ``read_color`` (linear RGBA,
alpha clamped to [0, 1]) / ``read_depth`` (distance along the ray, inf on miss)
/ ``read_normal`` / ``read_albedo``; rows start at the bottom of the frame.
Where the reference returns ``bool`` plus an error string, these methods
return ``(ok, error)`` tuples, and the ``*_or_raise`` helpers raise instead.
The accumulation itself (RenderBridge.cpp:828-928) runs in the library.
"""
import ctypes as C

import numpy as np

from . import abi
from ._lib import PhxError, fptr, lib
from .backend import Context


def convert_material(bridge_material):
    """convert_material (RenderBridge.cpp:324-370) through phx_convert_material:
    the pt-level Material the bridge renders an abi.BridgeMaterial with.  Only
    constant parameter sources are supported (PhxError otherwise)."""
    out = abi.Material()
    code = lib().phx_convert_material(C.byref(bridge_material), C.byref(out))
    if code != 0:
        raise PhxError(code, "phx_convert_material: attribute and sampler sources are not supported")
    return out


class Renderer:
    def __init__(self, device=0):
        self._device = device
        self._ctx = None
        self._size = (0, 0)

    def initialize(self):
        if self._ctx is None:
            try:
                self._ctx = Context(self._device)
            except Exception as e:  # the reference reports, it does not throw
                return False, str(e)
        return True, ""

    def backend_name(self):
        return lib().phx_backend_name().decode()

    def execution_space(self):
        return "HIP"

    def _call(self, code):
        if code != 0:
            msg = lib().phx_last_error(self._ctx.h)
            return False, (msg.decode() if msg else f"error {code}")
        return True, ""

    def set_scene(self, scene):
        ok, err = self.initialize()
        if not ok:
            return ok, err
        self._scene = scene
        self._desc = scene.desc()
        return self._call(lib().phx_set_scene(self._ctx.h, C.byref(self._desc)))

    def render(self, camera, settings, reset):
        ok, err = self.initialize()
        if not ok:
            return ok, err
        self._size = (settings.height, settings.width)
        return self._call(lib().phx_bridge_render(self._ctx.h, C.byref(camera), C.byref(settings), int(bool(reset))))

    def accumulated_samples(self):
        return 0 if self._ctx is None else lib().phx_bridge_accumulated_samples(self._ctx.h)

    def _read(self, fn, channels):
        if self._ctx is None:
            return False, "not initialized", None
        h, w = self._size
        out = np.empty((h, w, channels) if channels > 1 else (h, w), np.float32)
        ok, err = self._call(fn(self._ctx.h, fptr(out)))
        return ok, err, out if ok else None

    def read_color(self):
        return self._read(lib().phx_bridge_read_color, 4)

    def read_depth(self):
        return self._read(lib().phx_bridge_read_depth, 1)

    def read_normal(self):
        return self._read(lib().phx_bridge_read_normal, 3)

    def read_albedo(self):
        return self._read(lib().phx_bridge_read_albedo, 3)

    def read_color_or_raise(self):
        ok, err, out = self.read_color()
        if not ok:
            raise PhxError(-5, err)
        return out

    @staticmethod
    def shutdown():
        """Renderer::shutdown: contexts are released with their owners."""


def camera_data(position, direction, up=(0.0, 1.0, 0.0), fovy=1.0471976, aspect=0.0):
    c = abi.CameraData()
    c.position[:] = position
    c.direction[:] = direction
    c.up[:] = up
    c.orthographic = 0
    c.fovy = fovy
    c.height = 1.0
    c.aspect = aspect
    c.aperture_radius = 0.0
    c.focus_distance = 1.0
    c.image_region[:] = (0.0, 0.0, 1.0, 1.0)
    c.clip_near = 0.0
    c.clip_far = float("inf")
    return c


def render_settings(width, height, pixel_samples=1, max_depth=6):
    s = abi.RenderSettings()
    s.width, s.height, s.pixel_samples, s.max_depth = width, height, pixel_samples, max_depth
    s.background[:] = (0.0, 0.0, 0.0, 1.0)
    s.ambient_color[:] = (1.0, 1.0, 1.0)
    s.ambient_radiance = 0.0
    s.adaptive_sampling = 0
    s.throughput_clamp = 0.0
    s.max_diffuse_bounces = s.max_glossy_bounces = s.max_transmission_bounces = s.max_volume_bounces = -1
    s.clamp_direct = s.clamp_indirect = 0.0
    return s
