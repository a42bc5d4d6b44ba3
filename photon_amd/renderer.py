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
    the pt-level Material the bridge renders an abi.BridgeMaterial with."""
    out = abi.Material()
    code = lib().phx_convert_material(C.byref(bridge_material), C.byref(out))
    if code != 0:
        raise PhxError(code, "phx_convert_material")
    return out


def _f32(a, n=None):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32).reshape(-1 if n is None else n)


class SceneData:
    """bridge::SceneData (RenderBridge.h:112-125): the scene every front end
    hands the bridge, with numpy arrays for its std::vectors.  positions /
    normals (V, 3), attributes[slot] (V, 4) RGBA, indices (T, 3), material_ids
    (T,), materials (abi.BridgeMaterial), samplers (sampler_data()), lights
    (abi.LightData), environment (environment_data() or None), volumes
    (volume_data())."""

    def __init__(self, positions=None, indices=None, materials=(), material_ids=None, normals=None,
                 attributes=None, samplers=(), lights=(), environment=None, volumes=()):
        self.positions = np.zeros((0, 3), np.float32) if positions is None else _f32(positions, (-1, 3))
        self.indices = (np.zeros((0, 3), np.uint32) if indices is None
                        else np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3))
        self.materials = list(materials)
        self.material_ids = None if material_ids is None else np.ascontiguousarray(material_ids, dtype=np.uint32)
        self.normals = None if normals is None else _f32(normals, (-1, 3))
        self.attributes = [None] * 5
        for slot, a in (attributes or {}).items():
            self.attributes[slot] = _f32(a, (-1, 4))
        self.samplers = list(samplers)
        self.lights = list(lights)
        self.environment = environment
        self.volumes = list(volumes)

    def c_struct(self):
        """phx_scene_data over these arrays (kept alive by this object)."""
        d = abi.SceneData()
        keep = []

        def ptr(a):
            return None if a is None else a.ctypes.data

        d.positions, d.position_count = ptr(self.positions), self.positions.size
        if self.normals is not None:
            d.normals, d.normal_count = ptr(self.normals), self.normals.size
        for k, a in enumerate(self.attributes):
            if a is not None:
                d.attributes[k], d.attribute_count[k] = ptr(a), a.size
        d.indices, d.index_count = ptr(self.indices), self.indices.size
        if self.material_ids is not None:
            d.material_ids, d.material_id_count = ptr(self.material_ids), self.material_ids.size
        if self.materials:
            mats = (abi.BridgeMaterial * len(self.materials))(*self.materials)
            keep.append(mats)
            d.materials, d.material_count = C.cast(mats, C.POINTER(abi.BridgeMaterial)), len(self.materials)
        if self.samplers:
            sms = (abi.SamplerData * len(self.samplers))(*[sd.struct for sd in self.samplers])
            keep.append(sms)
            d.samplers, d.sampler_count = C.cast(sms, C.POINTER(abi.SamplerData)), len(self.samplers)
        if self.lights:
            ls = (abi.LightData * len(self.lights))(*self.lights)
            keep.append(ls)
            d.lights, d.light_count = C.cast(ls, C.POINTER(abi.LightData)), len(self.lights)
        if self.environment is not None:
            d.environment = self.environment.struct
        if self.volumes:
            vs = (abi.VolumeData * len(self.volumes))(*[v.struct for v in self.volumes])
            keep.append(vs)
            d.volumes, d.volume_count = C.cast(vs, C.POINTER(abi.VolumeData)), len(self.volumes)
        self._keep = keep
        return d


class _Held:
    """A ctypes struct together with the arrays its pointers refer to."""

    def __init__(self, struct, *arrays):
        self.struct = struct
        self.arrays = arrays


def sampler_data(type_=1, in_attribute=0, linear=True, wrap=(0, 0, 0), size=(1, 1, 1), texels=None,
                 in_transform=None, in_offset=(0, 0, 0, 0), out_transform=None, out_offset=(0, 0, 0, 0)):
    """bridge::SamplerData (RenderBridge.h:58-71); texels (N, 4) RGBA."""
    ident = (1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0)
    s = abi.SamplerData()
    s.type, s.in_attribute, s.linear = type_, in_attribute, 1 if linear else 0
    s.wrap[:] = wrap
    s.size[:] = size
    s.in_transform[:] = ident if in_transform is None else [float(x) for x in np.asarray(in_transform).reshape(-1)]
    s.in_offset[:] = in_offset
    s.out_transform[:] = ident if out_transform is None else [float(x) for x in np.asarray(out_transform).reshape(-1)]
    s.out_offset[:] = out_offset
    tx = None if texels is None else _f32(texels, (-1,))
    if tx is not None:
        s.texels, s.texel_float_count = tx.ctypes.data, tx.size
    return _Held(s, tx)


def environment_data(pixels, visible=True, world_to_map=None):
    """bridge::EnvironmentData (RenderBridge.h:87-97): pixels (H, W, 3) RGB."""
    px = _f32(pixels)
    h, w = np.asarray(pixels).shape[:2]
    e = abi.EnvironmentData()
    e.present, e.visible, e.width, e.height = 1, 1 if visible else 0, w, h
    e.pixels, e.pixel_float_count = px.ctypes.data, px.size
    e.world_to_map[:] = [float(x) for x in np.asarray(np.eye(3) if world_to_map is None else world_to_map).reshape(-1)]
    return _Held(e, px)


def volume_data(values, dims, transfer, value_range=(0.0, 1.0), density_scale=1.0, origin=(0, 0, 0),
                spacing=(1, 1, 1), world_to_local=None, linear=True):
    """bridge::VolumeData (RenderBridge.h:99-110): values x-fastest, transfer (N, 4) RGBA."""
    vals, tf = _f32(values, (-1,)), _f32(transfer, (-1,))
    v = abi.VolumeData()
    v.world_to_local[:] = ((1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0) if world_to_local is None
                           else [float(x) for x in np.asarray(world_to_local).reshape(-1)])
    v.origin[:], v.spacing[:], v.dims[:] = origin, spacing, dims
    v.values, v.value_count = vals.ctypes.data, vals.size
    v.transfer, v.transfer_float_count = tf.ctypes.data, tf.size
    v.value_range[:] = value_range
    v.density_scale = density_scale
    v.linear = 1 if linear else 0
    return _Held(v, vals, tf)


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
        """Renderer::set_scene: a SceneData (converted as the bridge does,
        phx_set_scene_data), or a scenes.Scene already at pt level."""
        ok, err = self.initialize()
        if not ok:
            return ok, err
        self._scene = scene
        if isinstance(scene, SceneData):
            self._desc = scene.c_struct()
            return self._call(lib().phx_set_scene_data(self._ctx.h, C.byref(self._desc)))
        self._desc = scene.desc()
        return self._call(lib().phx_set_scene(self._ctx.h, C.byref(self._desc)))

    def render(self, camera, settings, reset, background=None):
        """Renderer::render; `background` (H, W, 4) RGBA rows from the bottom is
        RenderSettings::background_image."""
        ok, err = self.initialize()
        if not ok:
            return ok, err
        self._size = (settings.height, settings.width)
        if background is not None:
            self._bg = _f32(background, (-1,))
            settings.background_image = self._bg.ctypes.data
            settings.background_height, settings.background_width = np.asarray(background).shape[:2]
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
