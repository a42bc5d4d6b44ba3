"""Host-side mirror of the reference's ray backend and path tracer interfaces.

* ``HipBackend`` mirrors ``RayBackend`` (reference include/photon/pt/backend/
  ray_backend.h:44-54): ``build_accel(scene)``, ``trace_closest(rays, hits)``,
  ``trace_occluded(rays, occluded)``, ``name()``.  Rays and results are device
  tensors (RayBatch-shaped: origins/directions (n, 3), tmin, tmax).
* ``create_best_backend()`` mirrors backend_factory.cpp:28-80 for the backends
  this package has: ``PHOTON_BACKEND=hip`` (or unset / ``auto``) gives the HIP
  backend; a name of a backend that is not built here raises.
* ``PathTracer`` mirrors ``PathTracer`` (pathtracer.h:65-73): ``set_scene``,
  ``render`` with ``params`` (RenderParams) and ``camera``.
"""
import ctypes as C
import os
import sys

import numpy as np

from . import abi
from ._lib import PhxError, check, fptr, lib

KNOWN_BACKENDS = ("hip", "kokkos", "embree", "optix", "hiprt")


def backend_verbose():
    """backend_factory.cpp:22-26: PHOTON_VERBOSE set."""
    return bool(os.environ.get("PHOTON_VERBOSE"))


class Context:
    """Owns one phx_context (one GPU, one stream)."""

    def __init__(self, device=0):
        self._lib = lib()
        h = C.c_void_p()
        code = self._lib.phx_create(C.byref(h), device)
        if code != 0:
            raise PhxError(code, "phx_create failed (no HIP device?)")
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            self._lib.phx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def check(self, code):
        return check(self.h, code)

    def stats(self):
        s = abi.Stats()
        self.check(self._lib.phx_get_stats(self.h, C.byref(s)))
        return s

    def set_tuning(self, **kw):
        t = abi.Tuning()
        self.check(self._lib.phx_get_tuning(self.h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, v)
        self.check(self._lib.phx_set_tuning(self.h, C.byref(t)))

    def synchronize(self):
        self.check(self._lib.phx_synchronize(self.h))


def _dptr(t):
    return C.c_void_p(t.data_ptr())


class HipBackend:
    """RayBackend over the gfx950 traversal kernels."""

    def __init__(self, context=None, device=0):
        self.ctx = context or Context(device)

    def name(self):
        return lib().phx_backend_name().decode()

    def build_accel(self, scene):
        self._scene = scene
        self._desc = scene.desc()
        self.ctx.check(lib().phx_build_accel(self.ctx.h, C.byref(self._desc)))

    def _rays(self, origins, directions, tmin, tmax):
        import torch
        n = origins.shape[0]
        for t, shape in ((origins, (n, 3)), (directions, (n, 3)), (tmin, (n,)), (tmax, (n,))):
            if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous() or tuple(t.shape) != shape:
                raise ValueError("rays must be contiguous float32 device tensors of shape (n,3),(n,3),(n,),(n,)")
        return n

    def trace_closest(self, origins, directions, tmin, tmax, hits, stream=None):
        """Full HitResult records into `hits` (uint8 tensor of n * 88 bytes)."""
        n = self._rays(origins, directions, tmin, tmax)
        if hits.numel() * hits.element_size() < n * C.sizeof(abi.Hit):
            raise ValueError("hits buffer too small")
        self.ctx.check(lib().phx_trace_closest(self.ctx.h, _dptr(origins), _dptr(directions), _dptr(tmin),
                                               _dptr(tmax), n, _dptr(hits), stream))

    def trace_closest_compact(self, origins, directions, tmin, tmax, hits, stream=None):
        """(t, prim_id, b1, b2) per ray into `hits` (float32 tensor (n, 4))."""
        n = self._rays(origins, directions, tmin, tmax)
        if hits.numel() < n * 4:
            raise ValueError("hits buffer too small")
        self.ctx.check(lib().phx_trace_closest_compact(self.ctx.h, _dptr(origins), _dptr(directions), _dptr(tmin),
                                                       _dptr(tmax), n, _dptr(hits), stream))

    def trace_occluded(self, origins, directions, tmin, tmax, occluded, stream=None):
        n = self._rays(origins, directions, tmin, tmax)
        if occluded.numel() < n:
            raise ValueError("occluded buffer too small")
        self.ctx.check(lib().phx_trace_occluded(self.ctx.h, _dptr(origins), _dptr(directions), _dptr(tmin),
                                                _dptr(tmax), n, _dptr(occluded), stream))


def create_best_backend(device=0):
    """backend_factory.cpp:28-80 for this package: the HIP backend."""
    name = os.environ.get("PHOTON_BACKEND", "").strip().lower()
    if name in ("", "auto", "hip"):
        if backend_verbose():
            sys.stderr.write("[photon] backend: hip (gfx950 wavefront)\n")
        return HipBackend(device=device)
    if name in KNOWN_BACKENDS:
        raise RuntimeError(f"PHOTON_BACKEND={name}: that backend is not part of this package (only 'hip')")
    raise RuntimeError(f"PHOTON_BACKEND={name}: unknown backend")


def make_camera(look_from, look_at, vup=(0.0, 1.0, 0.0), vfov=40.0, aspect=1.0, aperture=0.0, focus_dist=1.0):
    """Camera::make_perspective (camera.h:32-55), computed by the library."""
    cam = abi.Camera()
    f = np.asarray(look_from, np.float32)
    a = np.asarray(look_at, np.float32)
    u = np.asarray(vup, np.float32)
    code = lib().phx_camera_perspective(C.byref(cam), fptr(f), fptr(a), fptr(u), vfov, aspect, aperture, focus_dist)
    if code != 0:
        raise PhxError(code, "phx_camera_perspective")
    return cam


def make_orthographic_camera(look_from, look_at, vup=(0.0, 1.0, 0.0), height=1.0, aspect=1.0):
    """Camera::make_orthographic (camera.h:57-77), computed by the library."""
    cam = abi.Camera()
    f = np.asarray(look_from, np.float32)
    a = np.asarray(look_at, np.float32)
    u = np.asarray(vup, np.float32)
    code = lib().phx_camera_orthographic(C.byref(cam), fptr(f), fptr(a), fptr(u), height, aspect)
    if code != 0:
        raise PhxError(code, "phx_camera_orthographic")
    return cam


class BatchRenderer:
    """phx_render_batch: many views (of one or several scenes) in one wavefront
    with shared queues.  A view is (scene, camera, params) where scene is a
    PathTracer (or Context) holding the view's scene."""

    def __init__(self, context=None, device=0):
        self.ctx = context if context is not None else Context(device)
        self._views = None

    @staticmethod
    def _table(views):
        arr = (abi.View * len(views))()
        for i, (scene, cam, params) in enumerate(views):
            ctx = scene.ctx if hasattr(scene, "ctx") else scene
            arr[i].scene = ctx.h
            arr[i].camera = cam
            arr[i].params = params
        return arr

    def render(self, views):
        arr = self._table(views)
        self.ctx.check(lib().phx_render_batch(self.ctx.h, arr, len(views)))
        self._views = [(p.width, p.height) for _, _, p in views]

    def read(self, v):
        w, h = self._views[v]
        color = np.empty((h, w, 3), np.float32)
        alpha = np.empty((h, w), np.float32)
        depth = np.empty((h, w), np.float32)
        normal = np.empty((h, w, 3), np.float32)
        albedo = np.empty((h, w, 3), np.float32)
        self.ctx.check(lib().phx_read_batch_result(self.ctx.h, v, fptr(color), fptr(alpha), fptr(depth),
                                                   fptr(normal), fptr(albedo)))
        return RenderResult(color, alpha, depth, normal, albedo)

    def autotune(self, views, repeats=1):
        """phx_autotune over `views`: returns (tuning, milliseconds) and keeps it."""
        arr = self._table(views)
        t = abi.Tuning()
        ms = C.c_double()
        self.ctx.check(lib().phx_autotune(self.ctx.h, arr, len(views), repeats, C.byref(t), C.byref(ms)))
        return t, ms.value


class RenderResult:
    def __init__(self, color, alpha, depth, normal, albedo):
        self.color, self.alpha, self.depth, self.normal, self.albedo = color, alpha, depth, normal, albedo


class PathTracer:
    """PathTracer (pathtracer.h:65-73) on the gfx950 wavefront kernels."""

    def __init__(self, backend=None):
        self.backend = backend or create_best_backend()
        self.ctx = self.backend.ctx
        self.params = abi.default_render_params(512, 512, 8, 5)
        self.camera = None

    def set_scene(self, scene):
        self._scene = scene
        self._desc = scene.desc()
        self.ctx.check(lib().phx_set_scene(self.ctx.h, C.byref(self._desc)))

    def render(self, read=True, structure="wavefront"):
        """Renders one frame.  structure="reference" runs the baseline with the
        reference's algorithmic structure (full-width bounce loop over a binary
        median-split BVH) instead of the streaming wavefront."""
        if self.camera is None:
            raise RuntimeError("PathTracer.render: no camera")
        if structure == "wavefront":
            fn = lib().phx_render
        elif structure == "reference":
            fn = lib().phx_render_reference_structure
        else:
            raise ValueError(f"PathTracer.render: unknown structure {structure!r}")
        self.ctx.check(fn(self.ctx.h, C.byref(self.camera), C.byref(self.params)))
        if not read:
            return None
        return self.read_result()

    def read_result(self):
        h, w = self.params.height, self.params.width
        color = np.empty((h, w, 3), np.float32)
        alpha = np.empty((h, w), np.float32)
        depth = np.empty((h, w), np.float32)
        normal = np.empty((h, w, 3), np.float32)
        albedo = np.empty((h, w, 3), np.float32)
        self.ctx.check(lib().phx_read_result(self.ctx.h, fptr(color), fptr(alpha), fptr(depth), fptr(normal),
                                             fptr(albedo)))
        return RenderResult(color, alpha, depth, normal, albedo)

    def sample_counts(self):
        """Samples each pixel took in the last, adaptively sampled frame
        ((h, w) uint32, rows as rendered)."""
        h, w = self.params.height, self.params.width
        counts = np.empty((h, w), np.uint32)
        self.ctx.check(lib().phx_read_sample_counts(self.ctx.h, C.c_void_p(counts.ctypes.data)))
        return counts

    def device_color(self):
        p = C.c_void_p()
        self.ctx.check(lib().phx_result_device_color(self.ctx.h, C.byref(p)))
        return p.value

    def display(self, display=None):
        """The last frame through phx_display_transform on the GPU (default:
        photon_render's postprocess, ACES + sRGB curve with clamps): returns
        (float32 (h, w, 3), uint8 (h, w, 3)) torch tensors on the device."""
        import torch
        h, w = self.params.height, self.params.width
        out = torch.empty((h, w, 3), dtype=torch.float32, device=f"cuda:{self.ctx.device}")
        out8 = torch.empty((h, w, 3), dtype=torch.uint8, device=out.device)
        d = display if display is not None else abi.display()
        self.ctx.check(lib().phx_display_transform(self.ctx.h, self.device_color(), 3, h * w, C.byref(d),
                                                   out.data_ptr(), out8.data_ptr(), None))
        self.ctx.synchronize()
        return out, out8


def display_transform(ctx, rgb, display=None):
    """phx_display_transform of a (..., 3) or (..., 4) float32 CUDA tensor:
    returns (float32 (..., 3), uint8 (..., 3)) tensors on the same device."""
    import torch
    if not (rgb.is_cuda and rgb.dtype == torch.float32 and rgb.shape[-1] in (3, 4)):
        raise ValueError("display_transform: a float32 CUDA tensor of 3 or 4 channels")
    rgb = rgb.contiguous()
    ch = rgb.shape[-1]
    n = rgb.numel() // ch
    out = torch.empty(rgb.shape[:-1] + (3,), dtype=torch.float32, device=rgb.device)
    out8 = torch.empty(rgb.shape[:-1] + (3,), dtype=torch.uint8, device=rgb.device)
    d = display if display is not None else abi.display()
    ctx.check(lib().phx_display_transform(ctx.h, rgb.data_ptr(), ch, n, C.byref(d), out.data_ptr(),
                                          out8.data_ptr(), None))
    ctx.synchronize()
    return out, out8
