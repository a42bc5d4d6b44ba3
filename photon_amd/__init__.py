"""photon_amd - MI355X (gfx950) wavefront path-tracing backend.

A drop-in for the Kokkos BVH path of the Photon renderer: hand-written HIP
kernels behind the C ABI of include/phx.h, driven from this host package.
"""
from . import abi, scenes  # noqa: F401
from .backend import (Context, HipBackend, PathTracer, RenderResult, create_best_backend,  # noqa: F401
                      make_camera)
from .renderer import Renderer, camera_data, render_settings  # noqa: F401

__all__ = ["abi", "scenes", "Context", "HipBackend", "PathTracer", "RenderResult", "create_best_backend",
           "make_camera", "Renderer", "camera_data", "render_settings"]
