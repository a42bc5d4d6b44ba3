"""pbrt-v4 scene files: photon::pbrt::parse_pbrt_file + convert_pbrt_scene
(pbrt_parser.cpp:233, pbrt_to_photon.cpp:58-503) through phx_pbrt_load.

    sc = load_pbrt("scene.pbrt")
    pt = PathTracer(); pt.set_scene(sc)
    pt.camera, pt.params = sc.camera, sc.render_params()
"""
import ctypes as C

import numpy as np

from . import abi
from ._lib import PhxError, lib


class PbrtScene:
    """A converted pbrt scene; desc() is the phx_scene_desc that
    PathTracer.set_scene takes (its arrays live as long as this object)."""

    def __init__(self, path, use_instancing=False):
        h = C.c_void_p()
        code = lib().phx_pbrt_load(str(path).encode(), int(use_instancing), C.byref(h))
        if code != 0:
            raise PhxError(code, f"phx_pbrt_load({path})")
        self._h = h
        self._desc = abi.SceneDesc()
        self.camera = abi.Camera()
        self.info = abi.PbrtInfo()
        code = lib().phx_pbrt_scene_desc(self._h, C.byref(self._desc), C.byref(self.camera), C.byref(self.info))
        if code != 0:
            raise PhxError(code, "phx_pbrt_scene_desc")
        self.name = str(path)

    def desc(self):
        return self._desc

    def render_params(self):
        """RenderParams of the file's Film, Sampler and Integrator."""
        i = self.info
        return abi.default_render_params(i.width, i.height, i.spp, i.max_depth)

    # views of the converted arrays (copies), for checks
    def arrays(self):
        d = self._desc

        def take(ptr, n, dt):
            if not ptr or n == 0:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)
        nv, nt = d.vertex_count, d.triangle_count
        out = {
            "positions": take(d.positions, nv * 3, np.float32).reshape(-1, 3),
            "indices": take(d.indices, nt * 3, np.uint32).reshape(-1, 3),
            "normals": take(d.normals, nv * 3, np.float32).reshape(-1, 3),
            "texcoords": take(d.texcoords, nv * 2, np.float32).reshape(-1, 2),
            "material_ids": take(d.material_ids, nt, np.uint32),
            "emissive_ids": take(d.emissive_prim_ids, d.emissive_count, np.uint32),
            "emissive_areas": take(d.emissive_prim_areas, d.emissive_count, np.float32),
            "texture_pixels": take(d.texture_pixels, d.texture_pixel_count * 3, np.float32).reshape(-1, 3),
            "env_pixels": take(d.env_pixels, d.env_width * d.env_height * 3, np.float32),
            "env_size": (d.env_width, d.env_height),
            "env_world_to_tex": np.array(list(d.env_world_to_tex), np.float32),
        }
        out["materials"] = [abi.Material.from_buffer_copy(d.materials[i]) for i in range(d.material_count)]
        out["lights"] = [abi.Light.from_buffer_copy(d.lights[i]) for i in range(d.light_count)]
        out["texture_infos"] = [(d.texture_infos[i].offset, d.texture_infos[i].width, d.texture_infos[i].height)
                                for i in range(d.texture_count)]
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().phx_pbrt_free(h)
            self._h = None


def load_pbrt(path, use_instancing=False):
    return PbrtScene(path, use_instancing)
