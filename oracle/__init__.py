"""ORACLE - test infrastructure only.

ctypes wrapper of oracle/build/liboracle.so, the CPU restatement of the
reference's path-tracing semantics (see orc_core.h for what is restated and
how it is pinned).  Only tests, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, and only as the checker.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_scene_create.restype = P
        L.orc_scene_create.argtypes = [P, C.c_int]
        L.orc_scene_destroy.argtypes = [P]
        L.orc_scene_ray_epsilon.restype = C.c_float
        L.orc_scene_ray_epsilon.argtypes = [P]
        L.orc_scene_bvh_nodes.restype = C.c_uint32
        L.orc_scene_bvh_nodes.argtypes = [P]
        L.orc_scene_emissive.restype = C.c_uint32
        L.orc_scene_emissive.argtypes = [P, P, P, P]
        L.orc_trace_closest.argtypes = [P, P, P, P, P, C.c_uint32, P, P, P, P, C.c_int]
        L.orc_trace_closest_full.argtypes = [P, P, P, P, P, C.c_uint32, P]
        L.orc_trace_occluded.argtypes = [P, P, P, P, P, C.c_uint32, P, C.c_int]
        L.orc_bsdf_eval.argtypes = [P, P, P, P, P, C.c_uint32, P, P, C.c_int]
        L.orc_bsdf_sample.argtypes = [P, P, P, P, P, C.c_uint32, P, P, P, P]
        L.orc_rng.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, P, P]
        L.orc_hash.restype = C.c_uint32
        L.orc_hash.argtypes = [C.c_uint32]
        L.orc_power_heuristic.restype = C.c_float
        L.orc_power_heuristic.argtypes = [C.c_float, C.c_float]
        L.orc_onb.argtypes = [P, P, P]
        L.orc_env_sample.argtypes = [P, P, C.c_uint32, P, P]
        L.orc_env_eval.argtypes = [P, P, C.c_uint32, P, P]
        L.orc_env_cdf.argtypes = [P, P, P]
        L.orc_camera_perspective.argtypes = [P, P, P, P, C.c_float, C.c_float, C.c_float, C.c_float]
        L.orc_volume_majorants.restype = C.c_uint32
        L.orc_volume_majorants.argtypes = [P, P, P, C.c_uint32, P, C.c_uint32]
        L.orc_volume_transmittance.argtypes = [P, P, P, C.c_float, C.c_float, P, C.c_uint32, P]
        L.orc_volume_scatter.argtypes = [P, P, P, C.c_float, C.c_float, P, C.c_uint32, P]
        L.orc_render.restype = C.c_double
        L.orc_bsdf_eval_grad_fd.argtypes = [P, P, P, P, P, C.c_uint32, C.c_double, P]
        L.orc_render_detached.restype = C.c_double
        L.orc_render_detached.argtypes = [P, P, P, P, C.c_uint32, P, P, P, C.c_int]
        L.orc_render.argtypes = [P, P, P, P, P, P, P, P, P, C.c_int, C.c_int, C.c_uint32, C.c_uint32]
        L.orc_lobe_weights.argtypes = [P, C.c_double, C.c_int, P]
        L.orc_sample_light.argtypes = [P, P, C.c_uint32, P]
        L.orc_texture_sample.argtypes = [P, C.c_uint32, P, C.c_uint32, C.c_int, C.c_double, C.c_double, C.c_int, P]
        L.orc_sampler_sample.argtypes = [P, C.c_uint32, P, C.c_uint32, C.c_int, P, C.c_int, P]
        L.orc_surface_resolve.argtypes = [P, C.c_uint32, C.c_double, C.c_double, P, P, P, P, P, P]
        L.orc_display.argtypes = [P, C.c_uint32, C.c_uint64, P, P, P]
        _lib = L
    return _lib


def _p(a):
    """Pointer to an array's data; the caller holds the array for the call."""
    return None if a is None else C.c_void_p(a.ctypes.data)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class OracleScene:
    """The oracle's view of a photon_amd.scenes.Scene.  brute=True traces
    every ray against every triangle instead of the reference BVH."""

    def __init__(self, scene, brute=False):
        self.scene = scene
        self._desc = scene.desc()
        self.h = lib().orc_scene_create(C.addressof(self._desc), int(brute))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    @property
    def ray_epsilon(self):
        return lib().orc_scene_ray_epsilon(self.h)

    def emissive_tables(self):
        n = lib().orc_scene_emissive(self.h, None, None, None)
        cdf = np.zeros(max(n, 1), np.float32)
        pdf = np.zeros(self.scene.triangle_count, np.float32)
        ids = np.zeros(max(n, 1), np.uint32)
        lib().orc_scene_emissive(self.h, _p(cdf), _p(pdf) if n else None, _p(ids))
        return cdf[:n], pdf, ids[:n]

    def trace_closest(self, org, dir_, tmin, tmax, brute=False):
        org, dir_ = _c(org, np.float32), _c(dir_, np.float32)
        tmin, tmax = _c(tmin, np.float32), _c(tmax, np.float32)
        n = org.shape[0]
        t = np.zeros(n, np.float32)
        prim = np.zeros(n, np.uint32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        lib().orc_trace_closest(self.h, _p(org), _p(dir_), _p(tmin), _p(tmax), n, _p(t), _p(prim), _p(u), _p(v),
                                int(brute))
        return t, prim, u, v

    def trace_closest_full(self, org, dir_, tmin, tmax):
        from photon_amd import abi
        org, dir_ = _c(org, np.float32), _c(dir_, np.float32)
        tmin, tmax = _c(tmin, np.float32), _c(tmax, np.float32)
        n = org.shape[0]
        out = (abi.Hit * n)()
        lib().orc_trace_closest_full(self.h, _p(org), _p(dir_), _p(tmin), _p(tmax), n, C.addressof(out))
        return out

    def trace_occluded(self, org, dir_, tmin, tmax, brute=False):
        org, dir_ = _c(org, np.float32), _c(dir_, np.float32)
        tmin, tmax = _c(tmin, np.float32), _c(tmax, np.float32)
        n = org.shape[0]
        occ = np.zeros(n, np.uint32)
        lib().orc_trace_occluded(self.h, _p(org), _p(dir_), _p(tmin), _p(tmax), n, _p(occ), int(brute))
        return occ

    def env_sample(self, seeds):
        seeds = _c(seeds, np.uint32)
        n = seeds.size
        d = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32)
        lib().orc_env_sample(self.h, _p(seeds), n, _p(d), _p(pdf))
        return d, pdf

    def env_eval(self, dirs):
        dirs = _c(dirs, np.float32)
        n = dirs.shape[0]
        pdf = np.zeros(n, np.float32)
        rgb = np.zeros((n, 3), np.float32)
        lib().orc_env_eval(self.h, _p(dirs), n, _p(pdf), _p(rgb))
        return pdf, rgb

    def env_cdf(self):
        h, w = self.scene.env_pixels.shape[:2]
        m = np.zeros(h + 1, np.float32)
        c = np.zeros((h, w + 1), np.float32)
        lib().orc_env_cdf(self.h, _p(m), _p(c))
        return m, c

    def volume_transmittance(self, org, dir_, tmin, tmax, seeds):
        seeds = _c(seeds, np.uint32)
        out = np.zeros(seeds.size, np.float32)
        org, dir_ = _c(org, np.float32), _c(dir_, np.float32)  # held: the call reads them
        lib().orc_volume_transmittance(self.h, _p(org), _p(dir_), tmin, tmax,
                                       _p(seeds), seeds.size, _p(out))
        return out

    def volume_scatter(self, org, dir_, tmin, tmax, seeds):
        seeds = _c(seeds, np.uint32)
        out = np.zeros(seeds.size, np.float32)
        org, dir_ = _c(org, np.float32), _c(dir_, np.float32)
        lib().orc_volume_scatter(self.h, _p(org), _p(dir_), tmin, tmax, _p(seeds),
                                 seeds.size, _p(out))
        return out

    def surface_resolve(self, prim, b1, b2, p, sn, uv=(0.0, 0.0)):
        """resolve_material plus the normal maps at a point of primitive
        `prim`: (material, alpha, shading normal), in double."""
        from photon_amd import abi
        m = abi.Material()
        alpha = np.zeros(1, np.float64)
        n = np.zeros(3, np.float64)
        p, sn, uv = _c(p, np.float64), _c(sn, np.float64), _c(uv, np.float64)
        lib().orc_surface_resolve(self.h, prim, b1, b2, _p(p), _p(sn), _p(uv), C.addressof(m), _p(alpha), _p(n))
        return m, float(alpha[0]), n

    def render(self, camera, params, threads=0, use_double=False, pixel_range=(0, 0), aovs=False, nsamp=False):
        """PathTracer::render on the CPU: returns (color (h,w,3), extras dict, seconds).
        extras holds the AOVs with aovs=True and each pixel's sample count
        ("nsamp", (h,w) uint32) with nsamp=True."""
        h, w = params.height, params.width
        color = np.zeros((h, w, 3), np.float32)
        extras = {}
        if aovs:
            extras = {"alpha": np.zeros((h, w), np.float32), "depth": np.zeros((h, w), np.float32),
                      "normal": np.zeros((h, w, 3), np.float32), "albedo": np.zeros((h, w, 3), np.float32)}
        if nsamp:
            extras["nsamp"] = np.zeros((h, w), np.uint32)
        secs = lib().orc_render(self.h, C.addressof(camera), C.addressof(params), _p(color),
                                _p(extras.get("alpha")), _p(extras.get("depth")), _p(extras.get("normal")),
                                _p(extras.get("albedo")), _p(extras.get("nsamp")), int(threads), int(use_double), int(pixel_range[0]),
                                int(pixel_range[1]))
        return color, extras, secs


def _render_detached(self, camera, params, mat_params=None, texels=None, delta=None, image=False, threads=0):
    """The detached estimator of the path-replay adjoint in double (see
    detached_pixel in oracle.cpp): float sampling decisions, double weights with
    `mat_params` (materials x 5: base r, g, b, roughness, metallic) and atlas
    `texels` (pixels x 3) replacing the scene's.  Returns (sum delta * image or
    None, image (h, w, 3) float64 or None)."""
    h, w = params.height, params.width
    mp = None if mat_params is None else np.ascontiguousarray(mat_params, np.float64).reshape(-1)
    tx = None if texels is None else np.ascontiguousarray(texels, np.float64).reshape(-1)
    dl = None if delta is None else np.ascontiguousarray(delta, np.float32).reshape(-1)
    img = np.zeros((h, w, 3), np.float64) if image else None
    total = lib().orc_render_detached(self.h, C.addressof(camera), C.addressof(params), _p(mp),
                                      0 if mp is None else len(mp) // 5, _p(tx), _p(dl), _p(img), int(threads))
    return (total if delta is not None else None), img


OracleScene.render_detached = _render_detached


def bsdf_eval_grad_fd(mats, wo, wi, ng, sn, h=1e-4):
    """Central differences in double of the BSDF eval with respect to base r, g,
    b, roughness, metallic: (n, 5, 3)."""
    from photon_amd import abi
    n = len(mats)
    arr = (abi.Material * n)(*mats)
    wo, wi, ng, sn = (_c(x, np.float32) for x in (wo, wi, ng, sn))
    df = np.zeros((n, 5, 3), np.float64)
    lib().orc_bsdf_eval_grad_fd(C.addressof(arr), _p(wo), _p(wi), _p(ng), _p(sn), n, h, _p(df))
    return df


def bsdf_eval(mats, wo, wi, ng, sn, use_double=False):
    from photon_amd import abi
    n = len(mats)
    arr = (abi.Material * n)(*mats)
    wo, wi, ng, sn = (_c(x, np.float32) for x in (wo, wi, ng, sn))
    f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    lib().orc_bsdf_eval(C.addressof(arr), _p(wo), _p(wi), _p(ng), _p(sn), n, _p(f), _p(pdf), int(use_double))
    return f, pdf


def bsdf_sample(mats, wo, ng, sn, seeds):
    from photon_amd import abi
    n = len(mats)
    arr = (abi.Material * n)(*mats)
    wo, ng, sn = (_c(x, np.float32) for x in (wo, ng, sn))
    seeds = _c(seeds, np.uint32)
    wi = np.zeros((n, 3), np.float32)
    f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    flags = np.zeros(n, np.uint32)
    lib().orc_bsdf_sample(C.addressof(arr), _p(wo), _p(ng), _p(sn), _p(seeds), n, _p(wi), _p(f), _p(pdf), _p(flags))
    return wi, f, pdf, flags


def rng_stream(pixel, sample, dim, n):
    u = np.zeros(n, np.uint32)
    f = np.zeros(n, np.float32)
    lib().orc_rng(pixel, sample, dim, n, _p(u), _p(f))
    return u, f


def hash_u32(x):
    return lib().orc_hash(x)


def power_heuristic(a, b):
    return lib().orc_power_heuristic(a, b)


def onb(n):
    n = _c(n, np.float32)
    t = np.zeros(3, np.float32)
    b = np.zeros(3, np.float32)
    lib().orc_onb(_p(n), _p(t), _p(b))
    return t, b


def camera_perspective(look_from, look_at, vup, vfov, aspect, aperture=0.0, focus=1.0):
    from photon_amd import abi
    cam = abi.Camera()
    f, a, u = (_c(x, np.float32) for x in (look_from, look_at, vup))
    lib().orc_camera_perspective(C.addressof(cam), _p(f), _p(a), _p(u), vfov, aspect, aperture, focus)
    return cam


def volume_majorants(vol, values, tf):
    values = _c(values, np.float32)
    tf = _c(tf, np.float32)
    n = lib().orc_volume_majorants(C.addressof(vol), _p(values), _p(tf), tf.size // 4, None, 0)
    out = np.zeros(n, np.float32)
    lib().orc_volume_majorants(C.addressof(vol), _p(values), _p(tf), tf.size // 4, _p(out), n)
    return out


def lobe_weights(mat, nov, use_double=True):
    """disney_lobe_weights at NoV: (wd, ws, wt, wc)."""
    out = np.zeros(4, np.float64)
    lib().orc_lobe_weights(C.addressof(mat), nov, int(use_double), _p(out))
    return out


def sample_light(light, p, seed):
    """sample_light from point p with Rng(seed): (wi, Li, pdf, dist)."""
    out = np.zeros(8, np.float32)
    p = _c(p, np.float32)
    lib().orc_sample_light(C.addressof(light), _p(p), seed, _p(out))
    return out[:3], out[3:6], float(out[6]), float(out[7])


def texture_sample(pixels, infos, tex_id, u, v, use_double=True):
    """TextureAtlas::sample: pixels (N, 3), infos [(offset, width, height)]."""
    from photon_amd import abi
    px = _c(pixels, np.float32).reshape(-1, 3)
    arr = (abi.TextureInfo * max(len(infos), 1))(*[abi.TextureInfo(o, w, h, 0) for (o, w, h) in infos])
    out = np.zeros(3, np.float64)
    lib().orc_texture_sample(_p(px), px.shape[0], C.addressof(arr), len(infos), tex_id, u, v, int(use_double), _p(out))
    return out


def sampler_sample(infos, texels, sampler_id, attr, use_double=True):
    """SamplerTable::sample of abi.SamplerInfo list over RGBA texels (N, 4)."""
    from photon_amd import abi
    arr = (abi.SamplerInfo * max(len(infos), 1))(*infos)
    tx = None if texels is None else _c(texels, np.float32).reshape(-1, 4)
    out = np.zeros(4, np.float64)
    attr = _c(attr, np.float64)
    lib().orc_sampler_sample(C.addressof(arr), len(infos), _p(tx), 0 if tx is None else tx.shape[0], sampler_id,
                             _p(attr), int(use_double), _p(out))
    return out


def display(rgb, t):
    """orc_display: the display transform t (an abi.Display) of an (..., 3) or
    (..., 4) radiance array; returns (float rgb, uint8 rgb)."""
    rgb = _c(rgb, np.float32)
    ch = rgb.shape[-1]
    n = rgb.size // ch
    out = np.zeros(rgb.shape[:-1] + (3,), np.float32)
    out8 = np.zeros(rgb.shape[:-1] + (3,), np.uint8)
    lib().orc_display(_p(rgb), ch, n, C.addressof(t), _p(out), _p(out8))
    return out, out8
