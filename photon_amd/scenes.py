"""Scene container and the seeded procedural workloads.

``Scene`` holds everything pt::Scene holds (reference include/photon/pt/scene.h:41-72)
as numpy arrays and turns it into the ``phx_scene_desc`` of include/phx.h.

Workloads (all deterministic for a given seed):
  * ``cornell_box``  - 36 triangles: 2x2x2 room (10), two boxes rotated +-18 deg
    about y (24), 0.5x0.5 ceiling light (2) of radiance (17, 12, 4); diffuse
    Disney materials; 40 deg pinhole camera.
  * ``sphere_grid``  - 6x6 icospheres at subdivision 5 (20,480 triangles each)
    displaced along the normal by seeded 4-octave fBm (amplitude 0.05 r), a
    2-triangle ground plane, 4 rectangular area lights as emissive quads
    (8 triangles) and a 1024x512 sky + Gaussian sun (sigma 2 deg, 5e3)
    environment: 737,290 triangles.
  * ``textured_cornell`` - the Cornell box with a textured ground quad and 32
    materials (gradient workload).
  * ``fog_box``      - a 256^3 fBm density grid in an open box (volume workload).

Material convention: the sphere parameters are Disney perceptual values; as the
render bridge does (RenderBridge.cpp:324-370), the GGX alpha the BSDF uses is the
square of the perceptual roughness, transmission materials refract
(MATERIAL_FLAG_REFRACTIVE).
"""
import ctypes as C
import math

import numpy as np

from . import abi


def _ptr(a, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


class Scene:
    def __init__(self, positions, indices, materials, material_ids=None, normals=None, texcoords=None,
                 vertex_colors=None, lights=None):
        self.positions = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        self.indices = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3)
        self.materials = list(materials)
        self.material_ids = None if material_ids is None else np.ascontiguousarray(material_ids, dtype=np.uint32)
        self.normals = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
        self.texcoords = None if texcoords is None else np.ascontiguousarray(texcoords, dtype=np.float32).reshape(-1, 2)
        self.vertex_colors = None if vertex_colors is None else np.ascontiguousarray(vertex_colors, dtype=np.float32)
        self.lights = list(lights or [])
        self.env_pixels = None  # (H, W, 3) float32, row 0 = +Y of the map
        self.env_world_to_tex = np.eye(3, dtype=np.float32)
        self.env_visible = True
        self.texture_pixels = None  # (N, 3)
        self.texture_infos = []     # (offset, width, height)
        self.volumes = []           # abi.Volume
        self.volume_values = None
        self.volume_tf = None
        self.volume_majorants = None
        self.attributes = {}        # slot (0-4) -> (vertex_count, 4) float32 RGBA records
        self.samplers = []          # abi.SamplerInfo
        self.sampler_texels = None  # (N, 4) float32
        self.ray_epsilon = 0.0
        # explicit emissive primitives and their areas (None: derived from the
        # materials by the library)
        self.emissive_ids = None
        self.emissive_areas = None
        self.camera = None          # dict(look_from, look_at, vup, vfov, aspect)
        self.name = "scene"

    @property
    def triangle_count(self):
        return int(self.indices.shape[0])

    def add_texture(self, rgb):
        rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        h, w, _ = rgb.shape
        off = 0 if self.texture_pixels is None else self.texture_pixels.shape[0]
        flat = rgb.reshape(-1, 3)
        self.texture_pixels = flat if self.texture_pixels is None else np.concatenate([self.texture_pixels, flat])
        self.texture_infos.append((off, w, h))
        return len(self.texture_infos) - 1

    def desc(self):
        """phx_scene_desc over this scene's arrays (the arrays stay owned here)."""
        d = abi.SceneDesc()
        self._keep = []
        d.positions = _ptr(self.positions, C.c_float)
        d.vertex_count = self.positions.shape[0]
        d.indices = _ptr(self.indices, C.c_uint32)
        d.triangle_count = self.indices.shape[0]
        d.normals = _ptr(self.normals, C.c_float)
        d.texcoords = _ptr(self.texcoords, C.c_float)
        d.vertex_colors = _ptr(self.vertex_colors, C.c_float)
        d.material_ids = _ptr(self.material_ids, C.c_uint32)
        mats = (abi.Material * max(len(self.materials), 1))(*self.materials)
        self._keep.append(mats)
        d.materials = C.cast(mats, C.POINTER(abi.Material))
        d.material_count = len(self.materials)
        if self.lights:
            lights = (abi.Light * len(self.lights))(*self.lights)
            self._keep.append(lights)
            d.lights = C.cast(lights, C.POINTER(abi.Light))
        d.light_count = len(self.lights)
        if self.emissive_ids is not None and len(self.emissive_ids):
            self.emissive_ids = np.ascontiguousarray(self.emissive_ids, dtype=np.uint32)
            self.emissive_areas = np.ascontiguousarray(self.emissive_areas, dtype=np.float32)
            d.emissive_prim_ids = _ptr(self.emissive_ids, C.c_uint32)
            d.emissive_prim_areas = _ptr(self.emissive_areas, C.c_float)
            d.emissive_count = self.emissive_ids.size
        if self.texture_infos:
            infos = (abi.TextureInfo * len(self.texture_infos))(
                *[abi.TextureInfo(o, w, h, 0) for (o, w, h) in self.texture_infos])
            self._keep.append(infos)
            d.texture_infos = C.cast(infos, C.POINTER(abi.TextureInfo))
            d.texture_count = len(self.texture_infos)
            d.texture_pixels = _ptr(self.texture_pixels, C.c_float)
            d.texture_pixel_count = self.texture_pixels.shape[0]
        if self.env_pixels is not None:
            self.env_pixels = np.ascontiguousarray(self.env_pixels, dtype=np.float32)
            d.env_height, d.env_width = self.env_pixels.shape[:2]
            d.env_pixels = _ptr(self.env_pixels, C.c_float)
        d.env_world_to_tex[:] = [float(x) for x in np.asarray(self.env_world_to_tex, np.float32).reshape(-1)]
        d.env_visible = 1 if self.env_visible else 0
        if self.volumes:
            vols = (abi.Volume * len(self.volumes))(*self.volumes)
            self._keep.append(vols)
            d.volumes = C.cast(vols, C.POINTER(abi.Volume))
            d.volume_count = len(self.volumes)
            d.volume_values = _ptr(self.volume_values, C.c_float)
            d.volume_value_count = self.volume_values.size
            d.volume_tf = _ptr(self.volume_tf, C.c_float)
            d.volume_tf_count = self.volume_tf.size // 4
            d.volume_majorants = _ptr(self.volume_majorants, C.c_float)
            d.volume_majorant_count = self.volume_majorants.size
        for k in range(5):
            d.attribute_offset[k] = -1
        if self.attributes:
            recs, off = [], 0
            for slot in sorted(self.attributes):
                a = np.ascontiguousarray(self.attributes[slot], dtype=np.float32).reshape(-1, 4)
                d.attribute_offset[slot] = off
                recs.append(a)
                off += a.shape[0]
            allrec = np.ascontiguousarray(np.concatenate(recs), dtype=np.float32)
            self._keep.append(allrec)
            d.attributes = _ptr(allrec, C.c_float)
            d.attribute_count = allrec.shape[0]
        if self.samplers:
            sms = (abi.SamplerInfo * len(self.samplers))(*self.samplers)
            self._keep.append(sms)
            d.samplers = C.cast(sms, C.POINTER(abi.SamplerInfo))
            d.sampler_count = len(self.samplers)
            if self.sampler_texels is not None:
                self.sampler_texels = np.ascontiguousarray(self.sampler_texels, dtype=np.float32).reshape(-1, 4)
                d.sampler_texels = _ptr(self.sampler_texels, C.c_float)
                d.sampler_texel_count = self.sampler_texels.shape[0]
        d.ray_epsilon = self.ray_epsilon
        return d


# --------------------------------------------------------------------------- materials
def diffuse(rgb):
    """A bridge MaterialData{base_color=rgb, roughness=1} after convert_material
    (RenderBridge.cpp:324-370): alpha 1, no specular lobe, refractive flag."""
    m = abi.default_material()
    m.base_color[:] = rgb
    m.roughness = 1.0
    m.specular = 0.0
    m.emission_strength = 1.0
    m.alpha_mode = abi.PHX_ALPHA_OPAQUE
    m.flags = abi.PHX_MATERIAL_FLAG_REFRACTIVE
    m.roughness_src_squared = 1
    return m


def emitter(radiance, front_only=True):
    m = diffuse((0.0, 0.0, 0.0))
    m.emission[:] = radiance
    m.emission_strength = 1.0
    m.flags = abi.PHX_MATERIAL_FLAG_REFRACTIVE | abi.PHX_MATERIAL_FLAG_EMITTER_ONLY
    if not front_only:
        m.flags |= abi.PHX_MATERIAL_FLAG_EMIT_BOTH_SIDES
    return m


def disney(base, roughness, metallic=0.0, spec_trans=0.0, clearcoat=0.0, anisotropic=0.0, sheen=0.0):
    """Disney parameters (perceptual roughness) at pt level: alpha = r^2."""
    m = abi.default_material()
    m.base_color[:] = base
    m.roughness = float(roughness) ** 2
    m.metallic = metallic
    m.transmission = spec_trans
    m.clearcoat = clearcoat
    m.anisotropic = anisotropic
    m.sheen = sheen
    m.specular = 0.5
    m.emission_strength = 1.0
    m.alpha_mode = abi.PHX_ALPHA_OPAQUE
    m.flags = abi.PHX_MATERIAL_FLAG_REFRACTIVE
    return m


# --------------------------------------------------------------------------- geometry helpers
def _quad(p0, e1, e2, flip=False):
    """Two triangles of the parallelogram p0 + [0,1] e1 + [0,1] e2; the normal is
    e1 x e2 (or its opposite with flip)."""
    p0, e1, e2 = (np.asarray(v, np.float64) for v in (p0, e1, e2))
    v = np.stack([p0, p0 + e1, p0 + e1 + e2, p0 + e2])
    t = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    if flip:
        t = t[:, ::-1].copy()
    return v, t


def _box(size, angle_deg, center_xz, y0=0.0):
    """Closed box of size (sx, sy, sz) standing on y0, rotated about y, 12 tris
    with outward normals."""
    sx, sy, sz = size
    a = math.radians(angle_deg)
    ca, sa = math.cos(a), math.sin(a)
    verts, tris = [], []
    hx, hz = sx / 2, sz / 2
    faces = [  # (p0, e1, e2) with e1 x e2 outward, in box-local coordinates
        ((-hx, 0, hz), (2 * hx, 0, 0), (0, sy, 0)),     # +z
        ((hx, 0, -hz), (-2 * hx, 0, 0), (0, sy, 0)),    # -z
        ((hx, 0, hz), (0, 0, -2 * hz), (0, sy, 0)),     # +x
        ((-hx, 0, -hz), (0, 0, 2 * hz), (0, sy, 0)),    # -x
        ((-hx, sy, hz), (2 * hx, 0, 0), (0, 0, -2 * hz)),  # +y
        ((-hx, 0, -hz), (2 * hx, 0, 0), (0, 0, 2 * hz)),   # -y
    ]
    rot = np.array([[ca, 0, sa], [0, 1, 0], [-sa, 0, ca]])
    for p0, e1, e2 in faces:
        v, t = _quad(p0, e1, e2)
        v = v @ rot.T + np.array([center_xz[0], y0, center_xz[1]])
        tris.append(t + sum(len(x) for x in verts))
        verts.append(v)
    return np.concatenate(verts), np.concatenate(tris)


class _MeshBuilder:
    def __init__(self):
        self.v, self.t, self.m, self.n = [], [], [], []
        self.nv = 0

    def add(self, verts, tris, mat, normals=None):
        self.v.append(np.asarray(verts, np.float64))
        self.t.append(np.asarray(tris, np.int64) + self.nv)
        self.m.append(np.full(len(tris), mat, np.uint32))
        self.n.append(normals)
        self.nv += len(verts)

    def build(self, materials, with_normals=False):
        pos = np.concatenate(self.v).astype(np.float32)
        idx = np.concatenate(self.t).astype(np.uint32)
        mid = np.concatenate(self.m)
        normals = None
        if with_normals:
            normals = np.concatenate([n if n is not None else np.zeros((len(v), 3)) for v, n in zip(self.v, self.n)])
            normals = normals.astype(np.float32)
        return Scene(pos, idx, materials, material_ids=mid, normals=normals)


# --------------------------------------------------------------------------- Cornell box
CORNELL_WHITE = (0.73, 0.73, 0.73)
CORNELL_RED = (0.65, 0.05, 0.05)
CORNELL_GREEN = (0.12, 0.45, 0.15)


def cornell_box(wall_colors=None, light_radiance=(17.0, 12.0, 4.0)):
    """The 36-triangle Cornell box: room x,z in [-1,1], y in [0,2], open at +z."""
    white, red, green = wall_colors or (CORNELL_WHITE, CORNELL_RED, CORNELL_GREEN)
    mats = [diffuse(white), diffuse(red), diffuse(green), emitter(light_radiance)]
    b = _MeshBuilder()
    # walls, normals pointing into the room
    b.add(*_quad((-1, 0, 1), (2, 0, 0), (0, 0, -2)), 0)          # floor (+y)
    b.add(*_quad((-1, 2, -1), (2, 0, 0), (0, 0, 2)), 0)          # ceiling (-y)
    b.add(*_quad((-1, 0, -1), (2, 0, 0), (0, 2, 0)), 0)          # back wall (+z)
    b.add(*_quad((-1, 0, 1), (0, 0, -2), (0, 2, 0)), 1)          # left wall, red (+x)
    b.add(*_quad((1, 0, -1), (0, 0, 2), (0, 2, 0)), 2)           # right wall, green (-x)
    b.add(*_box((0.6, 0.6, 0.6), -18.0, (0.38, 0.35)), 0)        # short box
    b.add(*_box((0.6, 1.2, 0.6), 18.0, (-0.35, -0.30)), 0)       # tall box
    # ceiling light facing down, slightly below the ceiling
    b.add(*_quad((-0.25, 1.999, -0.25), (0.5, 0, 0), (0, 0, 0.5)), 3)
    s = b.build(mats)
    s.camera = dict(look_from=(0.0, 1.0, 3.7), look_at=(0.0, 1.0, 0.0), vup=(0.0, 1.0, 0.0), vfov=40.0, aspect=1.0)
    s.name = "cornell36"
    assert s.triangle_count == 36
    return s


# --------------------------------------------------------------------------- icospheres + fBm
def icosphere(level):
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    for _ in range(level):
        edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        edges.sort(axis=1)
        uniq, inv = np.unique(edges, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        nf = len(f)
        a = base + inv[:nf]
        b = base + inv[nf:2 * nf]
        c = base + inv[2 * nf:]
        f = np.concatenate([np.stack([f[:, 0], a, c], 1), np.stack([f[:, 1], b, a], 1),
                            np.stack([f[:, 2], c, b], 1), np.stack([a, b, c], 1)])
    return v, f


def _hash3(ix, iy, iz, seed):
    h = (ix.astype(np.int64) * 73856093) ^ (iy.astype(np.int64) * 19349663) ^ (iz.astype(np.int64) * 83492791)
    h = (h ^ (seed * 2654435761)) & 0xFFFFFFFF
    h = (h ^ (h >> 13)) * 1274126177 & 0xFFFFFFFF
    h = h ^ (h >> 16)
    return (h & 0xFFFFFF).astype(np.float64) / float(0xFFFFFF) * 2.0 - 1.0


def value_noise3(p, seed):
    """Trilinear value noise in [-1, 1] with a smoothstep fade."""
    i = np.floor(p).astype(np.int64)
    f = p - i
    u = f * f * (3.0 - 2.0 * f)
    out = 0.0
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = ((u[..., 0] if dx else 1 - u[..., 0]) * (u[..., 1] if dy else 1 - u[..., 1]) *
                     (u[..., 2] if dz else 1 - u[..., 2]))
                out = out + w * _hash3(i[..., 0] + dx, i[..., 1] + dy, i[..., 2] + dz, seed)
    return out


def fbm3(p, octaves, seed, lacunarity=2.0, gain=0.5):
    amp, freq, total, norm = 1.0, 1.0, 0.0, 0.0
    for o in range(octaves):
        total = total + amp * value_noise3(p * freq, seed + 7919 * o)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return total / norm


def _vertex_normals(v, f):
    fn = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    n = np.zeros_like(v)
    for k in range(3):
        np.add.at(n, f[:, k], fn)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    return n


def sky_envmap(width=1024, height=512, sun_dir=(0.45, 0.55, -0.7), sun_sigma_deg=2.0, sun_intensity=5e3):
    """Sky gradient plus a Gaussian sun, float32 (H, W, 3), row 0 = +Y."""
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    theta = v * np.pi
    phi = u * 2 * np.pi - np.pi
    st = np.sin(theta)[:, None]
    d = np.stack([st * np.cos(phi)[None, :], np.cos(theta)[:, None] * np.ones((1, width)),
                  st * np.sin(phi)[None, :]], -1)
    up = np.clip(d[..., 1], -1, 1)
    zenith = np.array([0.25, 0.45, 0.95])
    horizon = np.array([0.95, 0.95, 1.0])
    ground = np.array([0.18, 0.16, 0.14])
    t = np.clip(up, 0, 1)[..., None] ** 0.5
    sky = horizon * (1 - t) + zenith * t
    rgb = np.where(up[..., None] >= 0, sky, ground * (0.5 + 0.5 * np.clip(1 + up, 0, 1))[..., None])
    sd = np.asarray(sun_dir, np.float64)
    sd /= np.linalg.norm(sd)
    ang = np.arccos(np.clip(d @ sd, -1, 1))
    sig = math.radians(sun_sigma_deg)
    sun = sun_intensity * np.exp(-0.5 * (ang / sig) ** 2)
    rgb = rgb + sun[..., None] * np.array([1.0, 0.93, 0.82])
    return rgb.astype(np.float32)


def sphere_grid(seed=1234, level=5, grid=6, spacing=1.0, radius=0.4, env=True, env_size=(1024, 512)):
    """6x6 displaced icospheres on a ground quad with 4 quad lights and a sky."""
    rng = np.random.default_rng(seed)
    base_v, base_f = icosphere(level)
    mats = [disney((0.5, 0.5, 0.5), 1.0)]  # ground
    b = _MeshBuilder()
    b.add(*_quad((-20.0, 0.0, 20.0), (40.0, 0.0, 0.0), (0.0, 0.0, -40.0)), 0,
          normals=np.tile([[0.0, 1.0, 0.0]], (4, 1)))
    off = (grid - 1) * spacing / 2.0
    for gz in range(grid):
        for gx in range(grid):
            k = gz * grid + gx
            disp = fbm3(base_v * 2.0 + rng.uniform(-100, 100, size=3), 4, seed + 31 * k)
            v = base_v * (radius * (1.0 + 0.05 * disp))[:, None]
            center = np.array([gx * spacing - off, radius * 1.05, gz * spacing - off])
            v = v + center
            n = _vertex_normals(v, base_f)
            base = tuple(rng.uniform(0, 1, 3))
            rough = rng.uniform(0.05, 1.0)
            metal = 1.0 if rng.uniform() < 0.3 else 0.0
            trans = 1.0 if rng.uniform() < 0.2 else 0.0
            cc = rng.uniform(0, 1)
            an = rng.uniform(0, 0.8)
            sh = rng.uniform(0, 0.5)
            mats.append(disney(base, rough, metal, trans, cc, an, sh))
            b.add(v, base_f, len(mats) - 1, normals=n)
    light_mat = len(mats)
    mats.append(emitter((12.0, 11.0, 10.0)))
    for (cx, cz) in ((-2.0, -2.0), (2.0, -2.0), (-2.0, 2.0), (2.0, 2.0)):
        b.add(*_quad((cx - 0.5, 4.0, cz - 0.3), (1.0, 0, 0), (0, 0, 0.6)), light_mat,
              normals=np.tile([[0.0, -1.0, 0.0]], (4, 1)))
    s = b.build(mats, with_normals=True)
    if env:
        s.env_pixels = sky_envmap(*env_size)
    s.camera = dict(look_from=(0.0, 4.2, 7.6), look_at=(0.0, 0.25, 0.0), vup=(0.0, 1.0, 0.0), vfov=40.0, aspect=1.0)
    s.name = f"spheres{grid}x{grid}_l{level}"
    return s


def random_rays(scene, n, seed=0, inside_fraction=0.5):
    """Rays for hit-parity tests: half from the camera towards random scene
    points, half between random points of the scene bounds."""
    rng = np.random.default_rng(seed)
    lo = scene.positions.min(0).astype(np.float64)
    hi = scene.positions.max(0).astype(np.float64)
    ext = hi - lo
    n_in = int(n * inside_fraction)
    tri = rng.integers(0, scene.triangle_count, n)
    v = scene.positions[scene.indices[tri]].astype(np.float64)
    r1 = np.sqrt(rng.uniform(size=n))
    r2 = rng.uniform(size=n)
    target = v[:, 0] * (1 - r1)[:, None] + v[:, 1] * (r1 * (1 - r2))[:, None] + v[:, 2] * (r1 * r2)[:, None]
    org = lo + rng.uniform(size=(n, 3)) * ext
    if scene.camera is not None:
        org[n_in:] = np.asarray(scene.camera["look_from"]) + rng.normal(scale=0.05, size=(n - n_in, 3))
    d = target - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    tmin = np.zeros(n, np.float32)
    tmax = np.full(n, 1e30, np.float32)
    return org.astype(np.float32), d.astype(np.float32), tmin, tmax


# --------------------------------------------------------------------------- fast lattice fBm
def _lattice_noise_grid(n, freq, seed):
    """Value noise sampled on an n^3 grid over [0, 1)^3 at `freq` lattice cells
    per unit, with smoothstep-faded trilinear weights (separable)."""
    rng = np.random.default_rng(seed)
    m = int(math.ceil(freq)) + 2
    lat = rng.uniform(-1.0, 1.0, (m, m, m))
    x = (np.arange(n) + 0.5) / n * freq
    i = np.floor(x).astype(np.int64)
    f = x - i
    w = f * f * (3.0 - 2.0 * f)
    # interpolate axis by axis: lat[z, y, x]
    a = lat[:, :, i] * (1 - w) + lat[:, :, i + 1] * w            # (m, m, n)
    a = a[:, i, :] * (1 - w)[None, :, None] + a[:, i + 1, :] * w[None, :, None]  # (m, n, n)
    a = a[i, :, :] * (1 - w)[:, None, None] + a[i + 1, :, :] * w[:, None, None]  # (n, n, n)
    return a


def fbm_grid(n, octaves, seed, base_freq=4.0):
    total = np.zeros((n, n, n))
    amp, norm = 1.0, 0.0
    for o in range(octaves):
        total += amp * _lattice_noise_grid(n, base_freq * 2 ** o, seed + 101 * o)
        norm += amp
        amp *= 0.5
    return (total / norm).astype(np.float32)  # in about [-1, 1], z-major like the volume values


# --------------------------------------------------------------------------- gradient workload
def textured_cornell(n_materials=32, tex_size=512, texture=None, seed=0):
    """Cornell box whose floor is textured (512^2 RGB albedo) and whose other
    32 triangles cycle through the remaining materials; material 0 is the
    textured floor, materials 1..n-1 Disney diffuse/metal, the light last."""
    base = cornell_box()
    rng = np.random.default_rng(seed)
    mats = []
    floor = disney((0.5, 0.5, 0.5), 0.8)
    floor.base_color_tex = 0
    mats.append(floor)
    for k in range(1, n_materials):
        mats.append(disney(tuple(rng.uniform(0.2, 0.8, 3)), float(rng.uniform(0.3, 1.0)), float(rng.uniform(0, 0.3))))
    light = emitter((17.0, 12.0, 4.0))
    mats.append(light)
    mid = np.zeros(base.triangle_count, np.uint32)
    light_tris = np.nonzero(base.material_ids == 3)[0]
    floor_tris = np.array([0, 1])
    others = [t for t in range(base.triangle_count) if t not in set(light_tris) | set(floor_tris)]
    for k, t in enumerate(others):
        mid[t] = 1 + (k % (n_materials - 1))
    mid[light_tris] = n_materials
    uv = np.zeros((base.positions.shape[0], 2), np.float32)
    # floor quad vertices 0..3 span x, z in [-1, 1]
    fv = base.positions[:4]
    uv[:4, 0] = (fv[:, 0] + 1) / 2
    uv[:4, 1] = (fv[:, 2] + 1) / 2
    s = Scene(base.positions, base.indices, mats, material_ids=mid, texcoords=uv)
    tex = np.full((tex_size, tex_size, 3), 0.5, np.float32) if texture is None else texture
    s.add_texture(tex)
    s.camera = base.camera
    s.name = "cornell_textured"
    return s


def _lattice_noise_2d(n, freq, seed):
    rng = np.random.default_rng(seed)
    m = int(math.ceil(freq)) + 2
    lat = rng.uniform(-1.0, 1.0, (m, m))
    x = (np.arange(n) + 0.5) / n * freq
    i = np.floor(x).astype(np.int64)
    f = x - i
    w = f * f * (3.0 - 2.0 * f)
    a = lat[:, i] * (1 - w) + lat[:, i + 1] * w
    return a[i, :] * (1 - w)[:, None] + a[i + 1, :] * w[:, None]


def colored_noise_texture(size, seed, octaves=3):
    """Seeded 3-octave colored value noise in (0, 1), (size, size, 3)."""
    out = np.zeros((size, size, 3), np.float32)
    for c in range(3):
        g, amp, norm = 0.0, 1.0, 0.0
        for o in range(octaves):
            g = g + amp * _lattice_noise_2d(size, 6.0 * 2 ** o, seed + 17 * c + 101 * o)
            norm += amp
            amp *= 0.5
        out[..., c] = 0.5 + 0.45 * g / norm
    return np.clip(out, 0.02, 0.98)


# --------------------------------------------------------------------------- volume workload
def fog_box(n=256, seed=7, threshold=0.4, scale=20.0, macro=32, albedo=0.9, env=True, env_size=(1024, 512),
            majorants_fn=None, tf_size=256):
    """A 256^3 fBm density grid in an open box: floor, back and left walls,
    one quad area light and the sky environment.  The transfer function maps
    the normalized field to extinction (0 below `threshold`, ramping to 1),
    density_scale = `scale`; the scattering albedo is `albedo`, sampled in
    `tf_size` entries. `majorants_fn(volume, values, tf)` builds the macro-cell grid (the
    library's phx_volume_majorants is used when it is None)."""
    b = _MeshBuilder()
    mats = [disney((0.6, 0.6, 0.6), 1.0), emitter((30.0, 28.0, 25.0))]
    b.add(*_quad((-2, -1, 2), (4, 0, 0), (0, 0, -4)), 0)      # floor
    b.add(*_quad((-2, -1, -2), (4, 0, 0), (0, 3, 0)), 0)      # back wall
    b.add(*_quad((-2, -1, 2), (0, 0, -4), (0, 3, 0)), 0)      # left wall
    b.add(*_quad((0.2, 1.8, -0.6), (1.0, 0, 0), (0, 0, 1.0)), 1)  # light facing down
    s = b.build(mats)
    g = fbm_grid(n, 6, seed)
    # normalized field: the densest 35% of the grid lies above `threshold`
    q = np.quantile(g[::4, ::4, ::4], [0.65, 0.999])
    v = np.clip(threshold + (g - q[0]) * (1.0 - threshold) / (q[1] - q[0]), 0.0, 1.0).astype(np.float32).reshape(-1)
    vol = abi.Volume()
    vol.world_to_local[:] = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]
    vol.origin[:] = (-1.0, -1.0, -1.0)
    vol.spacing[:] = (2.0 / (n - 1),) * 3
    vol.dims[:] = (n, n, n)
    vol.data_offset = 0
    vol.tf_offset = 0
    vol.value_lo, vol.value_hi = 0.0, 1.0
    vol.density_scale = scale
    vol.linear = 1
    vol.macro_dims[:] = (macro, macro, macro)
    x = np.linspace(0, 1, tf_size)
    tf = np.zeros((tf_size, 4), np.float32)
    tf[:, :3] = albedo
    tf[:, 3] = np.clip((x - threshold) / (1 - threshold), 0, 1)
    vol.tf_size = tf_size
    if majorants_fn is None:
        from ._lib import volume_majorants
        majorants_fn = volume_majorants
    maj = majorants_fn(vol, v, tf)
    s.volumes = [vol]
    s.volume_values = v
    s.volume_tf = tf.reshape(-1)
    s.volume_majorants = maj
    if env:
        s.env_pixels = sky_envmap(*env_size)
    s.camera = dict(look_from=(0.0, 0.6, 4.2), look_at=(0.0, -0.1, 0.0), vup=(0.0, 1.0, 0.0), vfov=45.0, aspect=1.0)
    s.name = f"fog{n}"
    return s


# --------------------------------------------------------------------------- batched workloads
def orbit_eye(scene, angle):
    """The scene's camera position turned by `angle` about the vertical axis
    through its look-at point."""
    eye = np.asarray(scene.camera["look_from"], np.float64)
    at = np.asarray(scene.camera["look_at"], np.float64)
    d = eye - at
    c, sn = np.cos(angle), np.sin(angle)
    return at + np.array([c * d[0] + sn * d[2], d[1], -sn * d[0] + c * d[2]])


def batch_draw(seed=2024, views=16):
    """The batched workload's seeded draw: resolutions from {64..1024}^2 and
    spp from {4..64}; the first half of the views are the 737K-triangle
    sphere grid from orbit angles in [0, 2 pi), the second half Cornell boxes
    with random wall colours from angles in [-0.3, 0.3].  Returns one dict per
    view: scene ("spheres" or "cornell"), res, spp, angle, walls."""
    rng = np.random.default_rng(seed)
    res = rng.choice([64, 128, 256, 512, 1024], views)
    spp = rng.choice([4, 8, 16, 32, 64], views)
    out = []
    for v in range(views):
        if v < views // 2:
            out.append(dict(scene="spheres", res=int(res[v]), spp=int(spp[v]), angle=float(rng.uniform(0, 2 * np.pi)),
                            walls=None))
        else:
            walls = tuple(tuple(rng.uniform(0.05, 0.8, 3)) for _ in range(3))
            out.append(dict(scene="cornell", res=int(res[v]), spp=int(spp[v]), angle=float(rng.uniform(-0.3, 0.3)),
                            walls=walls))
    return out

