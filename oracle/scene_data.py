"""ORACLE - test infrastructure only.

The render bridge's scene import and camera set-up restated on the CPU, as
the checker of phx_set_scene_data and phx_bridge_render:
  scene_from_data  Renderer::Impl::build_scene (RenderBridge.cpp:441-826) ->
                   a pt-level photon_amd.scenes.Scene the oracle renders
  convert_material convert_source / convert_material (RenderBridge.cpp:314-370)
  bridge_view      the camera and RenderParams of Renderer::Impl::render
                   (RenderBridge.cpp:871-920)
Arithmetic that the bridge does in float is done in numpy float32.
"""
import numpy as np

from photon_amd import abi

f32 = np.float32


def convert_material(md):
    """RenderBridge.cpp:324-370 on an abi.BridgeMaterial."""
    m = abi.default_material()
    m.base_color[:] = tuple(md.base_color)
    m.metallic = md.metallic
    m.roughness = float(f32(md.roughness) * f32(md.roughness))
    m.ior = md.ior
    m.transmission = md.transmission
    m.specular = float(f32(0.5) * f32(md.specular))
    m.clearcoat = md.clearcoat
    m.clearcoat_roughness = md.clearcoat_roughness
    m.sheen = md.sheen
    m.emission[:] = tuple(md.emission)
    m.emission_strength = 1.0
    m.alpha = md.opacity
    pairs = (("base_color_src", "base_color_src"), ("opacity_src", "alpha_src"), ("metallic_src", "metallic_src"),
             ("roughness_src", "roughness_src"), ("emission_src", "emission_src"),
             ("transmission_src", "transmission_src"), ("normal_src", "normal_src"))
    for src, dst in pairs:
        s = getattr(md, src)
        kind = 1 if s.kind == 1 else (2 if s.kind == 2 else 0)
        getattr(m, dst).kind, getattr(m, dst).index = kind, s.index
    m.roughness_src_squared = 1
    m.base_color_alpha = 1 if md.base_color_alpha else 0
    if m.alpha_src.kind != 0:
        m.alpha = 1.0
    m.alpha_mode = {0: abi.PHX_ALPHA_BLEND, 2: abi.PHX_ALPHA_MASK}.get(md.alpha_mode, abi.PHX_ALPHA_OPAQUE)
    m.alpha_cutoff = md.alpha_cutoff
    m.thin = 1 if md.thin_walled else 0
    flags = abi.PHX_MATERIAL_FLAG_REFRACTIVE
    if md.emitter_only:
        flags |= abi.PHX_MATERIAL_FLAG_EMITTER_ONLY
    if md.hidden_to_camera:
        flags |= abi.PHX_MATERIAL_FLAG_HIDDEN_TO_CAMERA
    if md.emission_side == 2:
        flags |= abi.PHX_MATERIAL_FLAG_EMIT_BOTH_SIDES
    elif md.emission_side == 1:
        flags |= abi.PHX_MATERIAL_FLAG_EMIT_BACK_ONLY
    m.flags = flags
    return m


def scene_from_data(sd):
    """Impl::build_scene (RenderBridge.cpp:441-826) on a renderer.SceneData."""
    import oracle
    from photon_amd import scenes
    nv, nt = sd.positions.shape[0], sd.indices.shape[0]
    empty = nt == 0 or nv == 0
    if empty:  # one degenerate triangle (:452-456)
        pos = np.zeros((3, 3), np.float32)
        idx = np.array([[0, 1, 2]], np.uint32)
        nv, nt = 3, 1
    else:
        pos, idx = sd.positions.astype(np.float32), sd.indices.astype(np.uint32)
    mc = max(len(sd.materials), 1)
    mids = np.zeros(nt, np.uint32)
    if not empty and sd.material_ids is not None:
        k = min(nt, sd.material_ids.size)
        mids[:k] = sd.material_ids[:k]
    mids[mids >= mc] = 0
    normals = sd.normals if (not empty and sd.normals is not None and sd.normals.size >= nv * 3) else None
    mats = [convert_material(m) for m in sd.materials] or [abi.default_material()]
    s = scenes.Scene(pos, idx, mats, material_ids=mids, normals=normals)
    for slot in range(5):
        a = sd.attributes[slot]
        if not empty and a is not None and a.size >= nv * 4:
            s.attributes[slot] = a.reshape(-1, 4)[:nv]
    # samplers (:527-577)
    texels = []
    off = 0
    for h in sd.samplers:
        d = h.struct
        typ = {0: abi.PHX_SAMPLER_IMAGE1D, 2: abi.PHX_SAMPLER_IMAGE3D, 3: abi.PHX_SAMPLER_TRANSFORM}.get(
            d.type, abi.PHX_SAMPLER_IMAGE2D)
        size = tuple(max(int(x), 1) for x in d.size)
        wrap = tuple({1: abi.PHX_WRAP_REPEAT, 2: abi.PHX_WRAP_MIRROR}.get(int(w), abi.PHX_WRAP_CLAMP) for w in d.wrap)
        tx = h.arrays[0]
        n = 0 if tx is None else tx.size // 4
        if typ != abi.PHX_SAMPLER_TRANSFORM and n < size[0] * size[1] * size[2]:
            typ = abi.PHX_SAMPLER_TRANSFORM
        s.samplers.append(abi.sampler_info(type_=typ, in_attribute=d.in_attribute, linear=1 if d.linear else 0,
                                           wrap=wrap, offset=off, size=size, in_transform=tuple(d.in_transform),
                                           in_offset=tuple(d.in_offset), out_transform=tuple(d.out_transform),
                                           out_offset=tuple(d.out_offset)))
        if n:
            texels.append(tx.reshape(-1, 4)[:n])
        off += n
    if texels:
        s.sampler_texels = np.concatenate(texels).astype(np.float32)
    # lights (:579-597)
    for ld in sd.lights:
        l = abi.Light()
        l.type = {1: abi.PHX_LIGHT_DIRECTIONAL, 2: abi.PHX_LIGHT_SPOT}.get(ld.type, abi.PHX_LIGHT_POINT)
        l.position[:], l.direction[:], l.color[:] = tuple(ld.position), tuple(ld.direction), tuple(ld.color)
        l.intensity, l.radius, l.spot_angle, l.spot_falloff = ld.intensity, ld.radius, ld.spot_angle, ld.spot_falloff
        s.lights.append(l)
    # environment (:599-624): texels as given (non-finite and negative ones as 0)
    env = sd.environment
    if env is not None:
        e = env.struct
        px = env.arrays[0]
        if e.present and e.width > 0 and e.height > 0 and px.size >= e.width * e.height * 3:
            img = px[:e.width * e.height * 3].reshape(e.height, e.width, 3).astype(np.float32)
            s.env_pixels = np.where(np.isfinite(img) & (img > 0), img, 0).astype(np.float32)
            s.env_world_to_tex = np.array(list(e.world_to_map), np.float32).reshape(3, 3)
            s.env_visible = bool(e.visible)
    # volumes (:626-778), majorants by the oracle's restatement of the bridge's rule
    vals, tfs, majs = [], [], []
    voff = toff = moff = 0
    for h in sd.volumes:
        d = h.struct
        values, transfer = h.arrays
        v = abi.Volume()
        v.world_to_local[:] = tuple(d.world_to_local)
        v.origin[:], v.spacing[:] = tuple(d.origin), tuple(d.spacing)
        v.dims[:] = tuple(max(int(x), 1) for x in d.dims)
        v.data_offset, v.tf_offset, v.tf_size = voff, toff, transfer.size // 4
        v.value_lo, v.value_hi = d.value_range[0], d.value_range[1]
        v.density_scale = d.density_scale
        v.linear = 1 if d.linear else 0
        m = oracle.volume_majorants(v, values, transfer[:v.tf_size * 4])
        v.macro_offset = moff
        vals.append(values)
        tfs.append(transfer[:v.tf_size * 4])
        majs.append(m)
        voff += values.size
        toff += v.tf_size
        moff += m.size
        s.volumes.append(v)
    if s.volumes:
        s.volume_values = np.concatenate(vals).astype(np.float32)
        s.volume_tf = np.concatenate(tfs).astype(np.float32)
        s.volume_majorants = np.concatenate(majs).astype(np.float32)
    # ray offset from the bounds of every vertex (:815-820)
    eps = f32(1e-3)
    if not empty:
        lo, hi = pos.min(0), pos.max(0)
        dg = (hi - lo).astype(np.float32)
        e = f32(1e-4) * f32(np.sqrt(f32(dg[0] * dg[0] + dg[1] * dg[1] + dg[2] * dg[2])))
        eps = min(max(e, f32(1e-6)), f32(1e-2))
    s.ray_epsilon = float(eps)
    return s


def bridge_view(cam, settings, sample_offset=0):
    """The pt-level camera and RenderParams of Renderer::Impl::render
    (RenderBridge.cpp:871-920) for an abi.CameraData and abi.RenderSettings."""
    import oracle
    w, h = settings.width, settings.height
    pos = np.array(list(cam.position), np.float32)
    d = np.array(list(cam.direction), np.float32)
    ln = f32(np.sqrt(f32(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])))
    d = (d * (f32(1) / ln)).astype(np.float32) if ln > 0 else np.array([0, 0, -1], np.float32)
    up = np.array(list(cam.up), np.float32)
    cr = np.cross(d.astype(np.float64), up.astype(np.float64))
    if np.linalg.norm(cr) < 1e-6:
        up = np.array([0, 1, 0] if abs(d[1]) < 0.9 else [1, 0, 0], np.float32)
    aspect = cam.aspect if cam.aspect > 0 else float(f32(w) / f32(h))
    at = (pos + d).astype(np.float32)
    if cam.orthographic:
        c = _orthographic(pos, at, up, cam.height, aspect)
    else:
        fov = float(f32(f32(cam.fovy) * f32(180.0)) / f32(3.14159265358979))
        lens = cam.aperture_radius > 0 and cam.focus_distance > 0
        c = oracle.camera_perspective(pos, at, up, fov, aspect, float(f32(2) * f32(cam.aperture_radius)) if lens else 0.0,
                                      cam.focus_distance if lens else 1.0)
    c.clip_near = cam.clip_near if cam.clip_near > 0 else 0.0
    c.clip_far = cam.clip_far if cam.clip_far < 1e30 else 1e30
    p = abi.default_render_params(w, h, max(settings.pixel_samples, 1), max(settings.max_depth, 1))
    p.sample_offset = sample_offset
    p.background_color[:] = tuple(settings.background[:3])
    p.background_alpha = settings.background[3]
    p.ambient_color[:] = tuple(settings.ambient_color)
    p.ambient_radiance = settings.ambient_radiance
    p.throughput_clamp = settings.throughput_clamp
    p.max_diffuse_bounces = settings.max_diffuse_bounces
    p.max_glossy_bounces = settings.max_glossy_bounces
    p.max_transmission_bounces = settings.max_transmission_bounces
    p.max_volume_bounces = settings.max_volume_bounces
    p.clamp_direct_sum = float(f32(3) * f32(max(settings.clamp_direct, 0.0)))
    p.clamp_indirect_sum = float(f32(3) * f32(max(settings.clamp_indirect, 0.0)))
    p.bottom_up = 1
    p.image_region[:] = tuple(cam.image_region)
    return c, p


def _orthographic(pos, at, up, height, aspect):
    """Camera::make_orthographic (camera.h:57-77)."""
    import oracle
    c = oracle.camera_perspective(pos, at, up, 90.0, aspect, 0.0, 1.0)
    uu, vv = np.array(list(c.u), np.float32), np.array(list(c.v), np.float32)
    hor = (uu * f32(f32(height) * f32(aspect))).astype(np.float32)
    ver = (vv * f32(height)).astype(np.float32)
    c.horizontal[:] = tuple(float(x) for x in hor)
    c.vertical[:] = tuple(float(x) for x in ver)
    ll = ((pos - hor * f32(0.5)) - ver * f32(0.5)).astype(np.float32)
    c.lower_left[:] = tuple(float(x) for x in ll)
    c.focus_dist = 1.0
    c.lens_radius = 0.0
    c.is_ortho = 1
    return c
