"""GPU parity, through the C ABI, of the shading features the workloads leave
unused: analytic lights (point, spot, directional, area), alpha blend and
mask and the roughness / metallic / alpha / normal textures, light and
throughput clamps with finite bounce limits per kind, thin-lens,
orthographic, clipped and region cameras with bottom-up rows, parameter
sources (vertex attributes, image and transform samplers, sourced normal maps,
emission varying over an emitter), the background image, and the render
bridge's scene import (phx_set_scene_data) with lights, a rotated
environment, a volume and samplers.  Bars as in test_gpu_parity: relMSE
< 1e-3 against the oracle on the same sample streams and >= 90 % of pixels
equal to 1e-4 (relative), AOVs as exact as the colour.
"""
import math

import numpy as np
import pytest

import oracle
from oracle import scene_data as osd
from photon_amd import abi, scenes
from tests.helpers import rel_mse

pytestmark = pytest.mark.gpu


def _render_pair(ctx, s, cam, params, same_bar=0.9, brute=False):
    from photon_amd.backend import HipBackend, PathTracer
    pt = PathTracer(HipBackend(context=ctx))
    pt.set_scene(s)
    pt.camera, pt.params = cam, params
    res = pt.render()
    ref, ex, _ = oracle.OracleScene(s, brute=brute).render(cam, params, aovs=True)
    err = rel_mse(res.color, ref)
    same = np.mean(np.abs(res.color - ref) <= 1e-4 * (1 + np.abs(ref)))
    print(f"{s.name}: relMSE {err:.3e}, pixels equal to 1e-4 {same:.4f}, mean {ref.mean():.4f}", flush=True)
    assert np.all(np.isfinite(res.color))
    assert err < 1e-3
    assert same >= same_bar
    np.testing.assert_allclose(res.alpha, ex["alpha"], atol=1e-5)
    hit = ex["depth"] < 1e29
    assert np.mean(np.abs(res.depth[hit] - ex["depth"][hit]) <= 1e-4 * ex["depth"][hit]) > 0.99
    assert np.mean(np.all(np.abs(res.normal - ex["normal"]) <= 1e-4, axis=-1)) > 0.99
    assert np.mean(np.all(np.abs(res.albedo - ex["albedo"]) <= 1e-4, axis=-1)) > 0.99
    return res, ref


def _cornell_cam(w, h, **kw):
    from photon_amd.backend import make_camera
    return make_camera((0.0, 1.0, 3.7), (0.0, 1.0, 0.0), (0.0, 1.0, 0.0), 40.0, w / h, **kw)


def _light(**kw):
    l = abi.Light()
    l.color[:] = (1.0, 1.0, 1.0)
    l.intensity = 1.0
    l.direction[:] = (0.0, -1.0, 0.0)
    for k, v in kw.items():
        if k in ("position", "direction", "color", "edge1", "edge2"):
            getattr(l, k)[:] = v
        else:
            setattr(l, k, v)
    return l


def lights_cornell():
    """The Cornell box lit by analytic lights only: its ceiling quad does not
    emit; a point light of radius 0.05, a spot light, a directional light of
    angular radius 0.05 through the open front and an area light below the
    ceiling (no mesh primitives, so it stays an analytic light)."""
    s = scenes.cornell_box()
    s.materials[3] = scenes.diffuse((0.1, 0.1, 0.1))
    d = np.array([-0.2, -0.35, -1.0])
    d /= np.linalg.norm(d)
    s.lights = [
        _light(type=abi.PHX_LIGHT_POINT, position=(-0.3, 1.7, 0.2), intensity=2.0, radius=0.05, color=(1, 0.9, 0.8)),
        _light(type=abi.PHX_LIGHT_SPOT, position=(0.5, 1.9, 0.3), direction=(0, -1, 0), intensity=6.0, spot_angle=0.6,
               spot_falloff=0.2, color=(0.8, 0.9, 1.0)),
        _light(type=abi.PHX_LIGHT_DIRECTIONAL, direction=tuple(d), intensity=1.5, radius=0.05, color=(1, 1, 0.9)),
        _light(type=abi.PHX_LIGHT_AREA, position=(-0.25, 1.95, -0.25), edge1=(0.5, 0, 0), edge2=(0, 0, 0.5),
               intensity=8.0, color=(1.0, 0.8, 0.6)),
    ]
    s.name = "cornell_analytic_lights"
    return s


def test_analytic_lights_match_oracle(gpu_ctx):
    s = lights_cornell()
    res, ref = _render_pair(gpu_ctx, s, _cornell_cam(48, 48), abi.default_render_params(48, 48, 8, 6))
    assert ref.mean() > 0.02  # the analytic lights do light the box


def textures_cornell():
    """Textured Cornell box with every texture kind of resolve_material and
    pt_shade: roughness, metallic and alpha textures, alpha blend and mask,
    a tangent-space normal texture, texture coordinates on every wall."""
    s = scenes.textured_cornell(n_materials=8, tex_size=64, texture=scenes.colored_noise_texture(64, 4), seed=2)
    p = s.positions
    uv = np.stack([(p[:, 0] + p[:, 2]) * 0.5 + 0.5, p[:, 1] * 0.5], axis=1).astype(np.float32)
    uv[:4] = s.texcoords[:4]
    s.texcoords = uv
    noise = scenes.colored_noise_texture(64, 11)
    rough = s.add_texture(noise[..., [0, 0, 0]])
    metal = s.add_texture(noise[..., [1, 1, 1]])
    alpha = s.add_texture(noise[..., [2, 2, 2]])
    nmap = noise.copy()
    nmap[..., 0] = 0.5 + 0.6 * (noise[..., 0] - 0.5)
    nmap[..., 1] = 0.5 + 0.6 * (noise[..., 1] - 0.5)
    nmap[..., 2] = 1.0
    normal = s.add_texture(nmap)
    m = s.materials
    m[1].roughness_tex = rough
    m[2].metallic_tex = metal
    m[2].roughness = 0.3
    m[3].alpha_tex = alpha
    m[3].alpha_mode = abi.PHX_ALPHA_BLEND
    m[4].alpha_tex = alpha
    m[4].alpha_mode = abi.PHX_ALPHA_MASK
    m[4].alpha_cutoff = 0.5
    m[5].normal_tex = normal
    m[6].roughness_tex = rough
    m[6].metallic_tex = metal
    m[6].normal_tex = normal
    s.name = "cornell_all_textures"
    return s


def test_textures_alpha_modes_and_normal_maps_match_oracle(gpu_ctx):
    s = textures_cornell()
    _render_pair(gpu_ctx, s, _cornell_cam(48, 48), abi.default_render_params(48, 48, 8, 6))


@pytest.mark.parametrize("which", ["cornell", "spheres"])
def test_clamps_and_bounce_limits_match_oracle(gpu_ctx, which):
    # Cycles-style clamps (pathtracer.cpp:40-47), the throughput clamp and a
    # finite limit for every bounce kind (count_bounce, :51-58)
    if which == "cornell":
        s = scenes.cornell_box()
        cam = _cornell_cam(48, 48)
    else:
        s = scenes.sphere_grid(seed=9, level=2, grid=3, env_size=(128, 64))
        from photon_amd.backend import make_camera
        cam = make_camera(s.camera["look_from"], s.camera["look_at"], s.camera["vup"], s.camera["vfov"], 1.0)
    p = abi.default_render_params(48, 48, 8, 8)
    p.clamp_direct_sum = 1.5
    p.clamp_indirect_sum = 0.4
    p.throughput_clamp = 1.2
    p.max_diffuse_bounces = 2
    p.max_glossy_bounces = 1
    p.max_transmission_bounces = 2
    p.max_volume_bounces = 1
    s.name = f"{which}_clamped"
    res, ref = _render_pair(gpu_ctx, s, cam, p)
    # the clamps and limits do bite: the unclamped frame differs
    q = abi.default_render_params(48, 48, 8, 8)
    free, _, _ = oracle.OracleScene(s).render(cam, q)
    assert rel_mse(free, ref) > 1e-3


def test_fog_volume_bounce_limit_matches_oracle(gpu_ctx):
    s = scenes.fog_box(n=32, macro=8, env_size=(64, 32))
    from photon_amd.backend import make_camera
    cam = make_camera(s.camera["look_from"], s.camera["look_at"], s.camera["vup"], s.camera["vfov"], 1.0)
    p = abi.default_render_params(40, 40, 4, 12)
    p.max_volume_bounces = 2
    p.clamp_indirect_sum = 2.0
    s.name = "fog_volume_limit"
    _render_pair(gpu_ctx, s, cam, p)


@pytest.mark.parametrize("tf_size", [257, 512])
def test_transfer_function_beyond_lds_matches_oracle_and_batch(gpu_ctx, tf_size):
    # transfer functions of more than 256 entries stay in global memory (up
    # to 256 are staged in LDS): the fog box with a 257- and a 512-entry
    # table, whose albedo varies along the table so that a wrong entry shows,
    # against the oracle, and as a one-view batch against the plain render
    from photon_amd.backend import BatchRenderer, Context, make_camera
    s = scenes.fog_box(n=32, macro=8, env_size=(64, 32), tf_size=tf_size)
    tf = s.volume_tf.reshape(-1, 4)
    tf[:, 0] = np.linspace(0.3, 0.95, tf_size)
    tf[:, 2] = np.linspace(0.95, 0.4, tf_size)
    s.volume_tf = tf.reshape(-1)
    s.name = f"fog_tf{tf_size}"
    cam = make_camera(s.camera["look_from"], s.camera["look_at"], s.camera["vup"], s.camera["vfov"], 1.0)
    p = abi.default_render_params(40, 40, 4, 12)
    res, _ = _render_pair(gpu_ctx, s, cam, p)
    from photon_amd.backend import HipBackend, PathTracer
    pt = PathTracer(HipBackend(context=Context(0)))
    pt.set_scene(s)
    br = BatchRenderer(Context(0))
    br.render([(pt, cam, p)])
    np.testing.assert_array_equal(br.read(0).color, res.color)


@pytest.mark.parametrize("kind", ["thin_lens", "orthographic", "clipped", "region_bottom_up"])
def test_cameras_match_oracle(gpu_ctx, kind):
    from photon_amd.backend import make_orthographic_camera
    s = scenes.cornell_box()
    p = abi.default_render_params(48, 40, 8, 6)
    if kind == "thin_lens":
        cam = _cornell_cam(48, 40, aperture=0.25, focus_dist=3.0)
    elif kind == "orthographic":
        cam = make_orthographic_camera((0.0, 1.0, 3.7), (0.0, 1.0, 0.0), (0.0, 1.0, 0.0), 2.4, 48 / 40)
    elif kind == "clipped":
        cam = _cornell_cam(48, 40)
        cam.clip_near, cam.clip_far = 3.2, 5.5  # the near plane cuts into the boxes, the far one the back wall
    else:
        cam = _cornell_cam(48, 40)
        p.image_region[:] = (0.2, 0.1, 0.7, 0.8)
        p.bottom_up = 1
    s.name = f"cornell_{kind}"
    _render_pair(gpu_ctx, s, cam, p)


def sources_cornell():
    """Cornell box whose materials take parameters from vertex attributes and
    samplers: base colour and alpha from a colour attribute (base_color_alpha,
    alpha blend), roughness from a 2-D image sampler over a coordinate
    attribute (perceptual, squared), metallic from a transform of the world
    position, a sourced normal map, and the ceiling light's emission from an
    attribute, so it varies over the emitter."""
    s = scenes.cornell_box()
    p = s.positions.astype(np.float64)
    nv = p.shape[0]
    col = np.stack([0.2 + 0.3 * (p[:, 0] + 1), 0.2 + 0.3 * p[:, 1], 0.2 + 0.3 * (p[:, 2] + 1),
                    0.55 + 0.45 * np.cos(3 * p[:, 0])], axis=1)
    crd = np.stack([(p[:, 0] + p[:, 2]) * 0.7 + 0.3, p[:, 1] * 0.6, np.zeros(nv), np.ones(nv)], axis=1)
    emis = np.stack([12 + 40 * (p[:, 0] + 0.25), 10 + 20 * (p[:, 2] + 0.25), np.full(nv, 4.0), np.ones(nv)], axis=1)
    s.attributes = {0: col.astype(np.float32), 1: crd.astype(np.float32), 2: emis.astype(np.float32)}
    tex = scenes.colored_noise_texture(16, 21)
    rgba = np.concatenate([tex, np.ones((16, 16, 1), np.float32)], axis=2).reshape(-1, 4)
    nmap = rgba.copy()
    nmap[:, 0] = 0.5 + 0.5 * (rgba[:, 0] - 0.5)
    nmap[:, 1] = 0.5 + 0.5 * (rgba[:, 1] - 0.5)
    nmap[:, 2] = 1.0
    s.sampler_texels = np.concatenate([rgba, nmap]).astype(np.float32)
    rep = (abi.PHX_WRAP_REPEAT, abi.PHX_WRAP_MIRROR, abi.PHX_WRAP_CLAMP)
    s.samplers = [
        abi.sampler_info(type_=abi.PHX_SAMPLER_IMAGE2D, in_attribute=1, linear=1, wrap=rep, size=(16, 16, 1)),
        abi.sampler_info(type_=abi.PHX_SAMPLER_TRANSFORM, in_attribute=abi.PHX_ATTRIBUTE_WORLD_POSITION,
                         in_transform=(0, 0, 0, 0, 0, 0.4, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1),
                         out_transform=(0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1)),
        abi.sampler_info(type_=abi.PHX_SAMPLER_IMAGE2D, in_attribute=1, linear=0, wrap=(1, 1, 0), size=(16, 16, 1),
                         offset=256),
    ]
    white = s.materials[0]
    for m, (src, kind, idx) in zip([white, s.materials[1], s.materials[2]],
                                   [("base_color_src", 1, 0), ("roughness_src", 2, 0), ("metallic_src", 2, 1)]):
        getattr(m, src).kind, getattr(m, src).index = kind, idx
        m.roughness_src_squared = 1
    white.base_color_alpha = 1
    white.alpha_mode = abi.PHX_ALPHA_BLEND
    s.materials[2].specular = 0.5
    s.materials[1].normal_src.kind, s.materials[1].normal_src.index = 2, 2
    light = s.materials[3]
    light.emission_src.kind, light.emission_src.index = 1, 2
    s.name = "cornell_sources"
    return s


def test_parameter_sources_match_oracle(gpu_ctx):
    # The walls are mostly transparent here (their corners' alpha is 0.1), so
    # paths reach the boxes' bottom faces, which lie in the floor's plane: a
    # ray meets both at bitwise the same t.  The library gives such ties to
    # the larger primitive id, as a brute-force scan does; the reference's
    # tree gives them to whichever it visits last, which depends on its tree.
    # The oracle traces by brute force here, so ties fall the same way.
    s = sources_cornell()
    _render_pair(gpu_ctx, s, _cornell_cam(48, 48), abi.default_render_params(48, 48, 8, 6), brute=True)
    ref_tree, _, _ = oracle.OracleScene(s).render(_cornell_cam(48, 48), abi.default_render_params(48, 48, 8, 6))
    ref_scan, _, _ = oracle.OracleScene(s, brute=True).render(_cornell_cam(48, 48),
                                                               abi.default_render_params(48, 48, 8, 6))
    # the two tie rules part on a handful of pixels only
    assert np.mean(np.all(ref_tree == ref_scan, axis=-1)) > 0.98


def test_background_image_matches_oracle(gpu_ctx):
    # the camera sees past the open front of the box: those pixels take the
    # background image (rows from the bottom), bilinear at their centres
    import torch
    s = scenes.cornell_box()
    s.name = "cornell_background"
    rng = np.random.default_rng(1)
    img = rng.uniform(0.05, 1.0, (7, 9, 4)).astype(np.float32)
    dimg = torch.from_numpy(img).cuda()
    p = abi.default_render_params(48, 40, 4, 6)
    p.background_width, p.background_height = 9, 7
    from photon_amd.backend import HipBackend, PathTracer, make_camera
    cam = make_camera((0.0, 1.0, 5.5), (0.0, 1.0, 0.0), (0.0, 1.0, 0.0), 45.0, 48 / 40)
    for bottom_up in (0, 1):
        p.bottom_up = bottom_up
        pt = PathTracer(HipBackend(context=gpu_ctx))
        pt.set_scene(s)
        p.background_image = dimg.data_ptr()
        pt.camera, pt.params = cam, p
        res = pt.render()
        p.background_image = img.ctypes.data  # the oracle reads host memory
        ref, ex, _ = oracle.OracleScene(s).render(cam, p, aovs=True)
        assert rel_mse(res.color, ref) < 1e-6
        np.testing.assert_allclose(res.alpha, ex["alpha"], atol=1e-6)
        assert np.mean(ex["alpha"] < 0.999) > 0.1  # background pixels exist


def _bridge_scene_data(with_volume=True):
    """A SceneData as the ANARI device would hand it over: the Cornell
    geometry with MaterialData (perceptual roughness, specular weight), a
    sampled base colour, an emitter-only quad, point and spot LightData, an
    environment rotated by world_to_map and a small volume."""
    from photon_amd.renderer import SceneData, environment_data, sampler_data, volume_data
    base = scenes.cornell_box()
    mats = [abi.bridge_material(base_color=(0.73, 0.73, 0.73), roughness=0.9),
            abi.bridge_material(base_color=(0.65, 0.05, 0.05), roughness=0.5, specular=0.6),
            abi.bridge_material(base_color=(0.12, 0.45, 0.15), roughness=0.7, metallic=0.3,
                                base_color_src=(2, 0)),
            abi.bridge_material(base_color=(0, 0, 0), emission=(17.0, 12.0, 4.0), emitter_only=1, emission_side=0)]
    tex = scenes.colored_noise_texture(8, 5)
    smp = sampler_data(type_=3, in_attribute=abi.PHX_ATTRIBUTE_WORLD_POSITION,
                       in_transform=(0.3, 0, 0, 0, 0, 0.3, 0, 0, 0, 0, 0.3, 0, 0, 0, 0, 1),
                       in_offset=(0.3, 0.2, 0.4, 0))
    del tex
    lights = [abi.LightData(0, (0.3, 1.6, 0.4), (0, -1, 0), (1.0, 0.8, 0.6), 1.5, 0.05, 3.14159265, 0.0),
              abi.LightData(2, (-0.5, 1.9, 0.0), (0.2, -1, 0), (0.6, 0.8, 1.0), 4.0, 0.0, 0.5, 0.15)]
    env = scenes.sky_envmap(64, 32, sun_intensity=50.0)
    c, s_ = math.cos(0.7), math.sin(0.7)
    envd = environment_data(env, visible=True, world_to_map=((c, 0, s_), (0, 1, 0), (-s_, 0, c)))
    vols = []
    if with_volume:
        n = 12
        g = scenes.fbm_grid(n, 3, 4)
        vals = ((g - g.min()) / (g.max() - g.min())).astype(np.float32).reshape(-1)
        tf = np.zeros((16, 4), np.float32)
        tf[:, :3] = 0.8
        tf[:, 3] = np.clip(np.linspace(-0.5, 1.0, 16), 0, 1)
        vols.append(volume_data(vals, (n, n, n), tf, value_range=(0.0, 1.0), density_scale=3.0,
                                origin=(-0.6, 0.2, -0.6), spacing=(1.2 / (n - 1),) * 3))
    return SceneData(positions=base.positions, indices=base.indices, materials=mats, material_ids=base.material_ids,
                     samplers=[smp], lights=lights, environment=envd, volumes=vols)


@pytest.mark.parametrize("with_volume", [False, True])
def test_bridge_scene_import_matches_oracle(gpu_ctx, with_volume):
    from photon_amd import Renderer, camera_data, render_settings
    sd = _bridge_scene_data(with_volume)
    r = Renderer()
    ok, err = r.set_scene(sd)
    assert ok, err
    cam = camera_data((0.0, 1.0, 3.7), (0.0, 0.0, -1.0), fovy=np.radians(40.0), aspect=1.0)
    st = render_settings(40, 40, 6, 8)
    st.clamp_indirect = 2.0
    assert r.render(cam, st, True)[0]
    ok, err, rgba = r.read_color()
    assert ok, err
    s = osd.scene_from_data(sd)
    s.name = f"bridge_import{'_volume' if with_volume else ''}"
    ocam, p = osd.bridge_view(cam, st)
    ref, ex, _ = oracle.OracleScene(s).render(ocam, p, aovs=True)
    err = rel_mse(rgba[..., :3], ref)
    same = np.mean(np.abs(rgba[..., :3] - ref) <= 1e-4 * (1 + np.abs(ref)))
    print(f"{s.name}: relMSE {err:.3e}, pixels equal {same:.4f}", flush=True)
    assert err < 1e-3 and same > 0.9
    np.testing.assert_allclose(rgba[..., 3], np.clip(ex["alpha"], 0, 1), atol=1e-5)
    ok, err, depth = r.read_depth()
    hit = ex["depth"] < 1e29
    assert np.all(np.isinf(depth[~hit]))
    assert np.mean(np.abs(depth[hit] - ex["depth"][hit]) <= 1e-4 * ex["depth"][hit]) > 0.99


def test_bridge_empty_scene_renders_the_background(gpu_ctx):
    # build_scene turns an empty mesh into one degenerate triangle: every
    # camera ray misses and the frame is the background colour
    from photon_amd import Renderer, camera_data, render_settings
    from photon_amd.renderer import SceneData
    r = Renderer()
    ok, err = r.set_scene(SceneData())
    assert ok, err
    st = render_settings(16, 12, 2, 4)
    st.background[:] = (0.2, 0.4, 0.6, 0.5)
    assert r.render(camera_data((0, 0, 3), (0, 0, -1)), st, True)[0]
    ok, err, rgba = r.read_color()
    assert ok, err
    np.testing.assert_allclose(rgba.reshape(-1, 4), np.tile([0.1, 0.2, 0.3, 0.5], (16 * 12, 1)), rtol=1e-6)
