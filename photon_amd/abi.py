"""ctypes mirrors of the C ABI structs declared in include/phx.h.

The field order and sizes are those of the header; ``check_layout`` asserts the
sizes against the values the header fixes, so a drift fails loudly.
"""
import ctypes as C

F3 = C.c_float * 3


class ParamSource(C.Structure):
    """pt::ParamSource (material.h:16-19) / bridge::ParamSourceData."""
    _fields_ = [("kind", C.c_int32), ("index", C.c_int32)]


_SOURCES = ("base_color_src", "alpha_src", "metallic_src", "roughness_src", "emission_src", "transmission_src",
            "normal_src")


class Material(C.Structure):
    """pt::Material (reference include/photon/pt/material.h:35-82), 200 bytes."""
    _fields_ = [
        ("base_color", F3), ("metallic", C.c_float),
        ("roughness", C.c_float), ("ior", C.c_float), ("transmission", C.c_float), ("specular", C.c_float),
        ("clearcoat", C.c_float), ("clearcoat_roughness", C.c_float), ("sheen", C.c_float), ("sheen_tint", C.c_float),
        ("specular_tint", C.c_float), ("anisotropic", C.c_float),
        ("emission", F3), ("emission_strength", C.c_float), ("subsurface", C.c_float),
        ("subsurface_color", F3), ("subsurface_radius", C.c_float), ("alpha", C.c_float), ("thin", C.c_int32),
        ("base_color_tex", C.c_int32), ("normal_tex", C.c_int32), ("roughness_tex", C.c_int32),
        ("metallic_tex", C.c_int32), ("emission_tex", C.c_int32), ("alpha_tex", C.c_int32),
    ] + [(n, ParamSource) for n in _SOURCES] + [
        ("alpha_mode", C.c_int32), ("alpha_cutoff", C.c_float), ("flags", C.c_int32),
        ("base_color_alpha", C.c_int32), ("roughness_src_squared", C.c_int32),
    ]


class Light(C.Structure):
    """pt::Light (light.h:19-35), 92 bytes."""
    _fields_ = [
        ("type", C.c_uint32), ("position", F3), ("direction", F3), ("color", F3), ("intensity", C.c_float),
        ("spot_angle", C.c_float), ("spot_falloff", C.c_float), ("mesh_prim_begin", C.c_uint32),
        ("mesh_prim_count", C.c_uint32), ("area", C.c_float), ("edge1", F3), ("edge2", F3), ("radius", C.c_float),
    ]


class SamplerInfo(C.Structure):
    """pt::SamplerInfo (sampler.h:39-52), 200 bytes."""
    _fields_ = [
        ("type", C.c_uint32), ("in_attribute", C.c_int32), ("linear", C.c_uint32), ("wrap", C.c_uint32 * 3),
        ("offset", C.c_uint32), ("size", C.c_uint32 * 3), ("in_transform", C.c_float * 16),
        ("in_offset", C.c_float * 4), ("out_transform", C.c_float * 16), ("out_offset", C.c_float * 4),
    ]


class TextureInfo(C.Structure):
    _fields_ = [("offset", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32), ("pad_", C.c_uint32)]


class Volume(C.Structure):
    """pt::TfVolume (tf_volume.h:14-33), 132 bytes."""
    _fields_ = [
        ("world_to_local", C.c_float * 12), ("origin", F3), ("spacing", F3), ("dims", C.c_uint32 * 3),
        ("data_offset", C.c_uint32), ("tf_offset", C.c_uint32), ("tf_size", C.c_uint32),
        ("value_lo", C.c_float), ("value_hi", C.c_float), ("density_scale", C.c_float), ("majorant", C.c_float),
        ("linear", C.c_uint32), ("macro_offset", C.c_uint32), ("macro_dims", C.c_uint32 * 3),
    ]


class SceneDesc(C.Structure):
    _fields_ = [
        ("positions", C.POINTER(C.c_float)), ("vertex_count", C.c_uint32),
        ("indices", C.POINTER(C.c_uint32)), ("triangle_count", C.c_uint32),
        ("normals", C.POINTER(C.c_float)), ("texcoords", C.POINTER(C.c_float)),
        ("vertex_colors", C.POINTER(C.c_float)), ("material_ids", C.POINTER(C.c_uint32)),
        ("materials", C.POINTER(Material)), ("material_count", C.c_uint32),
        ("lights", C.POINTER(Light)), ("light_count", C.c_uint32),
        ("emissive_prim_ids", C.POINTER(C.c_uint32)), ("emissive_prim_areas", C.POINTER(C.c_float)),
        ("emissive_count", C.c_uint32),
        ("texture_pixels", C.POINTER(C.c_float)), ("texture_pixel_count", C.c_uint32),
        ("texture_infos", C.POINTER(TextureInfo)), ("texture_count", C.c_uint32),
        ("attributes", C.POINTER(C.c_float)), ("attribute_count", C.c_uint32), ("attribute_offset", C.c_int32 * 5),
        ("samplers", C.POINTER(SamplerInfo)), ("sampler_count", C.c_uint32),
        ("sampler_texels", C.POINTER(C.c_float)), ("sampler_texel_count", C.c_uint32),
        ("env_width", C.c_uint32), ("env_height", C.c_uint32), ("env_pixels", C.POINTER(C.c_float)),
        ("env_world_to_tex", C.c_float * 9), ("env_visible", C.c_int32),
        ("volumes", C.POINTER(Volume)), ("volume_count", C.c_uint32),
        ("volume_values", C.POINTER(C.c_float)), ("volume_value_count", C.c_uint64),
        ("volume_tf", C.POINTER(C.c_float)), ("volume_tf_count", C.c_uint32),
        ("volume_majorants", C.POINTER(C.c_float)), ("volume_majorant_count", C.c_uint32),
        ("ray_epsilon", C.c_float),
    ]


class Camera(C.Structure):
    """pt::Camera (camera.h:11-26), derived form."""
    _fields_ = [
        ("origin", F3), ("lower_left", F3), ("horizontal", F3), ("vertical", F3),
        ("u", F3), ("v", F3), ("w", F3),
        ("lens_radius", C.c_float), ("focus_dist", C.c_float), ("is_ortho", C.c_uint8), ("pad_", C.c_uint8 * 3),
        ("clip_near", C.c_float), ("clip_far", C.c_float),
    ]


class RenderParams(C.Structure):
    """pt::RenderParams (pathtracer.h:16-55); the background image is a device
    pointer to RGBA rows from the bottom of the frame."""
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32), ("samples_per_pixel", C.c_uint32),
        ("max_depth", C.c_uint32), ("sample_offset", C.c_uint32),
        ("background_color", F3), ("background_alpha", C.c_float),
        ("ambient_color", F3), ("ambient_radiance", C.c_float),
        ("adaptive_sampling", C.c_int32), ("throughput_clamp", C.c_float),
        ("max_diffuse_bounces", C.c_int32), ("max_glossy_bounces", C.c_int32),
        ("max_transmission_bounces", C.c_int32), ("max_volume_bounces", C.c_int32),
        ("clamp_direct_sum", C.c_float), ("clamp_indirect_sum", C.c_float),
        ("bottom_up", C.c_int32), ("image_region", C.c_float * 4),
        ("background_image", C.c_void_p), ("background_width", C.c_uint32), ("background_height", C.c_uint32),
    ]


class Hit(C.Structure):
    """pt::HitResult (ray_backend.h:15-29)."""
    _fields_ = [
        ("t", C.c_float), ("position", F3), ("normal", F3), ("shading_normal", F3),
        ("uv", C.c_float * 2), ("bary", C.c_float * 2), ("interpolated_color", F3),
        ("prim_id", C.c_uint32), ("geom_id", C.c_uint32), ("inst_id", C.c_uint32), ("material_id", C.c_uint32),
        ("hit", C.c_uint8), ("has_interpolated_color", C.c_uint8), ("pad_", C.c_uint8 * 2),
    ]


class HitCompact(C.Structure):
    _fields_ = [("t", C.c_float), ("prim_id", C.c_uint32), ("b1", C.c_float), ("b2", C.c_float)]


class Tuning(C.Structure):
    _fields_ = [
        ("trace_block", C.c_uint32), ("trace_groups_per_cu", C.c_uint32), ("refill_lanes", C.c_uint32),
        ("queue_chunk", C.c_uint32), ("paths_in_flight", C.c_uint32), ("sort_by_material", C.c_uint32),
        ("use_graph", C.c_uint32), ("time_kernels", C.c_uint32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("closest_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("shade_events", C.c_uint64),
        ("samples", C.c_uint64), ("pixels", C.c_uint64), ("scene_bytes", C.c_uint64),
        ("bvh_nodes", C.c_uint64), ("bvh_bytes", C.c_uint64), ("last_render_ms", C.c_double),
        ("extend_launches", C.c_uint64), ("extend_ms", C.c_double), ("shade_launches", C.c_uint64),
        ("shade_ms", C.c_double), ("shadow_launches", C.c_uint64), ("shadow_ms", C.c_double),
    ]


class BridgeMaterial(C.Structure):
    """bridge::MaterialData (RenderBridge.h:21-56), 132 bytes."""
    _fields_ = [
        ("base_color", F3), ("opacity", C.c_float), ("metallic", C.c_float), ("roughness", C.c_float),
        ("ior", C.c_float), ("transmission", C.c_float), ("specular", C.c_float), ("emission", F3),
        ("clearcoat", C.c_float), ("clearcoat_roughness", C.c_float), ("sheen", C.c_float),
        ("base_color_src", ParamSource), ("opacity_src", ParamSource), ("metallic_src", ParamSource),
        ("roughness_src", ParamSource), ("emission_src", ParamSource), ("transmission_src", ParamSource),
        ("normal_src", ParamSource), ("alpha_mode", C.c_int32), ("alpha_cutoff", C.c_float),
        ("base_color_alpha", C.c_uint8), ("thin_walled", C.c_uint8), ("emitter_only", C.c_uint8),
        ("hidden_to_camera", C.c_uint8), ("emission_side", C.c_int32),
    ]


def bridge_material(**fields):
    """A MaterialData with the bridge's defaults (RenderBridge.h:27-56), then
    `fields`; a source is given as (kind, index) or a kind alone."""
    m = BridgeMaterial()
    m.base_color[:] = (0.8, 0.8, 0.8)
    m.opacity, m.metallic, m.roughness, m.ior = 1.0, 0.0, 1.0, 1.5
    m.transmission, m.specular = 0.0, 0.0
    m.clearcoat, m.clearcoat_roughness, m.sheen = 0.0, 0.03, 0.0
    m.alpha_mode, m.alpha_cutoff = 1, 0.5
    m.emission_side = 2
    for k, v in fields.items():
        if k in ("base_color", "emission"):
            getattr(m, k)[:] = v
        elif k.endswith("_src"):
            kind, index = (v, 0) if isinstance(v, int) else v
            getattr(m, k).kind, getattr(m, k).index = kind, index
        else:
            setattr(m, k, v)
    return m


class CameraData(C.Structure):
    """bridge::CameraData (RenderBridge.h:128-144), 84 bytes."""
    _fields_ = [
        ("position", F3), ("direction", F3), ("up", F3), ("orthographic", C.c_uint8), ("pad_", C.c_uint8 * 3),
        ("fovy", C.c_float), ("height", C.c_float), ("aspect", C.c_float),
        ("aperture_radius", C.c_float), ("focus_distance", C.c_float), ("image_region", C.c_float * 4),
        ("clip_near", C.c_float), ("clip_far", C.c_float),
    ]


class RenderSettings(C.Structure):
    """bridge::RenderSettings (RenderBridge.h:146-170), 96 bytes; the background
    image is host memory."""
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32), ("pixel_samples", C.c_uint32), ("max_depth", C.c_uint32),
        ("background", C.c_float * 4), ("background_image", C.c_void_p), ("background_width", C.c_uint32),
        ("background_height", C.c_uint32), ("ambient_color", F3), ("ambient_radiance", C.c_float),
        ("adaptive_sampling", C.c_uint8), ("pad_", C.c_uint8 * 3), ("throughput_clamp", C.c_float),
        ("max_diffuse_bounces", C.c_int32), ("max_glossy_bounces", C.c_int32),
        ("max_transmission_bounces", C.c_int32), ("max_volume_bounces", C.c_int32),
        ("clamp_direct", C.c_float), ("clamp_indirect", C.c_float),
    ]


class LightData(C.Structure):
    """bridge::LightData (RenderBridge.h:73-85), 56 bytes."""
    _fields_ = [
        ("type", C.c_int32), ("position", F3), ("direction", F3), ("color", F3), ("intensity", C.c_float),
        ("radius", C.c_float), ("spot_angle", C.c_float), ("spot_falloff", C.c_float),
    ]


class SamplerData(C.Structure):
    """bridge::SamplerData (RenderBridge.h:58-71), the texel vector as pointer + float count."""
    _fields_ = [
        ("type", C.c_int32), ("in_attribute", C.c_int32), ("linear", C.c_uint8), ("pad_", C.c_uint8 * 3),
        ("wrap", C.c_int32 * 3), ("size", C.c_uint32 * 3), ("in_transform", C.c_float * 16),
        ("in_offset", C.c_float * 4), ("out_transform", C.c_float * 16), ("out_offset", C.c_float * 4),
        ("texels", C.c_void_p), ("texel_float_count", C.c_uint64),
    ]


class EnvironmentData(C.Structure):
    """bridge::EnvironmentData (RenderBridge.h:87-97)."""
    _fields_ = [
        ("present", C.c_uint8), ("visible", C.c_uint8), ("pad_", C.c_uint8 * 2), ("width", C.c_uint32),
        ("height", C.c_uint32), ("pixels", C.c_void_p), ("pixel_float_count", C.c_uint64),
        ("world_to_map", C.c_float * 9),
    ]


class VolumeData(C.Structure):
    """bridge::VolumeData (RenderBridge.h:99-110)."""
    _fields_ = [
        ("world_to_local", C.c_float * 12), ("origin", F3), ("spacing", F3), ("dims", C.c_uint32 * 3),
        ("values", C.c_void_p), ("value_count", C.c_uint64), ("transfer", C.c_void_p),
        ("transfer_float_count", C.c_uint64), ("value_range", C.c_float * 2), ("density_scale", C.c_float),
        ("linear", C.c_uint8), ("pad_", C.c_uint8 * 3),
    ]


class SceneData(C.Structure):
    """bridge::SceneData (RenderBridge.h:112-125), vectors as pointer + count."""
    _fields_ = [
        ("positions", C.c_void_p), ("position_count", C.c_uint64), ("normals", C.c_void_p),
        ("normal_count", C.c_uint64), ("attributes", C.c_void_p * 5), ("attribute_count", C.c_uint64 * 5),
        ("indices", C.c_void_p), ("index_count", C.c_uint64), ("material_ids", C.c_void_p),
        ("material_id_count", C.c_uint64), ("materials", C.POINTER(BridgeMaterial)), ("material_count", C.c_uint32),
        ("samplers", C.POINTER(SamplerData)), ("sampler_count", C.c_uint32),
        ("lights", C.POINTER(LightData)), ("light_count", C.c_uint32), ("environment", EnvironmentData),
        ("volumes", C.POINTER(VolumeData)), ("volume_count", C.c_uint32),
    ]


class View(C.Structure):
    """phx_view: one view of phx_render_batch."""
    _fields_ = [("scene", C.c_void_p), ("camera", Camera), ("params", RenderParams)]


class Display(C.Structure):
    """phx_display: exposure, tone map and transfer curve of phx_display_transform."""
    _fields_ = [("tonemap", C.c_int32), ("curve", C.c_int32), ("exposure", C.c_float), ("clamp", C.c_uint8),
                ("pad_", C.c_uint8 * 3)]


class PbrtInfo(C.Structure):
    """phx_pbrt_info: the pbrt file's Film/Sampler/Integrator settings and counts."""
    _fields_ = [(n, C.c_uint32) for n in ("width", "height", "spp", "max_depth", "meshes", "objects", "instances",
                                           "textures")]


PHX_TONEMAP_NONE, PHX_TONEMAP_ACES, PHX_TONEMAP_REINHARD = 0, 1, 2
PHX_CURVE_LINEAR, PHX_CURVE_SRGB, PHX_CURVE_GAMMA22 = 0, 1, 2


def display(tonemap=PHX_TONEMAP_ACES, curve=PHX_CURVE_SRGB, exposure=0.0, clamp=1):
    """A Display; the defaults are photon_render's postprocess (ACES, clamp,
    sRGB curve, clamp: photon_render.cpp:29-35)."""
    d = Display()
    d.tonemap, d.curve, d.exposure, d.clamp = tonemap, curve, exposure, clamp
    return d


EXPECTED_SIZES = {Material: 200, Light: 92, TextureInfo: 16, Volume: 132, SamplerInfo: 200, Hit: 88, HitCompact: 16,
                  Camera: 104, Tuning: 32, BridgeMaterial: 132, CameraData: 84, RenderSettings: 96, LightData: 56,
                  Display: 16, PbrtInfo: 32}


def check_layout():
    for cls, size in EXPECTED_SIZES.items():
        got = C.sizeof(cls)
        if got != size:
            raise AssertionError(f"{cls.__name__}: ctypes size {got} != {size} fixed by include/phx.h")


PHX_MATERIAL_FLAG_EMITTER_ONLY = 1
PHX_MATERIAL_FLAG_HIDDEN_TO_CAMERA = 2
PHX_MATERIAL_FLAG_EMIT_BOTH_SIDES = 4
PHX_MATERIAL_FLAG_REFRACTIVE = 8
PHX_MATERIAL_FLAG_EMIT_BACK_ONLY = 16
PHX_ALPHA_BLEND, PHX_ALPHA_OPAQUE, PHX_ALPHA_MASK = 0, 1, 2
PHX_LIGHT_POINT, PHX_LIGHT_DIRECTIONAL, PHX_LIGHT_SPOT, PHX_LIGHT_AREA = 0, 1, 2, 3
PHX_SOURCE_CONSTANT, PHX_SOURCE_ATTRIBUTE, PHX_SOURCE_SAMPLER = 0, 1, 2
PHX_ATTRIBUTE_COLOR, PHX_ATTRIBUTE_WORLD_POSITION, PHX_ATTRIBUTE_WORLD_NORMAL = 4, 5, 6
PHX_SAMPLER_IMAGE1D, PHX_SAMPLER_IMAGE2D, PHX_SAMPLER_IMAGE3D, PHX_SAMPLER_TRANSFORM = 0, 1, 2, 3
PHX_WRAP_CLAMP, PHX_WRAP_REPEAT, PHX_WRAP_MIRROR = 0, 1, 2

_IDENTITY4 = (1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0)


def sampler_info(type_=PHX_SAMPLER_IMAGE2D, in_attribute=0, linear=1, wrap=(0, 0, 0), offset=0, size=(1, 1, 1),
                 in_transform=_IDENTITY4, in_offset=(0, 0, 0, 0), out_transform=_IDENTITY4, out_offset=(0, 0, 0, 0)):
    """A SamplerInfo with the reference's defaults (sampler.h:39-52)."""
    si = SamplerInfo()
    si.type, si.in_attribute, si.linear, si.offset = type_, in_attribute, linear, offset
    si.wrap[:] = wrap
    si.size[:] = size
    si.in_transform[:] = in_transform
    si.in_offset[:] = in_offset
    si.out_transform[:] = out_transform
    si.out_offset[:] = out_offset
    return si


def default_material():
    """pt::Material{} defaults (material.h:35-82)."""
    m = Material()
    m.base_color[:] = (0.8, 0.8, 0.8)
    m.metallic = 0.0
    m.roughness = 0.5
    m.ior = 1.5
    m.transmission = 0.0
    m.specular = 0.5
    m.clearcoat = 0.0
    m.clearcoat_roughness = 0.03
    m.sheen = 0.0
    m.sheen_tint = 0.5
    m.specular_tint = 0.0
    m.anisotropic = 0.0
    m.emission[:] = (0.0, 0.0, 0.0)
    m.emission_strength = 0.0
    m.subsurface = 0.0
    m.subsurface_color[:] = (1.0, 1.0, 1.0)
    m.subsurface_radius = 1.0
    m.alpha = 1.0
    m.thin = 0
    m.flags = 0
    m.alpha_mode = PHX_ALPHA_BLEND
    m.alpha_cutoff = 0.5
    for k in ("base_color_tex", "normal_tex", "roughness_tex", "metallic_tex", "emission_tex", "alpha_tex"):
        setattr(m, k, -1)
    for k in _SOURCES:
        getattr(m, k).kind, getattr(m, k).index = PHX_SOURCE_CONSTANT, 0
    m.base_color_alpha = 0
    m.roughness_src_squared = 0
    return m


def default_render_params(width, height, spp, max_depth):
    """RenderParams with the settings every parity run fixes: adaptive sampling
    off, no throughput clamp, no light clamps, unlimited bounce kinds."""
    p = RenderParams()
    p.width, p.height, p.samples_per_pixel, p.max_depth = width, height, spp, max_depth
    p.sample_offset = 0
    p.background_color[:] = (0.0, 0.0, 0.0)
    p.background_alpha = 1.0
    p.ambient_color[:] = (1.0, 1.0, 1.0)
    p.ambient_radiance = 0.0
    p.adaptive_sampling = 0
    p.throughput_clamp = 0.0
    p.max_diffuse_bounces = p.max_glossy_bounces = p.max_transmission_bounces = p.max_volume_bounces = -1
    p.clamp_direct_sum = p.clamp_indirect_sum = 0.0
    p.bottom_up = 0
    p.image_region[:] = (0.0, 0.0, 1.0, 1.0)
    return p
