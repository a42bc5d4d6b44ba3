// Layout check of the C ABI against the reference's types.
//
// The namespace `ref` restates, member by member, the declarations of the
// reference's boundary types with the same member types (Vec3 = three floats,
// bool = C++ bool, enum class : u32 / i32 = 4 bytes).  Every static_assert
// compares the offset and size of one member with the phx_* counterpart, and
// the size of the whole type.  Compiling this file (g++ -std=c++17
// -fsyntax-only) is the test: a member out of order, a missing pad or a
// widened bool fails it.  Reference declarations restated here:
//   pt::Material, ParamSource     include/photon/pt/material.h:14-82
//   pt::Light                     include/photon/pt/light.h:11-35
//   pt::TfVolume                  include/photon/pt/tf_volume.h:14-33
//   pt::TextureInfo               include/photon/pt/texture.h:11-17
//   pt::SamplerInfo               include/photon/pt/sampler.h:36-52
//   pt::Camera                    include/photon/pt/camera.h:11-26
//   pt::HitResult                 include/photon/pt/backend/ray_backend.h:15-29
//   bridge::ParamSourceData, MaterialData, LightData, CameraData,
//   RenderSettings                src/photon/anari/RenderBridge.h:21-56,73-85,128-170
#include <cstddef>
#include <cstdint>

#include "../../include/phx.h"

namespace ref {
using f32 = float;
using u32 = unsigned;
using i32 = int;
struct Vec2 { f32 x, y; };
struct Vec3 { f32 x, y, z; };

enum class ParamSourceKind : i32 { Constant = 0, Attribute = 1, Sampler = 2 };
struct ParamSource { ParamSourceKind kind; i32 index; };
enum class AlphaMode : i32 { Blend = 0, Opaque = 1, Mask = 2 };
struct Material {
  Vec3 base_color;
  f32 metallic, roughness, ior, transmission, specular, clearcoat, clearcoat_roughness, sheen, sheen_tint,
      specular_tint, anisotropic;
  Vec3 emission;
  f32 emission_strength, subsurface;
  Vec3 subsurface_color;
  f32 subsurface_radius, alpha;
  i32 thin;
  i32 base_color_tex, normal_tex, roughness_tex, metallic_tex, emission_tex, alpha_tex;
  ParamSource base_color_src, alpha_src, metallic_src, roughness_src, emission_src, transmission_src, normal_src;
  AlphaMode alpha_mode;
  f32 alpha_cutoff;
  i32 flags, base_color_alpha, roughness_src_squared;
};

enum class LightType : u32 { Point = 0, Directional = 1, Spot = 2, Area = 3, Environment = 4 };
struct Light {
  LightType type;
  Vec3 position, direction, color;
  f32 intensity, spot_angle, spot_falloff;
  u32 mesh_prim_begin, mesh_prim_count;
  f32 area;
  Vec3 edge1, edge2;
  f32 radius;
};

struct TfVolume {
  f32 world_to_local[12];
  Vec3 origin, spacing;
  u32 dims[3];
  u32 data_offset, tf_offset, tf_size;
  f32 value_lo, value_hi, density_scale, majorant;
  u32 linear;
  u32 macro_offset;
  u32 macro_dims[3];
};

struct TextureInfo { u32 offset, width, height, pad; };

enum class SamplerType : u32 { Image1D = 0, Image2D = 1, Image3D = 2, Transform = 3 };
enum class WrapMode : u32 { ClampToEdge = 0, Repeat = 1, MirrorRepeat = 2 };
struct SamplerInfo {
  SamplerType type;
  i32 in_attribute;
  u32 linear;
  WrapMode wrap[3];
  u32 offset;
  u32 size[3];
  f32 in_transform[16], in_offset[4], out_transform[16], out_offset[4];
};

struct Camera {
  Vec3 origin, lower_left, horizontal, vertical, u, v, w;
  f32 lens_radius, focus_dist;
  bool is_ortho;
  f32 clip_near, clip_far;
};

struct HitResult {
  f32 t;
  Vec3 position, normal, shading_normal;
  Vec2 uv, bary;
  Vec3 interpolated_color;
  u32 prim_id, geom_id, inst_id, material_id;
  bool hit, has_interpolated_color;
};

namespace bridge {
struct ParamSourceData { int kind, index; };
struct MaterialData {
  float base_color[3];
  float opacity, metallic, roughness, ior, transmission, specular;
  float emission[3];
  float clearcoat, clearcoat_roughness, sheen;
  ParamSourceData base_color_src, opacity_src, metallic_src, roughness_src, emission_src, transmission_src, normal_src;
  int alpha_mode;
  float alpha_cutoff;
  bool base_color_alpha, thin_walled, emitter_only, hidden_to_camera;
  int emission_side;
};
struct LightData {
  int type;
  float position[3], direction[3], color[3];
  float intensity, radius, spot_angle, spot_falloff;
};
struct CameraData {
  float position[3], direction[3], up[3];
  bool orthographic;
  float fovy, height, aspect, aperture_radius, focus_distance;
  float image_region[4];
  float clip_near, clip_far;
};
struct RenderSettings {
  uint32_t width, height, pixel_samples, max_depth;
  float background[4];
  const float *background_image;
  uint32_t background_width, background_height;
  float ambient_color[3];
  float ambient_radiance;
  bool adaptive_sampling;
  float throughput_clamp;
  int max_diffuse_bounces, max_glossy_bounces, max_transmission_bounces, max_volume_bounces;
  float clamp_direct, clamp_indirect;
};
}  // namespace bridge
}  // namespace ref

#define SAME_SIZE(R, P) static_assert(sizeof(R) == sizeof(P), "size of " #P " differs from " #R)
#define SAME(R, P, f)                                                                   \
  static_assert(offsetof(R, f) == offsetof(P, f), "offset of " #P "::" #f " differs"); \
  static_assert(sizeof(((R *)0)->f) == sizeof(((P *)0)->f), "size of " #P "::" #f " differs")
#define SAME2(R, P, fr, fp)                                                               \
  static_assert(offsetof(R, fr) == offsetof(P, fp), "offset of " #P "::" #fp " differs"); \
  static_assert(sizeof(((R *)0)->fr) == sizeof(((P *)0)->fp), "size of " #P "::" #fp " differs")

SAME_SIZE(ref::ParamSource, phx_param_source);
SAME(ref::ParamSource, phx_param_source, kind);
SAME(ref::ParamSource, phx_param_source, index);

SAME_SIZE(ref::Material, phx_material);
SAME(ref::Material, phx_material, base_color);
SAME(ref::Material, phx_material, metallic);
SAME(ref::Material, phx_material, roughness);
SAME(ref::Material, phx_material, ior);
SAME(ref::Material, phx_material, transmission);
SAME(ref::Material, phx_material, specular);
SAME(ref::Material, phx_material, clearcoat);
SAME(ref::Material, phx_material, clearcoat_roughness);
SAME(ref::Material, phx_material, sheen);
SAME(ref::Material, phx_material, sheen_tint);
SAME(ref::Material, phx_material, specular_tint);
SAME(ref::Material, phx_material, anisotropic);
SAME(ref::Material, phx_material, emission);
SAME(ref::Material, phx_material, emission_strength);
SAME(ref::Material, phx_material, subsurface);
SAME(ref::Material, phx_material, subsurface_color);
SAME(ref::Material, phx_material, subsurface_radius);
SAME(ref::Material, phx_material, alpha);
SAME(ref::Material, phx_material, thin);
SAME(ref::Material, phx_material, base_color_tex);
SAME(ref::Material, phx_material, normal_tex);
SAME(ref::Material, phx_material, roughness_tex);
SAME(ref::Material, phx_material, metallic_tex);
SAME(ref::Material, phx_material, emission_tex);
SAME(ref::Material, phx_material, alpha_tex);
SAME(ref::Material, phx_material, base_color_src);
SAME(ref::Material, phx_material, alpha_src);
SAME(ref::Material, phx_material, metallic_src);
SAME(ref::Material, phx_material, roughness_src);
SAME(ref::Material, phx_material, emission_src);
SAME(ref::Material, phx_material, transmission_src);
SAME(ref::Material, phx_material, normal_src);
SAME(ref::Material, phx_material, alpha_mode);
SAME(ref::Material, phx_material, alpha_cutoff);
SAME(ref::Material, phx_material, flags);
SAME(ref::Material, phx_material, base_color_alpha);
SAME(ref::Material, phx_material, roughness_src_squared);

SAME_SIZE(ref::Light, phx_light);
SAME(ref::Light, phx_light, type);
SAME(ref::Light, phx_light, position);
SAME(ref::Light, phx_light, direction);
SAME(ref::Light, phx_light, color);
SAME(ref::Light, phx_light, intensity);
SAME(ref::Light, phx_light, spot_angle);
SAME(ref::Light, phx_light, spot_falloff);
SAME(ref::Light, phx_light, mesh_prim_begin);
SAME(ref::Light, phx_light, mesh_prim_count);
SAME(ref::Light, phx_light, area);
SAME(ref::Light, phx_light, edge1);
SAME(ref::Light, phx_light, edge2);
SAME(ref::Light, phx_light, radius);

SAME_SIZE(ref::TfVolume, phx_volume);
SAME(ref::TfVolume, phx_volume, world_to_local);
SAME(ref::TfVolume, phx_volume, origin);
SAME(ref::TfVolume, phx_volume, spacing);
SAME(ref::TfVolume, phx_volume, dims);
SAME(ref::TfVolume, phx_volume, data_offset);
SAME(ref::TfVolume, phx_volume, tf_offset);
SAME(ref::TfVolume, phx_volume, tf_size);
SAME(ref::TfVolume, phx_volume, value_lo);
SAME(ref::TfVolume, phx_volume, value_hi);
SAME(ref::TfVolume, phx_volume, density_scale);
SAME(ref::TfVolume, phx_volume, majorant);
SAME(ref::TfVolume, phx_volume, linear);
SAME(ref::TfVolume, phx_volume, macro_offset);
SAME(ref::TfVolume, phx_volume, macro_dims);

SAME_SIZE(ref::TextureInfo, phx_texture_info);
SAME(ref::TextureInfo, phx_texture_info, offset);
SAME(ref::TextureInfo, phx_texture_info, width);
SAME(ref::TextureInfo, phx_texture_info, height);
SAME2(ref::TextureInfo, phx_texture_info, pad, pad_);

SAME_SIZE(ref::SamplerInfo, phx_sampler_info);
SAME(ref::SamplerInfo, phx_sampler_info, type);
SAME(ref::SamplerInfo, phx_sampler_info, in_attribute);
SAME(ref::SamplerInfo, phx_sampler_info, linear);
SAME(ref::SamplerInfo, phx_sampler_info, wrap);
SAME(ref::SamplerInfo, phx_sampler_info, offset);
SAME(ref::SamplerInfo, phx_sampler_info, size);
SAME(ref::SamplerInfo, phx_sampler_info, in_transform);
SAME(ref::SamplerInfo, phx_sampler_info, in_offset);
SAME(ref::SamplerInfo, phx_sampler_info, out_transform);
SAME(ref::SamplerInfo, phx_sampler_info, out_offset);

SAME_SIZE(ref::Camera, phx_camera);
SAME(ref::Camera, phx_camera, origin);
SAME(ref::Camera, phx_camera, lower_left);
SAME(ref::Camera, phx_camera, horizontal);
SAME(ref::Camera, phx_camera, vertical);
SAME(ref::Camera, phx_camera, u);
SAME(ref::Camera, phx_camera, v);
SAME(ref::Camera, phx_camera, w);
SAME(ref::Camera, phx_camera, lens_radius);
SAME(ref::Camera, phx_camera, focus_dist);
SAME(ref::Camera, phx_camera, is_ortho);
SAME(ref::Camera, phx_camera, clip_near);
SAME(ref::Camera, phx_camera, clip_far);

SAME_SIZE(ref::HitResult, phx_hit);
SAME(ref::HitResult, phx_hit, t);
SAME(ref::HitResult, phx_hit, position);
SAME(ref::HitResult, phx_hit, normal);
SAME(ref::HitResult, phx_hit, shading_normal);
SAME(ref::HitResult, phx_hit, uv);
SAME(ref::HitResult, phx_hit, bary);
SAME(ref::HitResult, phx_hit, interpolated_color);
SAME(ref::HitResult, phx_hit, prim_id);
SAME(ref::HitResult, phx_hit, geom_id);
SAME(ref::HitResult, phx_hit, inst_id);
SAME(ref::HitResult, phx_hit, material_id);
SAME(ref::HitResult, phx_hit, hit);
SAME(ref::HitResult, phx_hit, has_interpolated_color);

using ref::bridge::MaterialData;
SAME_SIZE(ref::bridge::ParamSourceData, phx_param_source);
SAME_SIZE(MaterialData, phx_bridge_material);
SAME(MaterialData, phx_bridge_material, base_color);
SAME(MaterialData, phx_bridge_material, opacity);
SAME(MaterialData, phx_bridge_material, metallic);
SAME(MaterialData, phx_bridge_material, roughness);
SAME(MaterialData, phx_bridge_material, ior);
SAME(MaterialData, phx_bridge_material, transmission);
SAME(MaterialData, phx_bridge_material, specular);
SAME(MaterialData, phx_bridge_material, emission);
SAME(MaterialData, phx_bridge_material, clearcoat);
SAME(MaterialData, phx_bridge_material, clearcoat_roughness);
SAME(MaterialData, phx_bridge_material, sheen);
SAME(MaterialData, phx_bridge_material, base_color_src);
SAME(MaterialData, phx_bridge_material, opacity_src);
SAME(MaterialData, phx_bridge_material, metallic_src);
SAME(MaterialData, phx_bridge_material, roughness_src);
SAME(MaterialData, phx_bridge_material, emission_src);
SAME(MaterialData, phx_bridge_material, transmission_src);
SAME(MaterialData, phx_bridge_material, normal_src);
SAME(MaterialData, phx_bridge_material, alpha_mode);
SAME(MaterialData, phx_bridge_material, alpha_cutoff);
SAME(MaterialData, phx_bridge_material, base_color_alpha);
SAME(MaterialData, phx_bridge_material, thin_walled);
SAME(MaterialData, phx_bridge_material, emitter_only);
SAME(MaterialData, phx_bridge_material, hidden_to_camera);
SAME(MaterialData, phx_bridge_material, emission_side);

using ref::bridge::LightData;
SAME_SIZE(LightData, phx_light_data);
SAME(LightData, phx_light_data, type);
SAME(LightData, phx_light_data, position);
SAME(LightData, phx_light_data, direction);
SAME(LightData, phx_light_data, color);
SAME(LightData, phx_light_data, intensity);
SAME(LightData, phx_light_data, radius);
SAME(LightData, phx_light_data, spot_angle);
SAME(LightData, phx_light_data, spot_falloff);

using ref::bridge::CameraData;
SAME_SIZE(CameraData, phx_camera_data);
SAME(CameraData, phx_camera_data, position);
SAME(CameraData, phx_camera_data, direction);
SAME(CameraData, phx_camera_data, up);
SAME(CameraData, phx_camera_data, orthographic);
SAME(CameraData, phx_camera_data, fovy);
SAME(CameraData, phx_camera_data, height);
SAME(CameraData, phx_camera_data, aspect);
SAME(CameraData, phx_camera_data, aperture_radius);
SAME(CameraData, phx_camera_data, focus_distance);
SAME(CameraData, phx_camera_data, image_region);
SAME(CameraData, phx_camera_data, clip_near);
SAME(CameraData, phx_camera_data, clip_far);

using ref::bridge::RenderSettings;
SAME_SIZE(RenderSettings, phx_render_settings);
SAME(RenderSettings, phx_render_settings, width);
SAME(RenderSettings, phx_render_settings, height);
SAME(RenderSettings, phx_render_settings, pixel_samples);
SAME(RenderSettings, phx_render_settings, max_depth);
SAME(RenderSettings, phx_render_settings, background);
SAME(RenderSettings, phx_render_settings, background_image);
SAME(RenderSettings, phx_render_settings, background_width);
SAME(RenderSettings, phx_render_settings, background_height);
SAME(RenderSettings, phx_render_settings, ambient_color);
SAME(RenderSettings, phx_render_settings, ambient_radiance);
SAME(RenderSettings, phx_render_settings, adaptive_sampling);
SAME(RenderSettings, phx_render_settings, throughput_clamp);
SAME(RenderSettings, phx_render_settings, max_diffuse_bounces);
SAME(RenderSettings, phx_render_settings, max_glossy_bounces);
SAME(RenderSettings, phx_render_settings, max_transmission_bounces);
SAME(RenderSettings, phx_render_settings, max_volume_bounces);
SAME(RenderSettings, phx_render_settings, clamp_direct);
SAME(RenderSettings, phx_render_settings, clamp_indirect);

// The sizes the ABI documents.
static_assert(sizeof(phx_material) == 200, "phx_material is 200 bytes");
static_assert(sizeof(phx_hit) == 88, "phx_hit is 88 bytes");
static_assert(sizeof(phx_light) == 92, "phx_light is 92 bytes");
static_assert(sizeof(phx_volume) == 132, "phx_volume is 132 bytes");
static_assert(sizeof(phx_sampler_info) == 200, "phx_sampler_info is 200 bytes");
static_assert(sizeof(phx_camera) == 104, "phx_camera is 104 bytes");
static_assert(sizeof(phx_bridge_material) == 132, "phx_bridge_material is 132 bytes");
static_assert(sizeof(phx_light_data) == 56, "phx_light_data is 56 bytes");
static_assert(sizeof(phx_camera_data) == 84, "phx_camera_data is 84 bytes");
static_assert(sizeof(phx_render_settings) == 96, "phx_render_settings is 96 bytes");

// Reading a bool member of a reference object through the uint8_t member:
// the byte holds 0 or 1 (the x86-64 psABI representation of bool).
static_assert(sizeof(bool) == 1, "bool is one byte");
