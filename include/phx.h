/*
 * phx.h - C ABI of the MI355X (gfx950) wavefront path-tracing backend.
 *
 * Plain C types, plain pointers and sizes.  Every entry point returns PHX_OK (0)
 * or a negative PHX_ERR_* code; phx_last_error() gives the message of the last
 * failure on a context.  One context serves one host thread at a time (the
 * reference's backends are driven from a single thread, RenderBridge.cpp:66-77);
 * the library starts no host threads.  All kernels of a context run on the
 * context's HIP stream, or on the stream passed to a call that takes one; the
 * bounce loop of a render also runs its any-hit pass on a second stream of the
 * context, joined to the first before each pass that adds its results.  A
 * render of several runs (a batch, or a view larger than one run) streams
 * every other run on a second pair of streams of the context, created on
 * first use; its work is joined to the context's stream before the call
 * returns or resolves.
 *
 * Data types mirror the reference's scene model.  These have the reference
 * type's exact size, field order and field offsets (x86-64 and gfx950 agree),
 * so an array of the reference type can be handed over as is:
 *   phx_material   == pt::Material       (include/photon/pt/material.h:16-19,35-82)
 *   phx_light      == pt::Light          (include/photon/pt/light.h:19-35)
 *   phx_volume     == pt::TfVolume       (include/photon/pt/tf_volume.h:14-33)
 *   phx_texture_info == pt::TextureInfo  (include/photon/pt/texture.h:11-17)
 *   phx_sampler_info == pt::SamplerInfo  (include/photon/pt/sampler.h:39-52)
 *   phx_camera     == pt::Camera         (include/photon/pt/camera.h:11-26)
 *   phx_hit        == pt::HitResult      (include/photon/pt/backend/ray_backend.h:15-29)
 *   phx_bridge_material == bridge::MaterialData (src/photon/anari/RenderBridge.h:21-56)
 *   phx_light_data == bridge::LightData   (RenderBridge.h:73-85)
 *   phx_camera_data == bridge::CameraData (RenderBridge.h:128-144)
 *   phx_render_settings == bridge::RenderSettings (RenderBridge.h:146-170)
 * The C++ bool members of the reference types are uint8_t here.
 * phx_render_params follows pt::RenderParams (pathtracer.h:16-55) field by
 * field, with the background image as a device pointer and its size instead
 * of a Kokkos::View; the bridge's std::vector-holding types (SceneData,
 * SamplerData, EnvironmentData, VolumeData) become pointer + count structs.
 */
#ifndef PHX_H
#define PHX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHX_ABI_VERSION 3

enum {
  PHX_OK = 0,
  PHX_ERR_INVALID = -1,   /* bad argument or size */
  PHX_ERR_HIP = -2,       /* a HIP runtime call failed */
  PHX_ERR_NO_SCENE = -3,  /* render/trace before a scene (or accel) was set */
  PHX_ERR_OVERFLOW = -4,  /* a fixed-size device resource is too small */
  PHX_ERR_NO_FRAME = -5,  /* read before any frame was rendered */
  PHX_ERR_ALLOC = -6,     /* device or host allocation failed */
  PHX_ERR_UNSUPPORTED = -7
};

/* Material::flags bits (material.h:29-33). */
#define PHX_MATERIAL_FLAG_EMITTER_ONLY 1
#define PHX_MATERIAL_FLAG_HIDDEN_TO_CAMERA 2
#define PHX_MATERIAL_FLAG_EMIT_BOTH_SIDES 4
#define PHX_MATERIAL_FLAG_REFRACTIVE 8
#define PHX_MATERIAL_FLAG_EMIT_BACK_ONLY 16

/* AlphaMode (material.h:22-26). */
#define PHX_ALPHA_BLEND 0
#define PHX_ALPHA_OPAQUE 1
#define PHX_ALPHA_MASK 2

/* LightType (light.h:11-17). */
#define PHX_LIGHT_POINT 0
#define PHX_LIGHT_DIRECTIONAL 1
#define PHX_LIGHT_SPOT 2
#define PHX_LIGHT_AREA 3

/* ParamSourceKind (material.h:14). */
#define PHX_SOURCE_CONSTANT 0
#define PHX_SOURCE_ATTRIBUTE 1
#define PHX_SOURCE_SAMPLER 2

/* Attribute slots (sampler.h:11-19): 0-3 and the colour are per-vertex
 * attributes of the mesh; the world position and normal of the surface point
 * are computed. */
#define PHX_ATTRIBUTE_0 0
#define PHX_ATTRIBUTE_COLOR 4
#define PHX_ATTRIBUTE_MESH_COUNT 5
#define PHX_ATTRIBUTE_WORLD_POSITION 5
#define PHX_ATTRIBUTE_WORLD_NORMAL 6

/* pt::ParamSource (material.h:16-19): where a material parameter comes from. */
typedef struct phx_param_source {
  int32_t kind;  /* PHX_SOURCE_* */
  int32_t index; /* attribute slot or sampler id */
} phx_param_source;

/* pt::Material (material.h:35-82). 200 bytes. */
typedef struct phx_material {
  float base_color[3];
  float metallic;
  float roughness;      /* GGX alpha (the pt-level value, not perceptual) */
  float ior;
  float transmission;
  float specular;       /* 0.5 = Fresnel reflectance of ior */
  float clearcoat;
  float clearcoat_roughness;
  float sheen;
  float sheen_tint;
  float specular_tint;
  float anisotropic;
  float emission[3];
  float emission_strength;
  float subsurface;
  float subsurface_color[3];
  float subsurface_radius;
  float alpha;
  int32_t thin;
  int32_t base_color_tex;
  int32_t normal_tex;
  int32_t roughness_tex;
  int32_t metallic_tex;
  int32_t emission_tex;
  int32_t alpha_tex;
  phx_param_source base_color_src;
  phx_param_source alpha_src;
  phx_param_source metallic_src;
  phx_param_source roughness_src;
  phx_param_source emission_src;
  phx_param_source transmission_src;
  phx_param_source normal_src; /* tangent-space normal map: the value c encodes 2c - 1 */
  int32_t alpha_mode;          /* PHX_ALPHA_* */
  float alpha_cutoff;
  int32_t flags;               /* PHX_MATERIAL_FLAG_* */
  int32_t base_color_alpha;    /* 1: the base colour source's alpha multiplies alpha */
  int32_t roughness_src_squared; /* 1: a sourced roughness is perceptual and is squared */
} phx_material;

/* pt::Light (light.h:19-35). 92 bytes. */
typedef struct phx_light {
  uint32_t type;
  float position[3];
  float direction[3];
  float color[3];
  float intensity;
  float spot_angle;
  float spot_falloff;
  uint32_t mesh_prim_begin;
  uint32_t mesh_prim_count;
  float area;
  float edge1[3];
  float edge2[3];
  float radius;
} phx_light;

/* SamplerType and WrapMode (sampler.h:36-37). */
#define PHX_SAMPLER_IMAGE1D 0
#define PHX_SAMPLER_IMAGE2D 1
#define PHX_SAMPLER_IMAGE3D 2
#define PHX_SAMPLER_TRANSFORM 3
#define PHX_WRAP_CLAMP 0
#define PHX_WRAP_REPEAT 1
#define PHX_WRAP_MIRROR 2

/* pt::SamplerInfo (sampler.h:39-52): an image (1-3D, RGBA texels in the
 * scene's sampler texel array, row by row from coordinate 0) or a transform,
 * looked up with an attribute of the surface point.  200 bytes. */
typedef struct phx_sampler_info {
  uint32_t type;         /* PHX_SAMPLER_* */
  int32_t in_attribute;  /* PHX_ATTRIBUTE_* slot */
  uint32_t linear;
  uint32_t wrap[3];      /* PHX_WRAP_* */
  uint32_t offset;       /* first texel */
  uint32_t size[3];
  float in_transform[16];  /* row-major 4x4 */
  float in_offset[4];
  float out_transform[16];
  float out_offset[4];
} phx_sampler_info;

/* pt::TextureInfo (texture.h:11-17): texels are RGB floats in one atlas. */
typedef struct phx_texture_info {
  uint32_t offset;
  uint32_t width;
  uint32_t height;
  uint32_t pad_;
} phx_texture_info;

/* pt::TfVolume (tf_volume.h:14-33). 132 bytes. */
typedef struct phx_volume {
  float world_to_local[12]; /* row-major 3x4 */
  float origin[3];
  float spacing[3];
  uint32_t dims[3];
  uint32_t data_offset;
  uint32_t tf_offset;
  uint32_t tf_size;
  float value_lo;
  float value_hi;
  float density_scale;
  float majorant;
  uint32_t linear;
  uint32_t macro_offset;
  uint32_t macro_dims[3];
} phx_volume;

/* Everything pt::Scene (scene.h:41-72) holds, as host arrays. */
typedef struct phx_scene_desc {
  const float *positions;       /* vertex_count * 3 */
  uint32_t vertex_count;
  const uint32_t *indices;      /* triangle_count * 3 */
  uint32_t triangle_count;
  const float *normals;         /* NULL or vertex_count * 3 */
  const float *texcoords;       /* NULL or vertex_count * 2 */
  const float *vertex_colors;   /* NULL or vertex_count * 3 (x < 0: none) */
  const uint32_t *material_ids; /* NULL or triangle_count */

  const phx_material *materials;
  uint32_t material_count;
  const phx_light *lights;
  uint32_t light_count;

  /* Emissive primitives for NEE (Scene::emissive_prim_ids/areas). NULL: the
   * library derives them from the materials as the render bridge does
   * (RenderBridge.cpp:780-813). */
  const uint32_t *emissive_prim_ids;
  const float *emissive_prim_areas;
  uint32_t emissive_count;

  const float *texture_pixels; /* RGB */
  uint32_t texture_pixel_count;
  const phx_texture_info *texture_infos;
  uint32_t texture_count;

  /* Per-vertex attributes (MeshAttributes, sampler.h:25-34): RGBA records,
   * slot k's records start at attribute_offset[k] (-1: the slot is absent). */
  const float *attributes;
  uint32_t attribute_count; /* records */
  int32_t attribute_offset[PHX_ATTRIBUTE_MESH_COUNT];
  /* Samplers (SamplerTable, sampler.h:56-139) and their RGBA texels. */
  const phx_sampler_info *samplers;
  uint32_t sampler_count;
  const float *sampler_texels;
  uint32_t sampler_texel_count; /* texels */

  /* Equirect environment (environment_map.h:13-25): row 0 is +Y of the map. */
  uint32_t env_width;
  uint32_t env_height;
  const float *env_pixels; /* env_height * env_width * 3 */
  float env_world_to_tex[9];
  int32_t env_visible;

  const phx_volume *volumes;
  uint32_t volume_count;
  const float *volume_values;
  uint64_t volume_value_count;
  const float *volume_tf; /* RGBA entries */
  uint32_t volume_tf_count;
  const float *volume_majorants;
  uint32_t volume_majorant_count;

  float ray_epsilon; /* 0: derived from the bounds (pathtracer.cpp:489-502) */
} phx_scene_desc;

/* pt::Camera (camera.h:11-26) in its derived form. 104 bytes. */
typedef struct phx_camera {
  float origin[3];
  float lower_left[3];
  float horizontal[3];
  float vertical[3];
  float u[3], v[3], w[3];
  float lens_radius;
  float focus_dist;
  uint8_t is_ortho;
  uint8_t pad_[3];
  float clip_near;
  float clip_far;
} phx_camera;

/* pt::RenderParams (pathtracer.h:16-55).  The background image is a device
 * array of background_height rows of background_width RGBA floats, rows from
 * the bottom of the frame (NULL: the background colour). */
typedef struct phx_render_params {
  uint32_t width;
  uint32_t height;
  uint32_t samples_per_pixel;
  uint32_t max_depth;
  uint32_t sample_offset;
  float background_color[3];
  float background_alpha;
  float ambient_color[3];
  float ambient_radiance;
  int32_t adaptive_sampling; /* variance-based early exit (pathtracer.cpp:676-723); phx_render only */
  float throughput_clamp;
  int32_t max_diffuse_bounces;
  int32_t max_glossy_bounces;
  int32_t max_transmission_bounces;
  int32_t max_volume_bounces;
  float clamp_direct_sum;
  float clamp_indirect_sum;
  int32_t bottom_up;
  float image_region[4];
  const float *background_image;
  uint32_t background_width;
  uint32_t background_height;
} phx_render_params;

/* pt::HitResult (ray_backend.h:15-29). 88 bytes. */
typedef struct phx_hit {
  float t;
  float position[3];
  float normal[3];
  float shading_normal[3];
  float uv[2];
  float bary[2];
  float interpolated_color[3];
  uint32_t prim_id;
  uint32_t geom_id;
  uint32_t inst_id;
  uint32_t material_id;
  uint8_t hit;
  uint8_t has_interpolated_color;
  uint8_t pad_[2];
} phx_hit;

/* Compact closest-hit record used by the integrator: t, primitive, (b1, b2). */
typedef struct phx_hit_compact {
  float t;
  uint32_t prim_id; /* 0xffffffff: no hit */
  float b1;
  float b2;
} phx_hit_compact;

/* Launch shape and queue sizing, chosen by the autotuner (phx_autotune) or by
 * hand.  0 in a field selects its default. */
typedef struct phx_tuning {
  uint32_t trace_block;          /* work-group size of the traversal kernels: 128 or 256 (256) */
  uint32_t trace_groups_per_cu;  /* persistent traversal work-groups per CU, 1..32 (8) */
  uint32_t refill_lanes;         /* idle lanes of a traversal wave that make it take new rays, 1..64
                                    (0: 12 when the run's deepest BVH has 6+ levels, else 40) */
  uint32_t queue_chunk;          /* rays a wave takes per queue fetch, 64..4096, multiple of 64 (128) */
  uint32_t paths_in_flight;      /* path slots of the wavefront (2^26) */
  uint32_t sort_by_material;     /* material groups of sorted shading, 0..15 (0: queue order) */
  uint32_t use_graph;            /* 1: phx_render_batch replays its bounce loop as a captured hipGraph */
  uint32_t time_kernels;         /* 1: time each kernel launch with HIP events (phx_stats) */
} phx_tuning;

/* Counters kept by the kernels (active rays only). */
typedef struct phx_stats {
  uint64_t closest_rays;
  uint64_t shadow_rays;
  uint64_t shade_events;
  uint64_t samples;
  uint64_t pixels;
  uint64_t scene_bytes;
  uint64_t bvh_nodes;
  uint64_t bvh_bytes;
  double last_render_ms;
  /* With phx_tuning.time_kernels = 1: HIP-event time of every launch of the
   * closest-hit (extend), shade and shadow kernels of the last render.  The
   * shadow pass runs beside the next extend and shade launches, so the three
   * sums can exceed the render time. */
  uint64_t extend_launches;
  double extend_ms;
  uint64_t shade_launches;
  double shade_ms;
  uint64_t shadow_launches;
  double shadow_ms;
} phx_stats;

typedef struct phx_context phx_context;

/* ---- context ------------------------------------------------------------
 * Replaces create_best_backend() + KokkosBackend construction
 * (backend_factory.cpp:28-80, kokkos_backend.h:11-20) and Renderer::initialize
 * (RenderBridge.h:185). */
int phx_create(phx_context **out, int device);
void phx_destroy(phx_context *ctx);
const char *phx_last_error(const phx_context *ctx);
/* RayBackend::name() (ray_backend.h:53): "hip". */
const char *phx_backend_name(void);
int phx_abi_version(void);
int phx_set_tuning(phx_context *ctx, const phx_tuning *tuning);
int phx_get_tuning(const phx_context *ctx, phx_tuning *tuning);
int phx_get_stats(const phx_context *ctx, phx_stats *stats);

/* ---- seam A: RayBackend (ray_backend.h:44-54) ---------------------------
 * phx_build_accel replaces RayBackend::build_accel(const Scene&)
 * (kokkos_backend.cpp:11-19): only the geometry members of the descriptor
 * are read.  The trace calls replace trace_closest / trace_occluded
 * (kokkos_backend.cpp:21-86): device pointers, rays as in RayBatch (origins
 * and directions 3 floats each, tmin, tmax); ray i is inactive iff
 * !(tmax > tmin).  They are asynchronous on `stream` (NULL: the context's). */
int phx_build_accel(phx_context *ctx, const phx_scene_desc *scene);
int phx_trace_closest(phx_context *ctx, const float *d_origins, const float *d_directions,
                      const float *d_tmin, const float *d_tmax, uint32_t count, phx_hit *d_hits,
                      void *stream);
int phx_trace_closest_compact(phx_context *ctx, const float *d_origins, const float *d_directions,
                              const float *d_tmin, const float *d_tmax, uint32_t count,
                              phx_hit_compact *d_hits, void *stream);
int phx_trace_occluded(phx_context *ctx, const float *d_origins, const float *d_directions,
                       const float *d_tmin, const float *d_tmax, uint32_t count,
                       uint32_t *d_occluded, void *stream);
/* phx_trace_closest without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops).  The kernels' scratch (chunk counters and
 * traversal-stack spill) is a static 64 MiB device array per GPU, looked up
 * once per device by phx_create; the persistent grid is sized so that a BVH's
 * worst-case spill fits it.  Calls whose kernels overlap on one GPU (other
 * streams, threads or contexts) share that array and race.  A BVH too deep
 * for one work-group per CU gives PHX_ERR_OVERFLOW.  Results equal
 * phx_trace_closest's. */
int phx_trace_closest_noalloc(phx_context *ctx, const float *d_origins, const float *d_directions,
                              const float *d_tmin, const float *d_tmax, uint32_t count, phx_hit *d_hits,
                              void *stream);
int phx_synchronize(phx_context *ctx);

/* ---- integrator: PathTracer (pathtracer.h:65-73) ------------------------
 * phx_set_scene replaces PathTracer::set_scene + build_accel
 * (pathtracer.cpp:393-503, RenderBridge.cpp:822-825).  phx_render replaces
 * PathTracer::render (pathtracer.cpp:564-1159); its result is read with
 * phx_read_result (RenderResult, pathtracer.h:57-63; rows as rendered) into
 * host or device memory (NULL skips an output). */
int phx_set_scene(phx_context *ctx, const phx_scene_desc *scene);
int phx_render(phx_context *ctx, const phx_camera *camera, const phx_render_params *params);
int phx_read_result(phx_context *ctx, float *color_rgb, float *alpha, float *depth, float *normal_xyz,
                    float *albedo_rgb);
/* Device pointer of the resolved color (h*w*3 floats), valid until the next render. */
int phx_result_device_color(phx_context *ctx, const float **d_color);
/* Samples each pixel took (h*w uint32, rows as rendered, host or device
 * memory) in the last frame, when phx_render sampled it adaptively; else
 * PHX_ERR_NO_FRAME.  No counterpart in the reference, whose per-pixel counts
 * (PathTracer::m_pixel_samples, pathtracer.h:108) are private: it lets a
 * caller see where the variance-based early exit stopped. */
int phx_read_sample_counts(phx_context *ctx, uint32_t *counts);

/* ---- batched views ---------------------------------------------------------
 * Renders many views, of one or several scenes, in one wavefront: all views
 * share the path queues, so the kernels stay full across views of uneven size
 * and the per-frame tail is paid once.  The reference renders one view per
 * PathTracer::render call (pathtracer.cpp:564-1159); each view's result here
 * equals phx_render of that view bit for bit (same sample streams, same
 * summation order).  `scene` names the context holding the view's scene
 * (phx_set_scene); it must live on ctx's device and stay alive during the
 * call; it may be ctx itself.  With phx_tuning.use_graph = 1 the bounce loop
 * is captured once as a hipGraph and replayed.  At most 32768 views per
 * call. */
typedef struct phx_view {
  phx_context *scene;
  phx_camera camera;
  phx_render_params params;
} phx_view;
int phx_render_batch(phx_context *ctx, const phx_view *views, uint32_t count);
/* Result of view `view` of the last phx_render_batch, as phx_read_result. */
int phx_read_batch_result(phx_context *ctx, uint32_t view, float *color_rgb, float *alpha, float *depth,
                          float *normal_xyz, float *albedo_rgb);
/* Times phx_render_batch of `views` under a set of launch shapes (traversal
 * work-group size and groups per CU, queue chunk, refill threshold, material groups, path slots) and keeps the fastest in ctx's tuning.  `best` (optional) receives it,
 * `best_ms` (optional) its time.  Each candidate runs `repeats` times (>= 1). */
int phx_autotune(phx_context *ctx, const phx_view *views, uint32_t count, uint32_t repeats, phx_tuning *best,
                 double *best_ms);

/* ---- differentiable form: path-replay adjoint ------------------------------
 * No reference counterpart (the reference renders only).  Adds to
 * d_grad_materials[m * 5 + k] (k = base colour r, g, b, roughness, metallic of
 * material m; material_count * 5 floats) and, when d_grad_texture is not NULL,
 * to d_grad_texture[p * 3 + c] (texture-atlas texel p) the gradient of a
 * loss of the frame (camera, params) given d_adjoint = dLoss/dImage (RGB per
 * pixel, the layout of phx_read_result's color; all device memory).  The frame
 * is rendered with its own sample streams (params->sample_offset picks them;
 * use samples disjoint from the primal's for a decorrelated gradient) and
 * every path is replayed from its seed: sampling decisions are held fixed and
 * derivatives flow through the BSDF weights of next-event estimation and of
 * the continuing BSDF sample (delta and transmission lobes carry none).
 * Memory beyond the render's is one radiance value per sample in flight.
 * Roughness and metallic gradients are zero outside (0, 1), where the
 * material resolve clamps them; a texture-mapped base colour sends its
 * gradient to the texels.  Clamps (throughput, direct/indirect) are not
 * differentiated. */
int phx_adjoint(phx_context *ctx, const phx_camera *camera, const phx_render_params *params,
                const float *d_adjoint, float *d_grad_materials, float *d_grad_texture);
/* Writes base colour, roughness and metallic of materials 0..count-1 from a
 * device array of count * 5 floats (the gradient layout of phx_adjoint). */
int phx_set_material_params(phx_context *ctx, const float *d_params, uint32_t count);
/* Overwrites texture-atlas texels [first, first + count) from device RGB floats. */
int phx_set_texture_pixels(phx_context *ctx, const float *d_pixels, uint32_t first, uint32_t count);

/* Baseline: the same frame with the reference's algorithmic structure on this
 * GPU, standing in for the reference's Kokkos-HIP build (which needs Kokkos):
 * binary median-split BVH with 44-byte nodes (bvh.cpp:20-125), 64-entry
 * per-thread stack traversal (mesh_intersector.h:57-124), max_depth + 9
 * full-width iterations per sample of trace_closest (96-byte hit records),
 * pt_shade, trace_occluded and pt_nee_accum_pack with a host readback of the
 * active count after each (pathtracer.cpp:564-1159, kokkos_backend.cpp:21-86).
 * Same result interface as phx_render. */
int phx_render_reference_structure(phx_context *ctx, const phx_camera *camera, const phx_render_params *params);

/* Disney BSDF of the shading kernels on n tuples in device memory: materials,
 * wo, wi, geometric normal, shading normal (3 floats each).  phx_bsdf_eval
 * replaces disney_bsdf_eval + disney_bsdf_pdf (disney_bsdf.h:288-355): f*cos
 * (3 floats) and pdf per tuple.  phx_bsdf_sample replaces disney_bsdf_sample
 * (disney_bsdf.h:381-521) with an xorshift32 stream seeded per tuple (rng.h:13):
 * wi, f, pdf and flags (bit 0 specular, bit 1 diffuse).  Asynchronous. */
int phx_bsdf_eval(phx_context *ctx, const phx_material *d_materials, const float *d_wo, const float *d_wi,
                  const float *d_ng, const float *d_sn, uint32_t count, float *d_f, float *d_pdf, void *stream);
int phx_bsdf_sample(phx_context *ctx, const phx_material *d_materials, const float *d_wo, const float *d_ng,
                    const float *d_sn, const uint32_t *d_seeds, uint32_t count, float *d_wi, float *d_f,
                    float *d_pdf, uint32_t *d_flags, void *stream);

/* Derivatives of phx_bsdf_eval's f*cos with respect to base colour r, g, b,
 * roughness and metallic (the adjoint's parameters): df[i*15 + k*3 + c] =
 * d f_c / d param_k of tuple i; f as phx_bsdf_eval.  No reference
 * counterpart; a check of the path-replay adjoint's derivative code. */
int phx_bsdf_eval_grad(phx_context *ctx, const phx_material *d_materials, const float *d_wo, const float *d_wi,
                       const float *d_ng, const float *d_sn, uint32_t count, float *d_f, float *d_df, void *stream);

/* Camera::make_perspective / make_orthographic (camera.h:32-77). */
int phx_camera_perspective(phx_camera *out, const float look_from[3], const float look_at[3],
                           const float vup[3], float vfov_degrees, float aspect, float aperture,
                           float focus_dist);
int phx_camera_orthographic(phx_camera *out, const float look_from[3], const float look_at[3],
                            const float vup[3], float height, float aspect);

/* Macro-cell majorants of a volume (RenderBridge.cpp:656-778): fills
 * vol->macro_dims and vol->majorant and writes the cells to `out`
 * (capacity in floats; the needed count is returned in *out_count).
 * vol->macro_dims all 0: the reference's choice of cell counts; all > 0: that
 * coarse grid (e.g. 32^3 over a 256^3 volume), bounded like the reference's
 * coarse cells.  out = NULL only computes the count. */
int phx_volume_majorants(phx_volume *vol, const float *values, const float *tf_rgba,
                         uint32_t tf_size, float *out, uint32_t capacity, uint32_t *out_count);

/* ---- display: tone mapping and the denoiser seam --------------------------
 * The reference's display transforms, applied per pixel on the GPU to a
 * radiance image: exposure_adjust (c * 2^exposure), then aces_tonemap or
 * reinhard_tonemap (tone_mapping.h:7-21), then a transfer curve
 * (gamma_correct, the piecewise sRGB curve of tone_mapping.h:23-28, or
 * linear_to_srgb, c^(1/2.2) of color.h:9-12), with clamp01 (math.h:86-90)
 * after the tone map and after the curve when `clamp` is set, as
 * photon_render's postprocess does (photon_render.cpp:29-35: ACES, clamp,
 * sRGB curve, clamp).  The 8-bit output is to_u8 (photon_render.cpp:24-27):
 * clamp01(x) * 255.999, truncated. */
#define PHX_TONEMAP_NONE 0
#define PHX_TONEMAP_ACES 1
#define PHX_TONEMAP_REINHARD 2
#define PHX_CURVE_LINEAR 0
#define PHX_CURVE_SRGB 1  /* gamma_correct */
#define PHX_CURVE_GAMMA22 2 /* linear_to_srgb */
typedef struct phx_display {
  int32_t tonemap;  /* PHX_TONEMAP_* */
  int32_t curve;    /* PHX_CURVE_* */
  float exposure;   /* stops; 0 leaves the radiance as is */
  uint8_t clamp;    /* clamp01 after the tone map and after the curve */
  uint8_t pad_[3];
} phx_display;
/* d_rgb: `pixels` pixels of `in_channels` (3 or 4; a fourth channel is
 * skipped) floats in device memory; writes d_out_rgb (pixels * 3 floats)
 * and/or d_out_rgb8 (pixels * 3 bytes), either may be NULL.  Runs on
 * `stream` (NULL: the context's stream), asynchronously. */
int phx_display_transform(phx_context *ctx, const float *d_rgb, uint32_t in_channels, uint64_t pixels,
                          const phx_display *display, float *d_out_rgb, uint8_t *d_out_rgb8, void *stream);
/* Denoiser::is_available / denoise_buffer (denoiser.h:14-31): the reference
 * wraps Intel Open Image Denoise when it is built with it; this build has no
 * OIDN, so, as the reference without it, is_available is 0 and
 * denoise_buffer leaves the colour as it is (color, albedo and normal are
 * w*h*3 host floats). */
int phx_denoiser_is_available(void);
int phx_denoise_buffer(float *color, const float *albedo, const float *normal, uint32_t width, uint32_t height);

/* ---- pbrt-v4 scene files ----------------------------------------------------
 * parse_pbrt_file (pbrt_parser.cpp:233) followed by convert_pbrt_scene
 * (pbrt_to_photon.cpp:58-503): named materials, triangle meshes, PLY meshes,
 * spheres, area lights, image textures, an infinite light with an image,
 * object instances (flattened, within the reference's limits of 500
 * instances per object and 5e8 triangles; with use_instancing they are only
 * counted), Include/Import, the camera and the Film/Sampler/Integrator
 * settings.  Image files: PFM, binary PNM, Radiance HDR and PNG; others load
 * as missing textures (left out, as the reference does with files it cannot
 * read).  A missing scene file gives an empty scene, as in the reference:
 * its mesh is one degenerate triangle that no ray hits.  The handle owns
 * the converted arrays. */
typedef struct phx_pbrt_scene phx_pbrt_scene;
typedef struct phx_pbrt_info {
  uint32_t width, height, spp, max_depth; /* Film, Sampler, Integrator */
  uint32_t meshes;    /* meshes placed in the scene (instances flattened) */
  uint32_t objects;   /* ObjectBegin definitions */
  uint32_t instances; /* ObjectInstance uses */
  uint32_t textures;  /* image textures loaded */
} phx_pbrt_info;
int phx_pbrt_load(const char *path, int use_instancing, phx_pbrt_scene **out);
/* The converted scene: `desc` points into the handle's arrays (valid until
 * phx_pbrt_free; pass it to phx_set_scene), `camera` is
 * Camera::make_perspective of the file's camera; any output may be NULL. */
int phx_pbrt_scene_desc(const phx_pbrt_scene *scene, phx_scene_desc *desc, phx_camera *camera,
                        phx_pbrt_info *info);
void phx_pbrt_free(phx_pbrt_scene *scene);

/* ---- seam B: photon::bridge::Renderer (RenderBridge.h:175-214) ----------
 * Progressive accumulation over frames (RenderBridge.cpp:828-928): `reset`
 * drops the accumulated samples; sample_offset advances by the frame's spp.
 * read_* give the accumulated frame, rows starting at the bottom. */

/* bridge::CameraData (RenderBridge.h:128-144). 84 bytes. */
typedef struct phx_camera_data {
  float position[3];
  float direction[3];
  float up[3];
  uint8_t orthographic;
  uint8_t pad_[3];
  float fovy; /* radians */
  float height;
  float aspect;
  float aperture_radius;
  float focus_distance;
  float image_region[4];
  float clip_near;
  float clip_far;
} phx_camera_data;

/* bridge::RenderSettings (RenderBridge.h:146-170). 96 bytes.  The background
 * image is host memory: background_height rows of background_width RGBA
 * floats, rows from the bottom (NULL: the background colour). */
typedef struct phx_render_settings {
  uint32_t width;
  uint32_t height;
  uint32_t pixel_samples;
  uint32_t max_depth;
  float background[4];
  const float *background_image;
  uint32_t background_width;
  uint32_t background_height;
  float ambient_color[3];
  float ambient_radiance;
  uint8_t adaptive_sampling;
  uint8_t pad_[3];
  float throughput_clamp;
  int32_t max_diffuse_bounces;
  int32_t max_glossy_bounces;
  int32_t max_transmission_bounces;
  int32_t max_volume_bounces;
  float clamp_direct;
  float clamp_indirect;
} phx_render_settings;

int phx_bridge_render(phx_context *ctx, const phx_camera_data *camera, const phx_render_settings *settings,
                      int reset);

/* bridge::MaterialData (RenderBridge.h:21-56). 132 bytes.  Sources are
 * ParamSourceData {kind: 0 constant, 1 vertex attribute, 2 sampler; index}. */
typedef struct phx_bridge_material {
  float base_color[3];
  float opacity;
  float metallic;
  float roughness; /* perceptual: the GGX alpha is its square */
  float ior;
  float transmission;
  float specular; /* weight of the dielectric specular lobe, 0..1 */
  float emission[3];
  float clearcoat;
  float clearcoat_roughness;
  float sheen;
  phx_param_source base_color_src;
  phx_param_source opacity_src;
  phx_param_source metallic_src;
  phx_param_source roughness_src;
  phx_param_source emission_src;
  phx_param_source transmission_src;
  phx_param_source normal_src; /* tangent-space normal map, encoded as 0.5 + 0.5 n */
  int32_t alpha_mode; /* 0 blend, 1 opaque, 2 mask */
  float alpha_cutoff;
  uint8_t base_color_alpha;
  uint8_t thin_walled;
  uint8_t emitter_only;
  uint8_t hidden_to_camera;
  int32_t emission_side; /* 0 front, 1 back, 2 both */
} phx_bridge_material;

/* convert_material (RenderBridge.cpp:324-370): the pt-level material the
 * bridge renders a MaterialData with.  Roughness is squared (the BSDF takes
 * the GGX alpha), specular halved (0.5 is the full dielectric reflectance),
 * emission_strength is 1, alpha the opacity (1 when the opacity has a
 * source, which then scales it), the sources copied with
 * roughness_src_squared = 1, the alpha mode recoded to PHX_ALPHA_*,
 * REFRACTIVE always set and the emitter-only, hidden-to-camera and
 * emission-side flags added.  Fields the bridge leaves alone keep
 * pt::Material's defaults (material.h:35-82). */
int phx_convert_material(const phx_bridge_material *in, phx_material *out);

/* bridge::LightData (RenderBridge.h:73-85). 56 bytes. */
typedef struct phx_light_data {
  int32_t type; /* 0 point, 1 directional, 2 spot */
  float position[3];
  float direction[3];
  float color[3];
  float intensity;
  float radius;
  float spot_angle;
  float spot_falloff;
} phx_light_data;

/* bridge::SamplerData (RenderBridge.h:58-71); the texel vector as pointer +
 * count (floats: 4 per RGBA texel). */
typedef struct phx_sampler_data {
  int32_t type; /* 0 image1D, 1 image2D, 2 image3D, 3 transform */
  int32_t in_attribute;
  uint8_t linear;
  uint8_t pad_[3];
  int32_t wrap[3]; /* 0 clampToEdge, 1 repeat, 2 mirrorRepeat */
  uint32_t size[3];
  float in_transform[16];
  float in_offset[4];
  float out_transform[16];
  float out_offset[4];
  const float *texels;
  uint64_t texel_float_count;
} phx_sampler_data;

/* bridge::EnvironmentData (RenderBridge.h:87-97): RGB radiance rows, row 0
 * the bottom (-up) row, as pointer + count (floats). */
typedef struct phx_environment_data {
  uint8_t present;
  uint8_t visible;
  uint8_t pad_[2];
  uint32_t width;
  uint32_t height;
  const float *pixels;
  uint64_t pixel_float_count;
  float world_to_map[9]; /* world -> map rotation, row-major 3x3 */
} phx_environment_data;

/* bridge::VolumeData (RenderBridge.h:99-110), vectors as pointer + count. */
typedef struct phx_volume_data {
  float world_to_local[12];
  float origin[3];
  float spacing[3];
  uint32_t dims[3];
  const float *values; /* dims[0] * dims[1] * dims[2], x fastest */
  uint64_t value_count;
  const float *transfer; /* RGBA table spanning value_range, A the opacity */
  uint64_t transfer_float_count;
  float value_range[2];
  float density_scale;
  uint8_t linear;
  uint8_t pad_[3];
} phx_volume_data;

/* bridge::SceneData (RenderBridge.h:112-125): every std::vector as pointer +
 * element count (float counts for float vectors), so a caller passes
 * v.data(), v.size(). */
typedef struct phx_scene_data {
  const float *positions;
  uint64_t position_count; /* floats, 3 per vertex */
  const float *normals;
  uint64_t normal_count; /* floats; used when >= 3 per vertex */
  const float *attributes[PHX_ATTRIBUTE_MESH_COUNT];
  uint64_t attribute_count[PHX_ATTRIBUTE_MESH_COUNT]; /* floats; a slot is used when >= 4 per vertex */
  const uint32_t *indices;
  uint64_t index_count;
  const uint32_t *material_ids;
  uint64_t material_id_count;
  const phx_bridge_material *materials;
  uint32_t material_count;
  const phx_sampler_data *samplers;
  uint32_t sampler_count;
  const phx_light_data *lights;
  uint32_t light_count;
  phx_environment_data environment;
  const phx_volume_data *volumes;
  uint32_t volume_count;
} phx_scene_data;

/* Renderer::set_scene (RenderBridge.h:189) via Impl::build_scene
 * (RenderBridge.cpp:441-826): converts SceneData as the bridge does
 * (materials by convert_material, samplers, lights, the environment map,
 * volumes with macro-cell majorants, the emissive-triangle list, ray_epsilon
 * from the bounds; an empty mesh becomes one degenerate triangle) and sets the
 * result as the context's scene (phx_set_scene).  Also resets the
 * accumulated frame of seam B. */
int phx_set_scene_data(phx_context *ctx, const phx_scene_data *scene);
uint32_t phx_bridge_accumulated_samples(const phx_context *ctx);
int phx_bridge_read_color(phx_context *ctx, float *rgba);
int phx_bridge_read_depth(phx_context *ctx, float *depth);
int phx_bridge_read_normal(phx_context *ctx, float *xyz);
int phx_bridge_read_albedo(phx_context *ctx, float *rgb);

#ifdef __cplusplus
}
#endif

#endif /* PHX_H */
