// phx_set_scene_data: the render bridge's scene import, Renderer::set_scene
// via Impl::build_scene (RenderBridge.cpp:441-826, 1046-1058).  A SceneData-
// shaped description (phx_scene_data: the bridge's std::vectors as pointer +
// count) is converted into the pt-level scene and handed to phx_set_scene:
//  - mesh: an empty mesh becomes one degenerate triangle (never hit); material
//    ids beyond the material count become 0; normals and attribute slots are
//    used only when they cover every vertex;
//  - materials by convert_material (a default material when there is none);
//  - samplers, with images too small for their size demoted to transforms;
//  - lights (point, directional, spot);
//  - the environment map and its world-to-map rotation;
//  - volumes, with the bridge's macro-cell majorants (phx_volume_majorants);
//  - the ray offset from the bounds of all vertices.
// The emissive-triangle list is derived by phx_set_scene by the bridge's rule.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "context.h"

namespace {

using namespace phx_detail;

phx_light convert_light(const phx_light_data &ld) {
  phx_light l{};
  l.type = ld.type == 1 ? PHX_LIGHT_DIRECTIONAL : (ld.type == 2 ? PHX_LIGHT_SPOT : PHX_LIGHT_POINT);
  for (int k = 0; k < 3; ++k) {
    l.position[k] = ld.position[k];
    l.direction[k] = ld.direction[k];
    l.color[k] = ld.color[k];
  }
  l.intensity = ld.intensity;
  l.radius = ld.radius;
  l.spot_angle = ld.spot_angle;
  l.spot_falloff = ld.spot_falloff;
  return l;
}

}  // namespace

extern "C" int phx_set_scene_data(phx_context *c, const phx_scene_data *in) {
  if (!c || !in) return PHX_ERR_INVALID;
  const bool has_mesh = in->positions && in->indices && in->position_count >= 3 && in->index_count >= 3;
  const bool empty = !has_mesh;
  const size_t nv = empty ? 3 : size_t(in->position_count / 3);
  const size_t nt = empty ? 1 : size_t(in->index_count / 3);
  if (nv >= (size_t(1) << 32) || nt >= (size_t(1) << 28))
    return fail(c, PHX_ERR_INVALID, "set_scene_data: mesh too large");

  // Mesh (RenderBridge.cpp:445-485): positions and their bounds.
  std::vector<float> pos(nv * 3, 0.f);
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = 0; i < nv; ++i)
    for (int k = 0; k < 3; ++k) {
      const float v = empty ? 0.f : in->positions[3 * i + k];
      pos[3 * i + k] = v;
      lo[k] = std::min(lo[k], v);
      hi[k] = std::max(hi[k], v);
    }
  std::vector<uint32_t> idx(nt * 3);
  for (size_t i = 0; i < nt * 3; ++i) {
    idx[i] = empty ? uint32_t(i) : in->indices[i];
    if (idx[i] >= nv) return fail(c, PHX_ERR_INVALID, "set_scene_data: vertex index out of range");
  }
  const uint32_t material_count = std::max<uint32_t>(in->materials ? in->material_count : 0u, 1u);
  std::vector<uint32_t> mids(nt, 0u);
  for (size_t i = 0; i < nt; ++i) {
    const uint32_t id = (!empty && in->material_ids && i < in->material_id_count) ? in->material_ids[i] : 0u;
    mids[i] = id < material_count ? id : 0u;
  }
  std::vector<float> normals;
  if (!empty && in->normals && in->normal_count >= nv * 3) normals.assign(in->normals, in->normals + nv * 3);

  // Vertex attributes (:487-511): the slots that cover every vertex, one
  // after another.
  std::vector<float> attrs;
  phx_scene_desc d{};
  for (int slot = 0; slot < PHX_ATTRIBUTE_MESH_COUNT; ++slot) {
    d.attribute_offset[slot] = -1;
    if (empty || !in->attributes[slot] || in->attribute_count[slot] < nv * 4) continue;
    d.attribute_offset[slot] = int32_t(attrs.size() / 4);
    attrs.insert(attrs.end(), in->attributes[slot], in->attributes[slot] + nv * 4);
  }
  if (attrs.size() / 4 >= (size_t(1) << 31)) return fail(c, PHX_ERR_INVALID, "set_scene_data: too many attribute records");

  // Materials (:513-525).
  std::vector<phx_material> mats;
  for (uint32_t i = 0; in->materials && i < in->material_count; ++i) mats.push_back(convert_bridge_material(in->materials[i]));
  if (mats.empty()) mats.push_back(default_material());

  // Samplers (:527-577).
  std::vector<phx_sampler_info> samplers;
  std::vector<float> texels;
  for (uint32_t i = 0; in->samplers && i < in->sampler_count; ++i) {
    const phx_sampler_data &sd = in->samplers[i];
    phx_sampler_info si{};
    si.type = sd.type == 0 ? PHX_SAMPLER_IMAGE1D
                           : (sd.type == 2 ? PHX_SAMPLER_IMAGE3D : (sd.type == 3 ? PHX_SAMPLER_TRANSFORM : PHX_SAMPLER_IMAGE2D));
    si.in_attribute = sd.in_attribute;
    si.linear = sd.linear ? 1u : 0u;
    for (int k = 0; k < 3; ++k) {
      si.wrap[k] = sd.wrap[k] == 1 ? PHX_WRAP_REPEAT : (sd.wrap[k] == 2 ? PHX_WRAP_MIRROR : PHX_WRAP_CLAMP);
      si.size[k] = std::max<uint32_t>(sd.size[k], 1u);
    }
    si.offset = uint32_t(texels.size() / 4);
    std::memcpy(si.in_transform, sd.in_transform, sizeof si.in_transform);
    std::memcpy(si.in_offset, sd.in_offset, sizeof si.in_offset);
    std::memcpy(si.out_transform, sd.out_transform, sizeof si.out_transform);
    std::memcpy(si.out_offset, sd.out_offset, sizeof si.out_offset);
    const size_t n = sd.texels ? size_t(sd.texel_float_count / 4) : 0;
    const size_t expected = size_t(si.size[0]) * si.size[1] * si.size[2];
    // an incomplete image: a constant sampler keeps the lookups in bounds
    if (si.type != PHX_SAMPLER_TRANSFORM && n < expected) si.type = PHX_SAMPLER_TRANSFORM;
    if (n) texels.insert(texels.end(), sd.texels, sd.texels + n * 4);
    if (texels.size() / 4 >= (size_t(1) << 32)) return fail(c, PHX_ERR_INVALID, "set_scene_data: too many sampler texels");
    samplers.push_back(si);
  }

  // Lights (:579-597).
  std::vector<phx_light> lights;
  for (uint32_t i = 0; in->lights && i < in->light_count; ++i) lights.push_back(convert_light(in->lights[i]));

  // Environment map (:599-624).
  const phx_environment_data &env = in->environment;
  const bool has_env = env.present && env.width > 0 && env.height > 0 && env.pixels &&
                       env.pixel_float_count >= uint64_t(env.width) * env.height * 3;

  // Volumes (:626-778): values and tables one after another, majorants by the
  // bridge's macro-cell rule.
  std::vector<phx_volume> vols;
  std::vector<float> values, tf, majorants;
  for (uint32_t i = 0; in->volumes && i < in->volume_count; ++i) {
    const phx_volume_data &vd = in->volumes[i];
    phx_volume v{};
    std::memcpy(v.world_to_local, vd.world_to_local, sizeof v.world_to_local);
    for (int k = 0; k < 3; ++k) {
      v.origin[k] = vd.origin[k];
      v.spacing[k] = vd.spacing[k];
      v.dims[k] = std::max<uint32_t>(vd.dims[k], 1u);
    }
    const size_t nvals = vd.values ? size_t(vd.value_count) : 0;
    const uint32_t tf_size = vd.transfer ? uint32_t(vd.transfer_float_count / 4) : 0u;
    v.data_offset = uint32_t(values.size());
    v.tf_offset = uint32_t(tf.size() / 4);
    v.value_lo = vd.value_range[0];
    v.value_hi = vd.value_range[1];
    v.density_scale = vd.density_scale;
    v.linear = vd.linear ? 1u : 0u;
    v.macro_offset = uint32_t(majorants.size());
    const bool complete = nvals >= size_t(v.dims[0]) * v.dims[1] * v.dims[2];
    uint32_t cells = 0;
    int r = phx_volume_majorants(&v, nullptr, vd.transfer, tf_size, nullptr, 0, &cells);
    if (r) return fail(c, r, "set_scene_data: volume majorants");
    std::vector<float> m(cells);
    r = phx_volume_majorants(&v, complete ? vd.values : nullptr, vd.transfer, tf_size, m.data(), cells, &cells);
    if (r) return fail(c, r, "set_scene_data: volume majorants");
    v.macro_offset = uint32_t(majorants.size());
    v.data_offset = uint32_t(values.size());
    v.tf_offset = uint32_t(tf.size() / 4);
    majorants.insert(majorants.end(), m.begin(), m.end());
    if (nvals) values.insert(values.end(), vd.values, vd.values + nvals);
    if (tf_size) tf.insert(tf.end(), vd.transfer, vd.transfer + size_t(tf_size) * 4);
    if (values.size() >= (size_t(1) << 32) || majorants.size() >= (size_t(1) << 32))
      return fail(c, PHX_ERR_INVALID, "set_scene_data: volume data too large");
    vols.push_back(v);
  }

  d.positions = pos.data();
  d.vertex_count = uint32_t(nv);
  d.indices = idx.data();
  d.triangle_count = uint32_t(nt);
  d.normals = normals.empty() ? nullptr : normals.data();
  d.material_ids = mids.data();
  d.materials = mats.data();
  d.material_count = uint32_t(mats.size());
  d.lights = lights.empty() ? nullptr : lights.data();
  d.light_count = uint32_t(lights.size());
  d.attributes = attrs.empty() ? nullptr : attrs.data();
  d.attribute_count = uint32_t(attrs.size() / 4);
  d.samplers = samplers.empty() ? nullptr : samplers.data();
  d.sampler_count = uint32_t(samplers.size());
  d.sampler_texels = texels.empty() ? nullptr : texels.data();
  d.sampler_texel_count = uint32_t(texels.size() / 4);
  if (has_env) {
    d.env_width = env.width;
    d.env_height = env.height;
    d.env_pixels = env.pixels;
    std::memcpy(d.env_world_to_tex, env.world_to_map, sizeof d.env_world_to_tex);
    d.env_visible = env.visible ? 1 : 0;
  } else {
    d.env_visible = 1;
  }
  d.volumes = vols.empty() ? nullptr : vols.data();
  d.volume_count = uint32_t(vols.size());
  d.volume_values = values.empty() ? nullptr : values.data();
  d.volume_value_count = values.size();
  d.volume_tf = tf.empty() ? nullptr : tf.data();
  d.volume_tf_count = uint32_t(tf.size() / 4);
  d.volume_majorants = majorants.empty() ? nullptr : majorants.data();
  d.volume_majorant_count = uint32_t(majorants.size());
  // Ray offset relative to the bounds of every vertex (:815-820).
  float eps = 1e-3f;
  if (!empty && hi[0] >= lo[0]) {
    const float dg[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    const float e = 1e-4f * std::sqrt(dg[0] * dg[0] + dg[1] * dg[1] + dg[2] * dg[2]);
    eps = std::min(std::max(e, 1e-6f), 1e-2f);
  }
  d.ray_epsilon = eps;
  c->b_samples = 0;  // Renderer::set_scene drops the accumulated frame (:1051-1054)
  return phx_set_scene(c, &d);
}
