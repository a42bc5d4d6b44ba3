// Render-bridge layer (seam B: photon::bridge::Renderer, RenderBridge.h:175-214)
// and the volume majorant builder of the bridge's scene import.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "context.h"

namespace {

using namespace phx;
using namespace phx_detail;

// accumulate_frame (RenderBridge.cpp:386-412): color and alpha weighted by the
// frame's spp; the AOVs of the first frame are kept.
__global__ void k_frame_accumulate(FrameBuffers f, float *bc, float *ba, float *bd, float *bn, float *bl, uint32_t npix,
                                   float weight, uint32_t first) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  if (first) {
    for (int k = 0; k < 3; ++k) {
      bc[p * 3 + k] = f.color[p * 3 + k] * weight;
      bn[p * 3 + k] = f.normal[p * 3 + k];
      bl[p * 3 + k] = f.albedo[p * 3 + k];
    }
    ba[p] = f.alpha[p] * weight;
    bd[p] = f.depth[p];
  } else {
    for (int k = 0; k < 3; ++k) bc[p * 3 + k] = bc[p * 3 + k] + f.color[p * 3 + k] * weight;
    ba[p] = ba[p] + f.alpha[p] * weight;
  }
}

float len3(const float *v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

}  // namespace

extern "C" {

// RenderBridge.cpp:656-778.
int phx_volume_majorants(phx_volume *vol, const float *values, const float *tf, uint32_t tf_size, float *out,
                         uint32_t capacity, uint32_t *out_count) {
  if (!vol || !out_count) return PHX_ERR_INVALID;
  // A count-only call leaves the caller's volume untouched.
  phx_volume local = *vol;
  phx_volume *v = out ? vol : &local;
  float max_op = 0.f;
  for (uint32_t t = 0; t < tf_size; ++t) max_op = std::max(max_op, tf[4 * t + 3]);
  for (int d = 0; d < 3; ++d) v->dims[d] = std::max<uint32_t>(v->dims[d], 1);
  v->tf_size = tf_size;
  v->majorant = max_op * v->density_scale;
  const bool complete = values != nullptr;
  if (!complete) v->majorant = 0.f;
  uint32_t sub[3] = {0, 0, 0};
  const bool preset = v->macro_dims[0] > 0 && v->macro_dims[1] > 0 && v->macro_dims[2] > 0;
  if (preset) {
    // Caller-chosen coarse cells (each covering several grid cells); the
    // bounds use the coarse-cell rule below.
    for (int d = 0; d < 3; ++d) v->macro_dims[d] = std::min<uint32_t>(v->macro_dims[d], 1024);
  } else {
    for (int d = 0; d < 3; ++d) {
      const uint32_t cells = v->dims[d] > 1 ? v->dims[d] - 1 : 1;
      if (cells <= 32) {
        sub[d] = std::min<uint32_t>(std::max<uint32_t>((32 + cells - 1) / cells, 1), 16);
        v->macro_dims[d] = cells * sub[d];
      } else {
        v->macro_dims[d] = std::min<uint32_t>(std::max<uint32_t>((v->dims[d] + 2) / 4, 1), 128);
      }
    }
  }
  const bool refined = sub[0] > 0 && sub[1] > 0 && sub[2] > 0;
  if (!refined && !preset)
    for (int d = 0; d < 3; ++d)
      if (sub[d] > 0) v->macro_dims[d] = std::min<uint32_t>(std::max<uint32_t>((v->dims[d] + 2) / 4, 1), 128);
  const size_t n = size_t(v->macro_dims[0]) * v->macro_dims[1] * v->macro_dims[2];
  *out_count = uint32_t(n);
  if (!out) return PHX_OK;
  if (capacity < n) return PHX_ERR_OVERFLOW;
  const float range = v->value_hi - v->value_lo;
  const uint32_t *dims = v->dims;
  auto node = [&](uint32_t x, uint32_t y, uint32_t z) {
    x = std::min(x, dims[0] - 1);
    y = std::min(y, dims[1] - 1);
    z = std::min(z, dims[2] - 1);
    return values[(size_t(z) * dims[1] + y) * dims[0] + x];
  };
  auto interp = [&](float gx, float gy, float gz) {
    const uint32_t x = uint32_t(gx), y = uint32_t(gy), z = uint32_t(gz);
    const float fx = gx - float(x), fy = gy - float(y), fz = gz - float(z);
    const float c00 = node(x, y, z) * (1.f - fx) + node(x + 1, y, z) * fx;
    const float c10 = node(x, y + 1, z) * (1.f - fx) + node(x + 1, y + 1, z) * fx;
    const float c01 = node(x, y, z + 1) * (1.f - fx) + node(x + 1, y, z + 1) * fx;
    const float c11 = node(x, y + 1, z + 1) * (1.f - fx) + node(x + 1, y + 1, z + 1) * fx;
    return (c00 * (1.f - fy) + c10 * fy) * (1.f - fz) + (c01 * (1.f - fy) + c11 * fy) * fz;
  };
  size_t k = 0;
  for (uint32_t mz = 0; mz < v->macro_dims[2]; ++mz)
    for (uint32_t my = 0; my < v->macro_dims[1]; ++my)
      for (uint32_t mx = 0; mx < v->macro_dims[0]; ++mx) {
        float bound = 0.f;
        if (complete && tf_size > 0) {
          const uint32_t m[3] = {mx, my, mz};
          float vmin = std::numeric_limits<float>::max(), vmax = -std::numeric_limits<float>::max();
          if (refined) {
            for (int corner = 0; corner < 8; ++corner) {
              float g[3];
              for (int d = 0; d < 3; ++d) g[d] = float(m[d] + ((corner >> d) & 1)) / float(sub[d]);
              const float val = interp(g[0], g[1], g[2]);
              vmin = std::min(vmin, val);
              vmax = std::max(vmax, val);
            }
          } else {
            uint32_t first[3], last[3];
            for (int d = 0; d < 3; ++d) {
              const float cells = float(dims[d] > 1 ? dims[d] - 1 : 1);
              const float lo = float(m[d]) / float(v->macro_dims[d]) * cells;
              const float hi = float(m[d] + 1) / float(v->macro_dims[d]) * cells;
              first[d] = uint32_t(std::max(std::floor(lo) - 1.f, 0.f));
              last[d] = std::min<uint32_t>(uint32_t(std::ceil(hi)) + 1, dims[d] - 1);
            }
            for (uint32_t z = first[2]; z <= last[2]; ++z)
              for (uint32_t y = first[1]; y <= last[1]; ++y)
                for (uint32_t x = first[0]; x <= last[0]; ++x) {
                  const float val = values[(size_t(z) * dims[1] + y) * dims[0] + x];
                  vmin = std::min(vmin, val);
                  vmax = std::max(vmax, val);
                }
          }
          auto tpos = [&](float val) {
            float x = range != 0.f ? (val - v->value_lo) / range : 0.f;
            x = std::min(std::max(x, 0.f), 1.f);
            return x * float(tf_size - 1);
          };
          const size_t i0 = size_t(std::floor(tpos(vmin)));
          const size_t i1 = std::min(size_t(std::ceil(tpos(vmax))), size_t(tf_size) - 1);
          for (size_t t = i0; t <= i1; ++t) bound = std::max(bound, tf[4 * t + 3]);
          if (!(vmin <= vmax)) bound = max_op;
        }
        out[k++] = bound * v->density_scale * 1.01f;
      }
  return PHX_OK;
}

// Renderer::Impl::render (RenderBridge.cpp:828-928).
int phx_bridge_render(phx_context *c, const phx_camera_data *cd, const phx_render_settings *rs, int reset) {
  if (!c || !cd || !rs) return PHX_ERR_INVALID;
  const uint32_t w = rs->width, h = rs->height;
  if (w == 0 || h == 0) return fail(c, PHX_ERR_INVALID, "invalid frame size");
  if (!c->have_scene) return fail(c, PHX_ERR_NO_SCENE, "no scene");
  if (w != c->b_w || h != c->b_h) {
    const size_t n = size_t(w) * h;
    HIP_TRY(c, c->b_color.reserve(n * 12));
    HIP_TRY(c, c->b_alpha.reserve(n * 4));
    HIP_TRY(c, c->b_depth.reserve(n * 4));
    HIP_TRY(c, c->b_normal.reserve(n * 12));
    HIP_TRY(c, c->b_albedo.reserve(n * 12));
    c->b_w = w;
    c->b_h = h;
    reset = 1;
  }
  float pos[3] = {cd->position[0], cd->position[1], cd->position[2]};
  float dir[3] = {cd->direction[0], cd->direction[1], cd->direction[2]};
  float dl = len3(dir);
  if (dl > 0.f) {
    const float inv = 1.f / dl;
    for (float &x : dir) x *= inv;
  } else {
    dir[0] = 0.f; dir[1] = 0.f; dir[2] = -1.f;
  }
  float up[3] = {cd->up[0], cd->up[1], cd->up[2]};
  const float cr[3] = {dir[1] * up[2] - dir[2] * up[1], dir[2] * up[0] - dir[0] * up[2], dir[0] * up[1] - dir[1] * up[0]};
  if (len3(cr) < 1e-6f) {
    if (std::fabs(dir[1]) < 0.9f) { up[0] = 0.f; up[1] = 1.f; up[2] = 0.f; }
    else { up[0] = 1.f; up[1] = 0.f; up[2] = 0.f; }
  }
  const float at[3] = {pos[0] + dir[0], pos[1] + dir[1], pos[2] + dir[2]};
  const float aspect = cd->aspect > 0.f ? cd->aspect : float(w) / float(h);
  phx_camera cam{};
  int r;
  if (cd->orthographic) {
    r = phx_camera_orthographic(&cam, pos, at, up, cd->height, aspect);
  } else {
    const float fovy = cd->fovy * 180.f / 3.14159265358979f;
    const bool lens = cd->aperture_radius > 0.f && cd->focus_distance > 0.f;
    r = phx_camera_perspective(&cam, pos, at, up, fovy, aspect, lens ? 2.f * cd->aperture_radius : 0.f,
                               lens ? cd->focus_distance : 1.f);
  }
  if (r) return r;
  cam.clip_near = cd->clip_near > 0.f ? cd->clip_near : 0.f;
  cam.clip_far = cd->clip_far < 1e30f ? cd->clip_far : 1e30f;

  phx_render_params p{};
  p.width = w;
  p.height = h;
  p.samples_per_pixel = std::max<uint32_t>(rs->pixel_samples, 1);
  p.max_depth = std::max<uint32_t>(rs->max_depth, 1);
  p.sample_offset = c->b_counter;
  for (int i = 0; i < 3; ++i) p.background_color[i] = rs->background[i];
  p.background_alpha = rs->background[3];
  for (int i = 0; i < 3; ++i) p.ambient_color[i] = rs->ambient_color[i];
  p.ambient_radiance = rs->ambient_radiance;
  p.adaptive_sampling = rs->adaptive_sampling;
  p.throughput_clamp = rs->throughput_clamp;
  p.max_diffuse_bounces = rs->max_diffuse_bounces;
  p.max_glossy_bounces = rs->max_glossy_bounces;
  p.max_transmission_bounces = rs->max_transmission_bounces;
  p.max_volume_bounces = rs->max_volume_bounces;
  p.clamp_direct_sum = 3.f * std::max(rs->clamp_direct, 0.f);
  p.clamp_indirect_sum = 3.f * std::max(rs->clamp_indirect, 0.f);
  p.bottom_up = 1;
  for (int i = 0; i < 4; ++i) p.image_region[i] = cd->image_region[i];
  // Background image (RenderBridge.cpp:846-869): a device copy, uploaded again
  // when the caller's image changes.
  const size_t bg_n = size_t(rs->background_width) * rs->background_height * 4;
  if (rs->background_image && bg_n > 0) {
    if (c->b_bg_host.size() != bg_n || c->b_bg_w != rs->background_width ||
        std::memcmp(c->b_bg_host.data(), rs->background_image, bg_n * sizeof(float)) != 0) {
      c->b_bg_host.assign(rs->background_image, rs->background_image + bg_n);
      c->b_bg_w = rs->background_width;
      HIP_TRY(c, c->b_bg.reserve(bg_n * sizeof(float)));
      HIP_TRY(c, hipMemcpy(c->b_bg.p, c->b_bg_host.data(), bg_n * sizeof(float), hipMemcpyHostToDevice));
    }
    p.background_image = c->b_bg.as<float>();
    p.background_width = rs->background_width;
    p.background_height = rs->background_height;
  } else {
    c->b_bg_host.clear();
  }
  if ((r = phx_render(c, &cam, &p))) return r;

  const uint32_t spp = p.samples_per_pixel;
  c->b_counter += spp;
  const bool first = reset || c->b_samples == 0;
  FrameBuffers fb{c->f_accum.as<float4>(), c->f_color.as<float>(),  c->f_alpha.as<float>(),
                  c->f_depth.as<float>(),  c->f_normal.as<float>(), c->f_albedo.as<float>()};
  const uint32_t npix = w * h;
  hipLaunchKernelGGL(k_frame_accumulate, dim3((npix + 255) / 256), dim3(256), 0, c->stream, fb, c->b_color.as<float>(),
                     c->b_alpha.as<float>(), c->b_depth.as<float>(), c->b_normal.as<float>(), c->b_albedo.as<float>(),
                     npix, float(spp), first ? 1u : 0u);
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->b_samples = first ? spp : c->b_samples + spp;
  return PHX_OK;
}

uint32_t phx_bridge_accumulated_samples(const phx_context *c) { return c ? c->b_samples : 0u; }

int phx_bridge_read_color(phx_context *c, float *rgba) {
  if (!c || !rgba) return PHX_ERR_INVALID;
  if (c->b_samples == 0) return fail(c, PHX_ERR_NO_FRAME, "no frame has been rendered");
  const size_t n = size_t(c->b_w) * c->b_h;
  std::vector<float> col(n * 3), al(n);
  HIP_TRY(c, hipMemcpy(col.data(), c->b_color.p, n * 12, hipMemcpyDeviceToHost));
  HIP_TRY(c, hipMemcpy(al.data(), c->b_alpha.p, n * 4, hipMemcpyDeviceToHost));
  const float inv = 1.f / float(c->b_samples);
  for (size_t i = 0; i < n; ++i) {
    rgba[i * 4] = col[i * 3] * inv;
    rgba[i * 4 + 1] = col[i * 3 + 1] * inv;
    rgba[i * 4 + 2] = col[i * 3 + 2] * inv;
    const float a = al[i] * inv;
    rgba[i * 4 + 3] = a < 0.f ? 0.f : (a > 1.f ? 1.f : a);
  }
  return PHX_OK;
}
int phx_bridge_read_depth(phx_context *c, float *depth) {
  if (!c || !depth) return PHX_ERR_INVALID;
  if (c->b_samples == 0) return fail(c, PHX_ERR_NO_FRAME, "no frame has been rendered");
  const size_t n = size_t(c->b_w) * c->b_h;
  HIP_TRY(c, hipMemcpy(depth, c->b_depth.p, n * 4, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; ++i)
    if (!(depth[i] < 1e29f)) depth[i] = std::numeric_limits<float>::infinity();
  return PHX_OK;
}
int phx_bridge_read_normal(phx_context *c, float *xyz) {
  if (!c || !xyz) return PHX_ERR_INVALID;
  if (c->b_samples == 0) return fail(c, PHX_ERR_NO_FRAME, "no frame has been rendered");
  HIP_TRY(c, hipMemcpy(xyz, c->b_normal.p, size_t(c->b_w) * c->b_h * 12, hipMemcpyDeviceToHost));
  return PHX_OK;
}
int phx_bridge_read_albedo(phx_context *c, float *rgb) {
  if (!c || !rgb) return PHX_ERR_INVALID;
  if (c->b_samples == 0) return fail(c, PHX_ERR_NO_FRAME, "no frame has been rendered");
  HIP_TRY(c, hipMemcpy(rgb, c->b_albedo.p, size_t(c->b_w) * c->b_h * 12, hipMemcpyDeviceToHost));
  return PHX_OK;
}

}  // extern "C"
