// Seam A kernels: RayBackend::trace_closest / trace_occluded on caller-owned
// device arrays (kokkos_backend.cpp:21-86).  Rays are RayBatch-shaped: 3-float
// origins and directions, tmin, tmax; inactive iff !(tmax > tmin).
#include "dev_queue.h"
#include "dev_trace.h"
#include "integrator.h"

namespace phx {

constexpr uint32_t kRefill = 16;  // idle lanes that make a wave take new rays

PHX_DEV vec3 ldv(const float *p, uint32_t i) { return {p[i * 3], p[i * 3 + 1], p[i * 3 + 2]}; }

template <int MODE>  // 0: compact closest, 1: full HitResult, 2: occluded flag
// `spill_stride` ints of stack spill per thread: at least the entries the
// scene's BVH can push beyond the LDS stack ((kNodeWidth - 1) * depth + 1 -
// kShortStack).
__global__ __launch_bounds__(256) void k_trace(DevScene s, const float *org, const float *dir, const float *tmin,
                                               const float *tmax, uint32_t n, void *out, uint32_t *fetch,
                                               int *spill, uint32_t spill_stride) {
  __shared__ int lds[kShortStack * 256];
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  LaneStack st{&lds[threadIdx.x], 256, spill + size_t(gtid) * spill_stride, 0};
  traverse_stream<MODE == 2, true>(
      st, fetch, n, kChunk, kRefill,
      [&](uint32_t i, vec3 &o, vec3 &d, float &t0, float &t1, const BvhNode *&nodes, const TriRec *&tris) {
        o = ldv(org, i);
        d = ldv(dir, i);
        t0 = tmin[i];
        t1 = tmax[i];
        nodes = s.nodes;
        tris = s.tris;
      },
      [&](uint32_t i, bool hit, const TraceHit &h) {
        if (MODE == 0) {
          phx_hit_compact r;
          r.t = hit ? h.t : 0.f;
          r.prim_id = hit ? h.prim : 0xffffffffu;
          r.b1 = hit ? h.u : 0.f;
          r.b2 = hit ? h.v : 0.f;
          static_cast<phx_hit_compact *>(out)[i] = r;
        } else if (MODE == 2) {
          static_cast<uint32_t *>(out)[i] = hit ? 1u : 0u;
        } else {
          // Post-hit attributes (mesh_intersector.h:126-178).
          phx_hit r{};
          r.hit = hit ? 1u : 0u;
          if (hit) {
            const vec3 o = ldv(org, i), d = ldv(dir, i);
            const TriRec tr = s.tris[h.tri];
            const uint32_t prim = h.prim;
            const vec3 p = o + d * h.t;
            const vec3 nn = normalize(cross(vec3{tr.e1.x, tr.e1.y, tr.e1.z}, vec3{tr.e2.x, tr.e2.y, tr.e2.z}));
            const uint32_t i0 = s.indices[prim * 3], i1 = s.indices[prim * 3 + 1], i2 = s.indices[prim * 3 + 2];
            const float w0 = 1.f - h.u - h.v, w1 = h.u, w2 = h.v;
            vec3 sn = nn;
            if (s.normals) {
              sn = normalize(ldv(s.normals, i0) * w0 + ldv(s.normals, i1) * w1 + ldv(s.normals, i2) * w2);
              if (dot(sn, nn) < 0.f) sn = sn * -1.f;
            }
            r.t = h.t;
            r.position[0] = p.x; r.position[1] = p.y; r.position[2] = p.z;
            r.normal[0] = nn.x; r.normal[1] = nn.y; r.normal[2] = nn.z;
            r.shading_normal[0] = sn.x; r.shading_normal[1] = sn.y; r.shading_normal[2] = sn.z;
            if (s.texcoords) {
              const float *a = s.texcoords + size_t(i0) * 2, *b = s.texcoords + size_t(i1) * 2,
                          *c = s.texcoords + size_t(i2) * 2;
              r.uv[0] = a[0] * w0 + b[0] * w1 + c[0] * w2;
              r.uv[1] = a[1] * w0 + b[1] * w1 + c[1] * w2;
            }
            r.bary[0] = h.u;
            r.bary[1] = h.v;
            if (s.colors && s.colors[size_t(i0) * 3] >= 0.f) {
              const vec3 c = ldv(s.colors, i0) * w0 + ldv(s.colors, i1) * w1 + ldv(s.colors, i2) * w2;
              r.interpolated_color[0] = c.x; r.interpolated_color[1] = c.y; r.interpolated_color[2] = c.z;
              r.has_interpolated_color = 1u;
            }
            r.prim_id = prim;
            r.material_id = s.material_ids ? s.material_ids[prim] : 0u;
          }
          static_cast<phx_hit *>(out)[i] = r;
        }
      },
      s.nodes, s.tris);
}

// The scene's per-triangle normal records (DevScene::tri_ng), in leaf order.
__global__ void k_tri_normals(const TriRec *tris, float4 *out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const TriRec r = tris[i];
  const vec3 ng = tri_normal(r);
  out[i] = make_float4(ng.x, ng.y, ng.z, r.v0.w);
}

__global__ void k_clear_u32(uint32_t *p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

hipError_t launch_tri_normals(const TriRec *tris, float4 *out, uint32_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tri_normals, dim3((n + 255u) / 256u), dim3(256), 0, st, tris, out, n);
  return hipGetLastError();
}
hipError_t launch_clear_u32(uint32_t *p, uint32_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_clear_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
  return hipGetLastError();
}
hipError_t launch_trace_compact(const DevScene &s, const float *org, const float *dir, const float *tmin,
                                const float *tmax, uint32_t n, phx_hit_compact *out, uint32_t *fetch, int *spill,
                                uint32_t grid, hipStream_t st, uint32_t spill_stride) {
  hipLaunchKernelGGL(k_trace<0>, dim3(grid), dim3(256), 0, st, s, org, dir, tmin, tmax, n, (void *)out, fetch, spill,
                     spill_stride);
  return hipGetLastError();
}
hipError_t launch_trace_full(const DevScene &s, const float *org, const float *dir, const float *tmin,
                             const float *tmax, uint32_t n, phx_hit *out, uint32_t *fetch, int *spill, uint32_t grid,
                             hipStream_t st, uint32_t spill_stride) {
  hipLaunchKernelGGL(k_trace<1>, dim3(grid), dim3(256), 0, st, s, org, dir, tmin, tmax, n, (void *)out, fetch, spill,
                     spill_stride);
  return hipGetLastError();
}
hipError_t launch_trace_occluded(const DevScene &s, const float *org, const float *dir, const float *tmin,
                                 const float *tmax, uint32_t n, uint32_t *out, uint32_t *fetch, int *spill,
                                 uint32_t grid, hipStream_t st, uint32_t spill_stride) {
  hipLaunchKernelGGL(k_trace<2>, dim3(grid), dim3(256), 0, st, s, org, dir, tmin, tmax, n, (void *)out, fetch, spill,
                     spill_stride);
  return hipGetLastError();
}

}  // namespace phx
