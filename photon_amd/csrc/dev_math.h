// Device math, RNG streams and sampling routines of the gfx950 path tracer.
//
// Semantics follow the reference's single-precision host code: math.h,
// rng.h:9-30, pathtracer.cpp:20-34, sampling.h:16-100 (paths relative to the
// reference repository root).  The library is compiled without FP contraction
// so products and sums round exactly as the reference's x86-64 build does.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PHX_DEV __device__ __forceinline__
#define PHX_HD __host__ __device__ __forceinline__

namespace phx {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 1.f / kPi;
constexpr float kTwoPi = 2.f * kPi;
constexpr float kInvTwoPi = 1.f / kTwoPi;

struct vec3 {
  float x, y, z;
};
struct vec2 {
  float x, y;
};

PHX_HD vec3 mk(float x, float y, float z) { return vec3{x, y, z}; }
PHX_HD vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
PHX_HD vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
PHX_HD vec3 operator-(vec3 a) { return {-a.x, -a.y, -a.z}; }
PHX_HD vec3 operator*(vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
PHX_HD vec3 operator*(float s, vec3 a) { return {a.x * s, a.y * s, a.z * s}; }
PHX_HD vec3 operator*(vec3 a, vec3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
PHX_HD float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PHX_HD vec3 cross(vec3 a, vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
PHX_DEV float length(vec3 v) { return sqrtf(dot(v, v)); }
// Reciprocal then multiply, as the reference normalizes (math.h:56-60).
PHX_DEV vec3 normalize(vec3 v) {
  const float len = length(v);
  return len > 0.f ? v * (1.f / len) : vec3{0.f, 0.f, 0.f};
}
PHX_DEV vec3 reflect(vec3 i, vec3 n) { return i - (2.f * dot(i, n)) * n; }
PHX_DEV vec3 refract(vec3 i, vec3 n, float eta) {
  const float cosi = -dot(i, n);
  const float k = 1.f - eta * eta * (1.f - cosi * cosi);
  if (k < 0.f) return {0.f, 0.f, 0.f};
  return eta * i + (eta * cosi - sqrtf(k)) * n;
}
PHX_DEV vec3 lerp(vec3 a, vec3 b, float t) { return a + (b - a) * t; }
PHX_DEV float luminance(vec3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
PHX_DEV float max3(vec3 v) { return fmaxf(v.x, fmaxf(v.y, v.z)); }
PHX_DEV bool finite3(vec3 v) { return v.x == v.x && v.y == v.y && v.z == v.z; }
PHX_DEV vec3 ld3(const float *p) { return {p[0], p[1], p[2]}; }
PHX_DEV float saturate(float x) { return fminf(1.f, fmaxf(0.f, x)); }
PHX_DEV float clamp01(float x) { return x < 0.f ? 0.f : (x > 1.f ? 1.f : x); }

// ---- RNG: xorshift32 (rng.h:15-29) keyed per (pixel, sample, dimension)
// through a 3-level integer hash (pathtracer.cpp:20-34).
struct Rng {
  uint32_t s;
  PHX_DEV uint32_t next_u32() {
    uint32_t x = s;
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    s = x;
    return x;
  }
  PHX_DEV float next() { return float(next_u32() >> 8) * (1.0f / 16777216.0f); }
};
PHX_HD uint32_t hash_u32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}
PHX_DEV Rng make_rng(uint32_t pixel, uint32_t sample, uint32_t dim) {
  const uint32_t h = hash_u32(pixel + hash_u32(sample + hash_u32(dim + 0x9e3779b9U)));
  return Rng{h ? h : 1u};
}

// ---- sampling (sampling.h) --------------------------------------------------
PHX_DEV float power_heuristic(float f, float g) {
  const float f2 = f * f, g2 = g * g;
  return (f2 + g2) > 0.f ? f2 / (f2 + g2) : 0.f;
}
PHX_DEV void onb(vec3 n, vec3 &t, vec3 &b) {
  if (n.z < -0.9999999f) {
    t = {0.f, -1.f, 0.f};
    b = {-1.f, 0.f, 0.f};
    return;
  }
  const float a = 1.f / (1.f + n.z);
  const float bb = -n.x * n.y * a;
  t = normalize(vec3{1.f - n.x * n.x * a, bb, -n.x});
  b = normalize(vec3{bb, 1.f - n.y * n.y * a, -n.y});
}
// Sampled directions use the hardware sine and cosine (__sinf / __cosf,
// about 1e-6 absolute over one turn) instead of the correctly rounded libm:
// a sampled direction has to follow the stream's numbers to the parity
// tolerance, not to the last ulp, and the libm pair was a measurable share of
// the shade kernel.
PHX_DEV vec2 concentric_disk(Rng &rng) {
  const float u = 2.f * rng.next() - 1.f;
  const float v = 2.f * rng.next() - 1.f;
  if (u == 0.f && v == 0.f) return {0.f, 0.f};
  float r, th;
  if (fabsf(u) > fabsf(v)) {
    r = fabsf(u);
    th = (kPi / 4.f) * (v / u);
  } else {
    r = fabsf(v);
    th = (kPi / 2.f) - (kPi / 4.f) * (u / v);
  }
  return {r * __cosf(th), r * __sinf(th)};
}
PHX_DEV vec3 cosine_hemisphere(vec3 n, Rng &rng) {
  vec3 t, b;
  onb(n, t, b);
  const vec2 d = concentric_disk(rng);
  const float z = sqrtf(fmaxf(0.f, 1.f - d.x * d.x - d.y * d.y));
  return normalize(d.x * t + d.y * b + z * n);
}
PHX_DEV vec3 uniform_sphere(Rng &rng) {
  const float z = 1.f - 2.f * rng.next();
  const float r = sqrtf(fmaxf(0.f, 1.f - z * z));
  const float phi = kTwoPi * rng.next();
  return {r * __cosf(phi), r * __sinf(phi), z};
}

}  // namespace phx
