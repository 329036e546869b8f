// Scalar FP32 vector math shared by the host and gfx950 device code.
//
// The hot path must reproduce the reference's f32 results, so every helper
// keeps the operation order of the glam 0.29 scalar Vec3 the reference uses
// (crates/pyre/src/math.rs): dot is (x*x' + y*y') + z*z', normalize is
// v * (1/len), min/max ignore a NaN operand (Rust f32::min/max). Everything
// is compiled with -ffp-contract=off, so no multiply-add is fused.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#if defined(__clang__)
#pragma clang fp contract(off)
#endif

#define PH __host__ __device__ __forceinline__

namespace pyre {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTau = 6.28318530717958647692f;
constexpr float kFrac1Pi = 0.318309886183790671538f;
constexpr float kFracPi2 = 1.57079632679489661923f;
constexpr float kFracPi4 = 0.785398163397448309616f;
constexpr float kInf = __builtin_huge_valf();

struct V3 {
  float x, y, z;
};
struct V2 {
  float x, y;
};

PH V3 v3(float x, float y, float z) { return V3{x, y, z}; }
PH V3 splat(float s) { return V3{s, s, s}; }
PH V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
PH V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
PH V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
PH V3 operator/(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }
PH V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
PH V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
PH V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
PH V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
PH bool veq(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
PH bool is_zero(V3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }

PH float rmin(float a, float b) { return fminf(a, b); }
PH float rmax(float a, float b) { return fmaxf(a, b); }
// Rust f32::clamp (keeps NaN).
PH float rclamp(float x, float lo, float hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}

PH float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PH V3 cross(V3 a, V3 b) { return V3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
PH float length_squared(V3 a) { return dot(a, a); }
PH float length(V3 a) { return sqrtf(dot(a, a)); }
PH V3 normalize(V3 a) { return a * (1.0f / length(a)); }
PH V3 normalize_or_zero(V3 a) {
  float rcp = 1.0f / length(a);
  if (__builtin_isfinite(rcp) && rcp > 0.0f) return a * rcp;
  return V3{0.0f, 0.0f, 0.0f};
}
PH V3 vmin(V3 a, V3 b) { return V3{rmin(a.x, b.x), rmin(a.y, b.y), rmin(a.z, b.z)}; }
PH V3 vmax(V3 a, V3 b) { return V3{rmax(a.x, b.x), rmax(a.y, b.y), rmax(a.z, b.z)}; }
PH float max_element(V3 a) { return rmax(a.x, rmax(a.y, a.z)); }
PH float min_element(V3 a) { return rmin(a.x, rmin(a.y, a.z)); }
PH V3 lerp(V3 a, V3 b, float s) { return a * (1.0f - s) + b * s; }
PH float axis_value(V3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// Rust `f as i32`: saturating, NaN -> 0.
PH int32_t f2i32(float f) {
  if (f != f) return 0;
  if (f >= 2147483648.0f) return 2147483647;
  if (f <= -2147483648.0f) return (-2147483647 - 1);
  return (int32_t)f;
}
PH uint32_t f2u32(float f) {
  if (!(f > 0.0f)) return 0u;
  if (f >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)f;
}

// math.rs:81-105
struct Frame {
  V3 t, bt, n;
};
PH Frame frame_from_normal(V3 n) {
  V3 helper = fabsf(n.x) > 0.9f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
  V3 t = normalize(cross(helper, n));
  V3 bt = cross(n, t);
  return Frame{t, bt, n};
}
PH V3 to_local(const Frame& f, V3 v) { return v3(dot(v, f.t), dot(v, f.bt), dot(v, f.n)); }
PH V3 to_world(const Frame& f, V3 v) { return v.x * f.t + v.y * f.bt + v.z * f.n; }

// sampler.rs:41-49
PH uint64_t pixel_seed(uint32_t x, uint32_t y, uint32_t sample) {
  uint64_t h = ((uint64_t)x << 32) | (uint64_t)y;
  h ^= (uint64_t)sample * 0x9E3779B97F4A7C15ull;
  h ^= h >> 30;
  h *= 0xBF58476D1CE4E5B9ull;
  h ^= h >> 27;
  h *= 0x94D049BB133111EBull;
  return h ^ (h >> 31);
}

// IndependentSampler (sampler.rs:16-36): xoshiro256++ seeded with four
// SplitMix64 outputs (rand_xoshiro 0.7 seed_from_u64); next_f32 takes the top
// 24 bits of the upper 32 (rand 0.9 StandardUniform for f32).
struct Rng {
  uint64_t s0, s1, s2, s3;
};
PH uint64_t splitmix_next(uint64_t& st) {
  st += 0x9E3779B97F4A7C15ull;
  uint64_t z = st;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
PH Rng rng_seed(uint64_t seed) {
  uint64_t st = seed;
  Rng r;
  r.s0 = splitmix_next(st);
  r.s1 = splitmix_next(st);
  r.s2 = splitmix_next(st);
  r.s3 = splitmix_next(st);
  return r;
}
PH uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
PH uint64_t rng_next_u64(Rng& r) {
  uint64_t result = rotl64(r.s0 + r.s3, 23) + r.s0;
  uint64_t t = r.s1 << 17;
  r.s2 ^= r.s0;
  r.s3 ^= r.s1;
  r.s1 ^= r.s2;
  r.s0 ^= r.s3;
  r.s2 ^= t;
  r.s3 = rotl64(r.s3, 45);
  return result;
}
PH float rng_f32(Rng& r) {
  uint32_t v = (uint32_t)(rng_next_u64(r) >> 32) >> 8;
  return (float)v * (1.0f / 16777216.0f);
}
PH V2 rng_vec2(Rng& r) {
  float a = rng_f32(r);
  float b = rng_f32(r);
  return V2{a, b};
}

// Moller-Trumbore on a precomputed (v0, e1 = v1 - v0, e2 = v2 - v0); the
// subtractions are the same f32 operations as mesh.rs:154-155, so the result
// is bit-identical to moller_trumbore (mesh.rs:153-178). Accepts
// t in (t_min, t_hi] when inclusive_hi, else (t_min, t_hi).
// Every quantity is computed and the reference's rejection tests are
// combined at the end (no early returns): a wave's lanes stay converged, and
// an accepted hit's t/u/v come from the same operations in the same order.
PH bool mt_intersect(V3 o, V3 d, V3 v0, V3 e1, V3 e2, float t_min, float t_hi, bool inclusive_hi, float* t_out,
                     float* u_out, float* v_out) {
  V3 h = cross(d, e2);
  float a = dot(e1, h);
  float f = 1.0f / a;
  V3 s = o - v0;
  float u = f * dot(s, h);
  V3 q = cross(s, e1);
  float v = f * dot(d, q);
  float t = f * dot(e2, q);
  const bool ok = !(fabsf(a) < 1e-8f) & (u >= 0.0f) & (u <= 1.0f) & !(v < 0.0f) & !(u + v > 1.0f) & (t > t_min) &
                  (inclusive_hi ? t <= t_hi : t < t_hi);
  *t_out = t;  // meaningful only when the test passes
  *u_out = u;
  *v_out = v;
  return ok;
}

}  // namespace pyre
