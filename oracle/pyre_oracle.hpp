// TEST INFRASTRUCTURE ONLY -- the oracle.
//
// A CPU restatement of the Pyre reference renderer (Rust, crates/pyre) used to
// check the HIP backend. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it; the product library never links or calls it.
//
// Every function keeps the reference's operation order so that, compiled with
// -O2 -ffp-contract=off -fno-fast-math on x86-64 (SSE, no FMA, glibc libm), it
// reproduces the Rust binary's f32 arithmetic. Each function cites the
// reference file:line it restates (paths relative to the reference repo root).
//
// Third-party arithmetic that is not vendored in the reference and is restated
// from its published algorithm:
//   * rand_xoshiro 0.7.0 (Cargo.lock:1817): Xoshiro256PlusPlus::seed_from_u64 =
//     four SplitMix64 outputs as the state words; next_u32 = next_u64 >> 32.
//   * rand 0.9.4 (Cargo.lock:1788): random::<f32>() = (next_u32 >> 8) * 2^-24.
//   * glam 0.29.3 (Cargo.lock:721): scalar Vec3; dot left-to-right,
//     normalize = v * (1/len), lerp = a*(1-s) + b*s, min/max NaN-ignoring.
// These are pinned end to end by the milestone-8 beauty PNG byte comparison
// in tests/test_oracle_fixtures.py.
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

namespace orc {

// Rust f32 constants (std::f32::consts), rounded to f32.
static const float PI = 3.14159265358979323846f;
static const float TAU = 6.28318530717958647692f;
static const float FRAC_1_PI = 0.318309886183790671538f;
static const float FRAC_PI_2 = 1.57079632679489661923f;
static const float FRAC_PI_4 = 0.785398163397448309616f;

// Rust f32::min / f32::max ignore a NaN operand, like C fminf/fmaxf.
static inline float rmin(float a, float b) { return std::fmin(a, b); }
static inline float rmax(float a, float b) { return std::fmax(a, b); }
// Rust f32::clamp keeps NaN (math.rs uses it via glam and std).
static inline float rclamp(float x, float lo, float hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
// Rust `f as i32` / `f as usize`: saturating, NaN -> 0.
static inline int32_t f2i32(float f) {
  if (f != f) return 0;
  if (f >= 2147483648.0f) return INT32_MAX;
  if (f <= -2147483648.0f) return INT32_MIN;
  return (int32_t)f;
}
static inline uint64_t f2usize(float f) {
  if (!(f > 0.0f)) return 0;
  if (f >= 18446744073709551616.0f) return UINT64_MAX;
  return (uint64_t)f;
}
static inline uint8_t f2u8(float f) {
  if (!(f > 0.0f)) return 0;
  if (f >= 255.0f) return 255;
  return (uint8_t)f;
}

// glam::Vec3 (scalar f32).
struct V3 {
  float x, y, z;
};
static inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
static inline V3 splat(float s) { return V3{s, s, s}; }
static inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline V3 operator/(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }
static inline V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
static inline V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
static inline V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
static inline V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
static inline bool operator==(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
static inline bool operator!=(V3 a, V3 b) { return !(a == b); }
static inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
static inline float length_squared(V3 a) { return dot(a, a); }
static inline float length(V3 a) { return std::sqrt(dot(a, a)); }
static inline V3 normalize(V3 a) { return a * (1.0f / length(a)); }
static inline V3 normalize_or_zero(V3 a) {
  float rcp = 1.0f / length(a);
  if (std::isfinite(rcp) && rcp > 0.0f) return a * rcp;
  return V3{0.0f, 0.0f, 0.0f};
}
static inline V3 vmin(V3 a, V3 b) { return V3{rmin(a.x, b.x), rmin(a.y, b.y), rmin(a.z, b.z)}; }
static inline V3 vmax(V3 a, V3 b) { return V3{rmax(a.x, b.x), rmax(a.y, b.y), rmax(a.z, b.z)}; }
static inline float max_element(V3 a) { return rmax(a.x, rmax(a.y, a.z)); }
static inline float min_element(V3 a) { return rmin(a.x, rmin(a.y, a.z)); }
static inline V3 lerp(V3 a, V3 b, float s) { return a * (1.0f - s) + b * s; }
static inline float axis_value(V3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

struct V2 {
  float x, y;
};

// math.rs:3-29
struct Ray {
  V3 o, d;
  float t_min, t_max, time;
};
static inline Ray ray_new(V3 o, V3 d) { return Ray{o, normalize(d), 1e-4f, INFINITY, 0.0f}; }
static inline V3 ray_at(const Ray& r, float t) { return r.o + r.d * t; }

// math.rs:31-77
struct Bounds3 {
  V3 min, max;
};
static inline Bounds3 bounds_empty() { return Bounds3{splat(INFINITY), splat(-INFINITY)}; }
static inline Bounds3 bunion(const Bounds3& a, const Bounds3& b) {
  return Bounds3{vmin(a.min, b.min), vmax(a.max, b.max)};
}
static inline Bounds3 bextend(const Bounds3& a, V3 p) { return Bounds3{vmin(a.min, p), vmax(a.max, p)}; }
static inline float surface_area(const Bounds3& b) {
  V3 d = b.max - b.min;
  return 2.0f * (d.x * d.y + d.x * d.z + d.y * d.z);
}
static inline V3 bcentroid(const Bounds3& b) { return 0.5f * (b.min + b.max); }

// math.rs:81-105
struct Frame {
  V3 t, bt, n;
};
static inline Frame frame_from_normal(V3 n) {
  V3 helper = std::fabs(n.x) > 0.9f ? v3(0, 1, 0) : v3(1, 0, 0);
  V3 t = normalize(cross(helper, n));
  V3 bt = cross(n, t);
  return Frame{t, bt, n};
}
static inline V3 to_local(const Frame& f, V3 v) { return v3(dot(v, f.t), dot(v, f.bt), dot(v, f.n)); }
static inline V3 to_world(const Frame& f, V3 v) { return v.x * f.t + v.y * f.bt + v.z * f.n; }

// sampler.rs:41-49
uint64_t pixel_seed(uint32_t x, uint32_t y, uint32_t sample);

// sampler.rs:16-36 over rand_xoshiro / rand (see header comment).
struct Sampler {
  uint64_t s[4];
  explicit Sampler(uint64_t seed);
  uint64_t next_u64();
  float next_f32() { return (float)((uint32_t)(next_u64() >> 32) >> 8) * (1.0f / 16777216.0f); }
  V2 next_vec2() {
    float a = next_f32();
    float b = next_f32();
    return V2{a, b};
  }
};

// distribution.rs:12-23
V2 concentric_disk(V2 u);

// distribution.rs:29-113
struct Distribution1D {
  std::vector<float> cdf;
  float integral = 0.0f;
  size_t n = 0;
  Distribution1D() = default;
  explicit Distribution1D(const float* w, size_t n);
  void sample_continuous(float u, float* t, float* pdf, size_t* bin) const;
  float pdf_at(float t) const;
};

// distribution.rs:119-172
struct Distribution2D {
  size_t width = 0, height = 0;
  std::vector<Distribution1D> conditionals;
  Distribution1D marginal;
  Distribution2D() = default;
  Distribution2D(const float* w, size_t width, size_t height);
  void sample_continuous(V2 s, V2* uv, float* pdf) const;
  float pdf_at(V2 uv) const;
};

// geometry/bvh.rs:18-30
struct BvhNode {
  Bounds3 bounds;
  uint32_t offset;
  uint16_t n_primitives;
  uint8_t axis;
  uint8_t pad;
};

struct BvhParams {
  uint32_t bins = 12;       // bvh.rs:13
  uint32_t max_leaf = 8;    // bvh.rs:14
  float traversal_cost = 1.0f;  // bvh.rs:15
  float intersect_cost = 2.0f;  // bvh.rs:16
};

struct TraversalStats {
  uint64_t nodes = 0;  // node fetches (stack pops)
  uint64_t tris = 0;   // Moller-Trumbore tests
};

// Work of the reference traversal counted per scene query while a render
// runs with statistics (orc_render_stats): extension queries
// (Scene::intersect) and shadow queries (Scene::occluded / occluded_dir).
struct RenderStats {
  uint64_t ext = 0, shadow = 0;
  TraversalStats ext_st, shadow_st;
  uint64_t ext_surface = 0;  // extension queries that hit a surface (kind 1)
  uint64_t first_hit = 0;    // samples whose camera ray hit a surface or a light
  bool first = false;        // the next extension query is a sample's camera ray
};
extern thread_local RenderStats* g_render_stats;

struct Bvh {
  std::vector<BvhNode> nodes;
  std::vector<uint32_t> prim;
  int max_depth = 0;
  void build(const std::vector<Bounds3>& pb, const std::vector<V3>& pc, const BvhParams& p);
};

// geometry/mesh.rs:5-41, 43-60
struct TriangleMesh {
  std::vector<V3> positions;
  std::vector<uint32_t> indices;
  std::vector<V3> normals;
  bool has_normals = false;
  size_t triangle_count() const { return indices.size() / 3; }
};

struct SurfaceInteraction {
  float t;
  V3 position, normal;
  uint32_t tri;  // surfaced for parity (the reference drops it, mesh.rs:124-142)
  float u, v;
};

struct MeshInstance {
  TriangleMesh mesh;
  Bvh bvh;
  bool has_motion = false;
  V3 off0{0, 0, 0}, off1{0, 0, 0};
  void build(const BvhParams& p);
  Bounds3 world_bounds() const;
  bool intersect(const Ray& ray, SurfaceInteraction* out, TraversalStats* st) const;
};

bool moller_trumbore(const Ray& r, V3 v0, V3 v1, V3 v2, float* t, float* u, float* v);

// material.rs
struct Material {
  int kind;  // 0 Lambertian, 1 Disney
  V3 base;   // Lambertian albedo or Disney base_color
  float metallic, roughness, specular;
};
struct BsdfSample {
  V3 wi, f;
  float pdf;
};
V3 bsdf_eval(const Material& m, V3 wo, V3 wi);
float bsdf_pdf(const Material& m, V3 wo, V3 wi);
bool bsdf_sample(const Material& m, V3 wo, V2 u, BsdfSample* out);
V3 bsdf_albedo(const Material& m);
V3 cosine_sample_hemisphere(V2 u);

// light.rs:46-162
struct QuadLight {
  V3 p0, eu, ev, cross_, normal, emission;
  float inv_cross_len_sq, area;
  QuadLight(V3 p0, V3 eu, V3 ev, V3 em);
  bool intersect_plane(V3 o, V3 d, float tmin, float tmax, float* t, V3* p) const;
  bool sample(V3 from, V2 u, V3* pos, V3* wi, float* dist, V3* li, float* pdf) const;
  float pdf(V3 from, V3 wi) const;
  V3 le(V3 pos, V3 n, V3 wo) const;
};

// Emissive triangle mesh behind the Light trait (light.rs:36-42): the
// general emitter that the reference's PBRT builder drops
// (io/pbrt/builder.rs:463-485,513-516). The reference has no implementation,
// so this states the design the device follows operation for operation:
//  - triangle i: v0, e1 = v1 - v0, e2 = v2 - v0 (mesh.rs:154-155),
//    normal = normalize(e1 x e2) (one-sided, emitting along +normal as
//    DiffuseAreaQuadLight does), weight 0.5 |e1 x e2|;
//  - dist: Distribution1D over the weights (distribution.rs:39-62); A = its
//    integral (the total area);
//  - sample(from, u): the triangle of sample_continuous's bin search on u.x
//    (distribution.rs:78-101), du = (u.x - cdf[i]) / max(cdf[i+1] - cdf[i],
//    1e-20) clamped to [0, 1], s = sqrt(du), point (v0 + (s (1 - u.y)) e1) +
//    (s u.y) e2 (uniform in area); then DiffuseAreaQuadLight::sample's steps
//    with pdf = dist^2 / (cos_light A) (light.rs:111-133);
//  - intersect: Moller-Trumbore (mesh.rs:153-178) over the triangles in
//    order, t in (t_min, t_max) with t_max tightened by each hit (the first
//    of equal t wins);
//  - pdf(from, wi): the hit of Ray{from, wi, 1e-4, inf}, t^2 / (cos_light A);
//  - le: emission when normal . wo > 0 (light.rs:146-152).
struct MeshLight {
  std::vector<V3> v0, v1, v2, e1, e2, normal;
  Distribution1D dist;
  float area = 0.0f;
  V3 emission;
  MeshLight(const float* pos, uint32_t nv, const uint32_t* idx, uint32_t nt, V3 em);
  bool intersect(V3 o, V3 d, float tmin, float tmax, float* t, uint32_t* tri) const;
  bool sample(V3 from, V2 u, V3* pos, V3* wi, float* dist, V3* li, float* pdf, uint32_t* tri = nullptr) const;
  float pdf(V3 from, V3 wi) const;
  V3 le(V3 pos, V3 n, V3 wo) const;
};

// light.rs:205-350
struct Hdri {
  uint32_t width = 0, height = 0;
  std::vector<V3> pixels;
  Distribution2D dist;
  float intensity = 1.0f;
  Hdri(uint32_t w, uint32_t h, std::vector<V3> px, float intensity);
  static V2 dir_to_uv(V3 wi);
  static V3 uv_to_dir(V2 uv);
  V3 sample_texture(V2 uv) const;
  V3 le(V3 wi) const;
  void sample(V2 u, V3* wi, V3* li, float* pdf) const;
  float pdf(V3 wi) const;
};
std::vector<V3> hdri_gradient_pixels(V3 zenith, V3 horizon, uint32_t* w, uint32_t* h);

// camera.rs
struct Camera {
  bool thin_lens = false;
  V3 origin, forward, right, up;
  float half_width, half_height, aperture_radius = 0.0f, focus_distance = 0.0f;
  void look_at(V3 origin, V3 target, V3 up, float vfov_deg, float aspect);
  Ray generate_ray(V2 ndc, V2 lens, float time) const;
};

// scene.rs
struct Scene {
  std::vector<MeshInstance> prims;
  std::vector<uint32_t> material_id;
  std::vector<Material> materials;
  std::vector<QuadLight> lights;
  std::vector<MeshLight> mesh_lights;  // light ids after the quads'
  Hdri* env = nullptr;
  ~Scene();
  // kind: 0 miss, 1 surface, 2 light
  int intersect(const Ray& ray, SurfaceInteraction* it, uint32_t* id) const;
  bool occluded_dir(V3 from, V3 dir, float time) const;
  bool occluded(V3 from, V3 to, float time) const;
  Bounds3 bounds() const;
};

// integrator.rs:14-29
struct PixelSample {
  V3 radiance, albedo, normal;
  float depth;
};
PixelSample integrate(const Scene& sc, Ray ray, Sampler& s, uint32_t max_depth, uint32_t min_rr);
float power_heuristic(float a, float b);

}  // namespace orc
