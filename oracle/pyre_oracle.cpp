// TEST INFRASTRUCTURE ONLY -- the oracle (see pyre_oracle.hpp).
// CPU restatement of crates/pyre/src/*.rs and crates/pyre-cli/src/main.rs.
#include "pyre_oracle.hpp"

#include <algorithm>
#include <atomic>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <thread>

namespace orc {

// ---------------------------------------------------------------- sampler.rs
// sampler.rs:41-49
uint64_t pixel_seed(uint32_t x, uint32_t y, uint32_t sample) {
  uint64_t h = ((uint64_t)x << 32) | (uint64_t)y;
  h ^= (uint64_t)sample * 0x9E3779B97F4A7C15ull;
  h ^= h >> 30;
  h *= 0xBF58476D1CE4E5B9ull;
  h ^= h >> 27;
  h *= 0x94D049BB133111EBull;
  return h ^ (h >> 31);
}

// rand_xoshiro 0.7 Xoshiro256PlusPlus::seed_from_u64: SplitMix64 fills the
// 32-byte seed with four little-endian outputs (Blackman & Vigna).
Sampler::Sampler(uint64_t seed) {
  uint64_t st = seed;
  for (int i = 0; i < 4; ++i) {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    s[i] = z ^ (z >> 31);
  }
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// xoshiro256++ next (Blackman & Vigna 2019).
uint64_t Sampler::next_u64() {
  uint64_t result = rotl(s[0] + s[3], 23) + s[0];
  uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return result;
}

// ----------------------------------------------------------- distribution.rs
// distribution.rs:12-23
V2 concentric_disk(V2 u) {
  V2 off{2.0f * u.x - 1.0f, 2.0f * u.y - 1.0f};
  if (off.x == 0.0f && off.y == 0.0f) return V2{0.0f, 0.0f};
  float r, theta;
  if (std::fabs(off.x) > std::fabs(off.y)) {
    r = off.x;
    theta = FRAC_PI_4 * (off.y / off.x);
  } else {
    r = off.y;
    theta = FRAC_PI_2 - FRAC_PI_4 * (off.x / off.y);
  }
  return V2{r * std::cos(theta), r * std::sin(theta)};
}

// distribution.rs:39-62
Distribution1D::Distribution1D(const float* w, size_t n_) : n(n_) {
  cdf.reserve(n + 1);
  cdf.push_back(0.0f);
  float acc = 0.0f;
  for (size_t i = 0; i < n; ++i) {
    acc += rmax(w[i], 0.0f);
    cdf.push_back(acc);
  }
  integral = acc;
  if (integral > 0.0f) {
    for (auto& c : cdf) c /= integral;
  } else {
    for (size_t i = 0; i < cdf.size(); ++i) cdf[i] = (float)i / (float)n;
  }
}

// distribution.rs:78-101
void Distribution1D::sample_continuous(float u, float* t, float* pdf, size_t* bin_out) const {
  size_t lo = 1, hi = cdf.size() - 1;
  while (lo < hi) {
    size_t mid = (lo + hi) / 2;
    if (cdf[mid] <= u)
      lo = mid + 1;
    else
      hi = mid;
  }
  size_t bin = std::min(lo - 1, n - 1);
  float cdf_lo = cdf[bin];
  float cdf_hi = cdf[bin + 1];
  float bin_size = rmax(cdf_hi - cdf_lo, 1e-20f);
  float du = (u - cdf_lo) / bin_size;
  *t = ((float)bin + du) / (float)n;
  *pdf = bin_size * (float)n;
  *bin_out = bin;
}

// distribution.rs:104-112
float Distribution1D::pdf_at(float t) const {
  if (t < 0.0f || t >= 1.0f || n == 0) return 0.0f;
  size_t bin = std::min((size_t)f2usize(t * (float)n), n - 1);
  return (cdf[bin + 1] - cdf[bin]) * (float)n;
}

// distribution.rs:130-147
Distribution2D::Distribution2D(const float* w, size_t width_, size_t height_)
    : width(width_), height(height_) {
  std::vector<float> row_integrals;
  conditionals.reserve(height);
  for (size_t y = 0; y < height; ++y) {
    conditionals.emplace_back(w + y * width, width);
    row_integrals.push_back(conditionals.back().integral);
  }
  marginal = Distribution1D(row_integrals.data(), height);
}

// distribution.rs:151-155
void Distribution2D::sample_continuous(V2 s, V2* uv, float* pdf) const {
  float v, pdf_v, u, pdf_u;
  size_t row, col;
  marginal.sample_continuous(s.y, &v, &pdf_v, &row);
  conditionals[row].sample_continuous(s.x, &u, &pdf_u, &col);
  *uv = V2{u, v};
  *pdf = pdf_u * pdf_v;
}

// distribution.rs:157-163
float Distribution2D::pdf_at(V2 uv) const {
  if (uv.x < 0.0f || uv.x >= 1.0f || uv.y < 0.0f || uv.y >= 1.0f) return 0.0f;
  size_t row = std::min((size_t)f2usize(uv.y * (float)height), height - 1);
  return conditionals[row].pdf_at(uv.x) * marginal.pdf_at(uv.y);
}

// ------------------------------------------------------------ geometry/bvh.rs
static int max_axis(V3 v) {  // bvh.rs:164-172
  if (v.x >= v.y && v.x >= v.z) return 0;
  if (v.y >= v.z) return 1;
  return 2;
}

struct BuildCtx {
  std::vector<BvhNode>* nodes;
  std::vector<uint32_t>* idx;
  const std::vector<Bounds3>* pb;
  const std::vector<V3>* pc;
  BvhParams p;
  int max_depth;
};

static uint32_t push_leaf(BuildCtx& c, Bounds3 b, uint32_t start, uint16_t n) {  // bvh.rs:348-358
  uint32_t i = (uint32_t)c.nodes->size();
  c.nodes->push_back(BvhNode{b, start, n, 0, 0});
  return i;
}

static uint32_t build_recursive(BuildCtx& c, uint32_t start, uint32_t end, int depth);

static uint32_t push_split(BuildCtx& c, Bounds3 b, uint8_t axis, uint32_t start, uint32_t mid,
                           uint32_t end, int depth) {  // bvh.rs:360-387
  uint32_t node_idx = (uint32_t)c.nodes->size();
  c.nodes->push_back(BvhNode{b, 0, 0, axis, 0});
  build_recursive(c, start, mid, depth + 1);
  uint32_t right_idx = (uint32_t)c.nodes->size();
  build_recursive(c, mid, end, depth + 1);
  (*c.nodes)[node_idx].offset = right_idx;
  return node_idx;
}

static uint32_t partition_by_axis(BuildCtx& c, uint32_t start, uint32_t end, int axis,
                                  float split_pos) {  // bvh.rs:389-409
  auto& idx = *c.idx;
  size_t left = start, right = end;
  while (left < right) {
    float cv = axis_value((*c.pc)[idx[left]], axis);
    if (cv < split_pos) {
      left += 1;
    } else {
      right -= 1;
      std::swap(idx[left], idx[right]);
    }
  }
  return (uint32_t)left;
}

struct Bin {
  uint32_t count;
  Bounds3 bounds;
};

// bvh.rs:180-346
static uint32_t build_recursive(BuildCtx& c, uint32_t start, uint32_t end, int depth) {
  if (depth > c.max_depth) c.max_depth = depth;
  size_t n = end - start;
  const auto& pb = *c.pb;
  const auto& pc = *c.pc;
  auto& idx = *c.idx;
  const uint32_t BINS = c.p.bins;
  Bounds3 bounds = bounds_empty(), cb = bounds_empty();
  for (uint32_t k = start; k < end; ++k) {
    bounds = bunion(bounds, pb[idx[k]]);
    cb = bextend(cb, pc[idx[k]]);
  }
  if (n == 1) return push_leaf(c, bounds, start, 1);
  V3 cdiag = cb.max - cb.min;
  int axis = max_axis(cdiag);
  float extent = axis_value(cdiag, axis);
  if (extent < 1e-9f) return push_leaf(c, bounds, start, (uint16_t)std::min(n, (size_t)65535));

  std::vector<Bin> bins(BINS, Bin{0, bounds_empty()});
  float centroid_min = axis_value(cb.min, axis);
  float scale = (float)BINS / extent;
  for (uint32_t k = start; k < end; ++k) {
    float cv = axis_value(pc[idx[k]], axis);
    int32_t b = f2i32((cv - centroid_min) * scale);
    if (b < 0) b = 0;
    if (b >= (int32_t)BINS) b = (int32_t)BINS - 1;
    bins[b].count += 1;
    bins[b].bounds = bunion(bins[b].bounds, pb[idx[k]]);
  }
  std::vector<uint32_t> lc(BINS - 1), rc(BINS - 1);
  std::vector<Bounds3> lb(BINS - 1), rb(BINS - 1);
  uint32_t acc_c = 0;
  Bounds3 acc_b = bounds_empty();
  for (uint32_t i = 0; i < BINS - 1; ++i) {
    acc_c += bins[i].count;
    acc_b = bunion(acc_b, bins[i].bounds);
    lc[i] = acc_c;
    lb[i] = acc_b;
  }
  acc_c = 0;
  acc_b = bounds_empty();
  for (int i = (int)BINS - 2; i >= 0; --i) {
    acc_c += bins[i + 1].count;
    acc_b = bunion(acc_b, bins[i + 1].bounds);
    rc[i] = acc_c;
    rb[i] = acc_b;
  }
  float parent_area = rmax(surface_area(bounds), 1e-9f);
  float best_cost = INFINITY;
  uint32_t best_split = 0;
  for (uint32_t i = 0; i < BINS - 1; ++i) {
    if (lc[i] == 0 || rc[i] == 0) continue;
    float cost = c.p.traversal_cost +
                 c.p.intersect_cost *
                     ((float)lc[i] * surface_area(lb[i]) + (float)rc[i] * surface_area(rb[i])) /
                     parent_area;
    if (cost < best_cost) {
      best_cost = cost;
      best_split = i;
    }
  }
  float leaf_cost = c.p.intersect_cost * (float)n;
  if (!std::isfinite(best_cost)) {
    float mid_pos = centroid_min + 0.5f * extent;
    uint32_t mid = partition_by_axis(c, start, end, axis, mid_pos);
    if (mid != start && mid != end) return push_split(c, bounds, (uint8_t)axis, start, mid, end, depth);
    if (n <= c.p.max_leaf) return push_leaf(c, bounds, start, (uint16_t)n);
    mid = start + (uint32_t)n / 2;
    return push_split(c, bounds, (uint8_t)axis, start, mid, end, depth);
  }
  if (n <= c.p.max_leaf && best_cost >= leaf_cost) return push_leaf(c, bounds, start, (uint16_t)n);
  float split_pos = centroid_min + (float)(best_split + 1) * extent / (float)BINS;
  uint32_t mid = partition_by_axis(c, start, end, axis, split_pos);
  if (mid == start || mid == end) {
    if (n <= c.p.max_leaf) return push_leaf(c, bounds, start, (uint16_t)n);
    mid = start + (uint32_t)n / 2;
    return push_split(c, bounds, (uint8_t)axis, start, mid, end, depth);
  }
  return push_split(c, bounds, (uint8_t)axis, start, mid, end, depth);
}

// bvh.rs:45-64
void Bvh::build(const std::vector<Bounds3>& pb, const std::vector<V3>& pc, const BvhParams& p) {
  size_t n = pb.size();
  prim.resize(n);
  for (size_t i = 0; i < n; ++i) prim[i] = (uint32_t)i;
  nodes.clear();
  nodes.reserve(2 * std::max(n, (size_t)1));
  BuildCtx c{&nodes, &prim, &pb, &pc, p, 0};
  if (n > 0) build_recursive(c, 0, (uint32_t)n, 0);
  max_depth = c.max_depth;
}

// bvh.rs:146-153
static inline bool ray_aabb(const Ray& r, const Bounds3& b, V3 inv_dir, float t_max) {
  V3 t1 = (b.min - r.o) * inv_dir;
  V3 t2 = (b.max - r.o) * inv_dir;
  float tmin = rmax(max_element(vmin(t1, t2)), r.t_min);
  float tmax = rmin(min_element(vmax(t1, t2)), t_max);
  return tmin <= tmax;
}

// ----------------------------------------------------------- geometry/mesh.rs
// mesh.rs:153-178
bool moller_trumbore(const Ray& r, V3 v0, V3 v1, V3 v2, float* t_out, float* u_out, float* v_out) {
  V3 e1 = v1 - v0;
  V3 e2 = v2 - v0;
  V3 h = cross(r.d, e2);
  float a = dot(e1, h);
  if (std::fabs(a) < 1e-8f) return false;
  float f = 1.0f / a;
  V3 s = r.o - v0;
  float u = f * dot(s, h);
  if (!(u >= 0.0f && u <= 1.0f)) return false;
  V3 q = cross(s, e1);
  float v = f * dot(r.d, q);
  if (v < 0.0f || u + v > 1.0f) return false;
  float t = f * dot(e2, q);
  if (t > r.t_min && t < r.t_max) {
    *t_out = t;
    *u_out = u;
    *v_out = v;
    return true;
  }
  return false;
}

// mesh.rs:77-87
void MeshInstance::build(const BvhParams& p) {
  size_t n = mesh.triangle_count();
  std::vector<Bounds3> b(n);
  std::vector<V3> c(n);
  for (size_t t = 0; t < n; ++t) {
    V3 v0 = mesh.positions[mesh.indices[3 * t]];
    V3 v1 = mesh.positions[mesh.indices[3 * t + 1]];
    V3 v2 = mesh.positions[mesh.indices[3 * t + 2]];
    b[t] = bextend(bextend(Bounds3{v0, v0}, v1), v2);  // mesh.rs:32-35
    c[t] = (v0 + v1 + v2) / 3.0f;                      // mesh.rs:37-40
  }
  bvh.build(b, c, p);
}

// mesh.rs:97-106
Bounds3 MeshInstance::world_bounds() const {
  Bounds3 base = bvh.nodes.empty() ? bounds_empty() : bvh.nodes[0].bounds;
  if (!has_motion) return base;
  Bounds3 b0{base.min + off0, base.max + off0};
  Bounds3 b1{base.min + off1, base.max + off1};
  return bunion(b0, b1);
}

// mesh.rs:109-148 wrapped around bvh.rs:86-143 (the visit closure inlined).
bool MeshInstance::intersect(const Ray& ray, SurfaceInteraction* out, TraversalStats* st) const {
  V3 offset{0, 0, 0};
  if (has_motion) offset = lerp(off0, off1, rclamp(ray.time, 0.0f, 1.0f));
  Ray lr = ray;
  bool moved = offset != V3{0, 0, 0};
  if (moved) lr.o = ray.o - offset;
  if (bvh.nodes.empty()) return false;
  V3 inv_dir = V3{1.0f, 1.0f, 1.0f} / lr.d;
  bool dir_neg[3] = {inv_dir.x < 0.0f, inv_dir.y < 0.0f, inv_dir.z < 0.0f};
  bool found = false;
  SurfaceInteraction closest{};
  float t_max = lr.t_max;
  uint32_t stack[64];
  stack[0] = 0;
  size_t sp = 1;
  while (sp > 0) {
    sp -= 1;
    uint32_t ni = stack[sp];
    const BvhNode& node = bvh.nodes[ni];
    if (st) st->nodes++;
    if (!ray_aabb(lr, node.bounds, inv_dir, t_max)) continue;
    if (node.n_primitives > 0) {
      for (uint32_t k = node.offset; k < node.offset + node.n_primitives; ++k) {
        uint32_t tri = bvh.prim[k];
        Ray tr = lr;
        tr.t_max = t_max;
        uint32_t a = mesh.indices[3 * tri], b = mesh.indices[3 * tri + 1], c = mesh.indices[3 * tri + 2];
        V3 v0 = mesh.positions[a], v1 = mesh.positions[b], v2 = mesh.positions[c];
        float t, u, v;
        if (st) st->tris++;
        if (!moller_trumbore(tr, v0, v1, v2, &t, &u, &v)) continue;
        SurfaceInteraction hit;
        hit.t = t;
        hit.position = ray_at(tr, t);
        if (mesh.has_normals) {
          float w = 1.0f - u - v;
          hit.normal = normalize(w * mesh.normals[a] + u * mesh.normals[b] + v * mesh.normals[c]);
        } else {
          hit.normal = normalize(cross(v1 - v0, v2 - v0));
        }
        hit.tri = tri;
        hit.u = u;
        hit.v = v;
        if (hit.t >= lr.t_min && hit.t < t_max) {
          t_max = hit.t;
          closest = hit;
          found = true;
        }
      }
    } else {
      if (sp + 2 > 64) {
        // The reference indexes a fixed [u32; 64] stack and panics here.
        std::fprintf(stderr, "oracle: BVH traversal stack overflow\n");
        std::abort();
      }
      uint32_t left = ni + 1, right = node.offset;
      if (dir_neg[node.axis]) {
        stack[sp++] = left;
        stack[sp++] = right;
      } else {
        stack[sp++] = right;
        stack[sp++] = left;
      }
    }
  }
  if (!found) return false;
  if (moved) closest.position = closest.position + offset;
  *out = closest;
  return true;
}

// --------------------------------------------------------------- material.rs
static inline float luminance(V3 c) { return dot(c, v3(0.2126f, 0.7152f, 0.0722f)); }  // :232-235
static inline float powi5(float x) { return x * ((x * x) * (x * x)); }  // llvm.powi(x, 5) expansion

V3 cosine_sample_hemisphere(V2 u) {  // material.rs:239-246
  float r = std::sqrt(u.x);
  float phi = 2.0f * PI * u.y;
  float x = r * std::cos(phi);
  float y = r * std::sin(phi);
  float z = std::sqrt(rmax(1.0f - u.x, 0.0f));
  return v3(x, y, z);
}

static float ggx_d(V3 h, float alpha) {  // :249-256
  if (h.z <= 0.0f) return 0.0f;
  float a2 = alpha * alpha;
  float denom = h.z * h.z * (a2 - 1.0f) + 1.0f;
  return a2 / (PI * denom * denom);
}

static V3 schlick_fresnel(float cos_d, V3 f0) {  // :259-262
  float c = rclamp(cos_d, 0.0f, 1.0f);
  return f0 + (splat(1.0f) - f0) * powi5(1.0f - c);
}

static float ggx_lambda(V3 v, float alpha) {  // :264-272
  float cos2 = v.z * v.z;
  if (cos2 >= 1.0f) return 0.0f;
  float a2 = alpha * alpha;
  float tan2 = (1.0f - cos2) / rmax(cos2, 1e-8f);
  return 0.5f * (-1.0f + std::sqrt(1.0f + a2 * tan2));
}

static float ggx_g1(V3 v, float alpha) { return 1.0f / (1.0f + ggx_lambda(v, alpha)); }  // :274-276
static float ggx_g2(V3 wo, V3 wi, float alpha) {  // :279-281
  return 1.0f / (1.0f + ggx_lambda(wo, alpha) + ggx_lambda(wi, alpha));
}
static V3 reflect(V3 v, V3 n) { return 2.0f * dot(v, n) * n - v; }  // :285-287

static V3 sample_ggx_vndf(V3 wo, float alpha, V2 u) {  // :291-317
  V3 vh = normalize(v3(alpha * wo.x, alpha * wo.y, wo.z));
  float lensq = vh.x * vh.x + vh.y * vh.y;
  V3 t1 = lensq > 0.0f ? v3(-vh.y, vh.x, 0.0f) / std::sqrt(lensq) : v3(1, 0, 0);
  V3 t2 = cross(vh, t1);
  float r = std::sqrt(u.x);
  float phi = 2.0f * PI * u.y;
  float t1c = r * std::cos(phi);
  float t2c = r * std::sin(phi);
  float s = 0.5f * (1.0f + vh.z);
  t2c = (1.0f - s) * std::sqrt(rmax(1.0f - t1c * t1c, 0.0f)) + s * t2c;
  V3 nh = t1c * t1 + t2c * t2 + std::sqrt(rmax(1.0f - t1c * t1c - t2c * t2c, 0.0f)) * vh;
  return normalize(v3(alpha * nh.x, alpha * nh.y, rmax(nh.z, 0.0f)));
}

static float disney_alpha(const Material& m) { return rmax(m.roughness * m.roughness, 1e-3f); }  // :97-102
static V3 disney_f0(const Material& m) { return lerp(splat(0.08f * m.specular), m.base, m.metallic); }  // :104-107
static void disney_lobe_weights(const Material& m, float* pd, float* ps) {  // :112-121
  float diff = (1.0f - m.metallic) * luminance(m.base);
  float spec = rmax(luminance(disney_f0(m)), 0.04f);
  float total = diff + spec;
  if (total > 0.0f) {
    *pd = diff / total;
    *ps = spec / total;
  } else {
    *pd = 1.0f;
    *ps = 0.0f;
  }
}
static V3 disney_eval_diffuse(const Material& m, V3 wo, V3 wi) {  // :125-138
  if (wo.z <= 0.0f || wi.z <= 0.0f || m.metallic >= 1.0f) return splat(0.0f);
  V3 h = normalize_or_zero(wo + wi);
  if (length_squared(h) < 1e-10f) return splat(0.0f);
  float cos_d = rclamp(dot(wi, h), 0.0f, 1.0f);
  float fd90 = 0.5f + 2.0f * m.roughness * cos_d * cos_d;
  float fd_v = 1.0f + (fd90 - 1.0f) * powi5(rmax(1.0f - wo.z, 0.0f));
  float fd_l = 1.0f + (fd90 - 1.0f) * powi5(rmax(1.0f - wi.z, 0.0f));
  return m.base * (1.0f - m.metallic) / PI * fd_v * fd_l;
}
static V3 disney_eval_specular(const Material& m, V3 wo, V3 wi) {  // :142-155
  if (wo.z <= 0.0f || wi.z <= 0.0f) return splat(0.0f);
  V3 h = normalize_or_zero(wo + wi);
  if (length_squared(h) < 1e-10f || h.z <= 0.0f) return splat(0.0f);
  float alpha = disney_alpha(m);
  float d = ggx_d(h, alpha);
  V3 f = schlick_fresnel(rmax(dot(wi, h), 0.0f), disney_f0(m));
  float g = ggx_g2(wo, wi, alpha);
  return d * f * g / (4.0f * wo.z * wi.z);
}
static float disney_pdf_diffuse(V3 wi) { return wi.z <= 0.0f ? 0.0f : wi.z / PI; }  // :157-163
static float disney_pdf_specular(const Material& m, V3 wo, V3 wi) {  // :168-180
  if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
  V3 h = normalize_or_zero(wo + wi);
  if (length_squared(h) < 1e-10f || h.z <= 0.0f) return 0.0f;
  float alpha = disney_alpha(m);
  float d = ggx_d(h, alpha);
  float g1 = ggx_g1(wo, alpha);
  return g1 * d / (4.0f * wo.z);
}

V3 bsdf_albedo(const Material& m) { return m.base; }  // :48-50, :184-186

V3 bsdf_eval(const Material& m, V3 wo, V3 wi) {
  if (m.kind == 0) {  // :52-58
    if (wo.z <= 0.0f || wi.z <= 0.0f) return splat(0.0f);
    return m.base / PI;
  }
  return disney_eval_diffuse(m, wo, wi) + disney_eval_specular(m, wo, wi);  // :188-190
}

float bsdf_pdf(const Material& m, V3 wo, V3 wi) {
  if (m.kind == 0) {  // :60-66
    if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
    return wi.z / PI;
  }
  float pd, ps;  // :192-195
  disney_lobe_weights(m, &pd, &ps);
  return pd * disney_pdf_diffuse(wi) + ps * disney_pdf_specular(m, wo, wi);
}

bool bsdf_sample(const Material& m, V3 wo, V2 u, BsdfSample* out) {
  if (m.kind == 0) {  // :68-76
    if (wo.z <= 0.0f) return false;
    V3 wi = cosine_sample_hemisphere(u);
    out->wi = wi;
    out->pdf = wi.z / PI;
    out->f = m.base / PI;
    return true;
  }
  if (wo.z <= 0.0f) return false;  // :197-225
  float pd, ps;
  disney_lobe_weights(m, &pd, &ps);
  if (pd + ps <= 0.0f) return false;
  V3 wi;
  if (u.x < pd) {
    V2 ur{rmin(u.x / pd, 1.0f - 1e-7f), u.y};
    wi = cosine_sample_hemisphere(ur);
  } else {
    V2 ur{rmin((u.x - pd) / ps, 1.0f - 1e-7f), u.y};
    V3 h = sample_ggx_vndf(wo, disney_alpha(m), ur);
    wi = reflect(wo, h);
  }
  if (wi.z <= 0.0f) return false;
  float pdf = bsdf_pdf(m, wo, wi);
  if (pdf <= 0.0f) return false;
  out->wi = wi;
  out->f = bsdf_eval(m, wo, wi);
  out->pdf = pdf;
  return true;
}

// ------------------------------------------------------------------ light.rs
QuadLight::QuadLight(V3 p0_, V3 eu_, V3 ev_, V3 em) : p0(p0_), eu(eu_), ev(ev_), emission(em) {  // :61-76
  cross_ = cross(eu, ev);
  area = length(cross_);
  inv_cross_len_sq = 1.0f / length_squared(cross_);
  normal = normalize(cross_);
}

bool QuadLight::intersect_plane(V3 o, V3 d, float tmin, float tmax, float* t_out, V3* p_out) const {  // :86-107
  float denom = dot(d, normal);
  if (std::fabs(denom) < 1e-9f) return false;
  float t = dot(p0 - o, normal) / denom;
  if (t < tmin || t > tmax) return false;
  V3 p = o + t * d;
  V3 dd = p - p0;
  float alpha = dot(cross(dd, ev), cross_) * inv_cross_len_sq;
  if (!(alpha >= 0.0f && alpha <= 1.0f)) return false;
  float beta = dot(cross(eu, dd), cross_) * inv_cross_len_sq;
  if (!(beta >= 0.0f && beta <= 1.0f)) return false;
  *t_out = t;
  *p_out = p;
  return true;
}

bool QuadLight::sample(V3 from, V2 u, V3* pos, V3* wi_out, float* dist_out, V3* li, float* pdf) const {  // :111-133
  V3 p = p0 + u.x * eu + u.y * ev;
  V3 to_light = p - from;
  float dist2 = length_squared(to_light);
  if (dist2 <= 0.0f) return false;
  float dist = std::sqrt(dist2);
  V3 wi = to_light / dist;
  float cos_light = -dot(wi, normal);
  if (cos_light <= 0.0f) return false;
  *pdf = dist2 / (cos_light * area);
  *pos = p;
  *wi_out = wi;
  *dist_out = dist;
  *li = emission;
  return true;
}

float QuadLight::pdf(V3 from, V3 wi) const {  // :135-144
  float t;
  V3 p;
  if (!intersect_plane(from, wi, 1e-4f, INFINITY, &t, &p)) return 0.0f;
  float cos_light = -dot(wi, normal);
  if (cos_light <= 0.0f) return 0.0f;
  return t * t / (cos_light * area);
}

V3 QuadLight::le(V3, V3 n, V3 wo) const { return dot(n, wo) > 0.0f ? emission : splat(0.0f); }  // :146-152

// Emissive triangle mesh (no reference counterpart; the design is stated at
// its declaration).
MeshLight::MeshLight(const float* pos, uint32_t, const uint32_t* idx, uint32_t nt, V3 em) : emission(em) {
  std::vector<float> w(nt);
  for (uint32_t i = 0; i < nt; ++i) {
    const float* pa = pos + 3 * (size_t)idx[3 * i];
    const float* pb = pos + 3 * (size_t)idx[3 * i + 1];
    const float* pc = pos + 3 * (size_t)idx[3 * i + 2];
    V3 a = v3(pa[0], pa[1], pa[2]), b = v3(pb[0], pb[1], pb[2]), c = v3(pc[0], pc[1], pc[2]);
    V3 x = b - a, y = c - a;
    V3 cr = cross(x, y);
    v0.push_back(a);
    v1.push_back(b);
    v2.push_back(c);
    e1.push_back(x);
    e2.push_back(y);
    normal.push_back(normalize(cr));
    w[i] = 0.5f * length(cr);
  }
  dist = Distribution1D(w.data(), nt);
  area = dist.integral;
}

bool MeshLight::intersect(V3 o, V3 d, float tmin, float tmax, float* t_out, uint32_t* tri_out) const {
  bool any = false;
  for (size_t i = 0; i < v0.size(); ++i) {
    Ray r{o, d, tmin, tmax, 0.0f};
    float t, u, v;
    if (moller_trumbore(r, v0[i], v1[i], v2[i], &t, &u, &v)) {
      tmax = t;
      *t_out = t;
      *tri_out = (uint32_t)i;
      any = true;
    }
  }
  return any;
}

bool MeshLight::sample(V3 from, V2 u, V3* pos, V3* wi_out, float* dist_out, V3* li, float* pdf,
                       uint32_t* tri_out) const {
  size_t lo = 1, hi = dist.cdf.size() - 1;  // sample_continuous's bin search (distribution.rs:80-90)
  while (lo < hi) {
    size_t mid = (lo + hi) / 2;
    if (dist.cdf[mid] <= u.x)
      lo = mid + 1;
    else
      hi = mid;
  }
  size_t i = std::min(lo - 1, dist.n - 1);
  if (tri_out) *tri_out = (uint32_t)i;
  float du = (u.x - dist.cdf[i]) / rmax(dist.cdf[i + 1] - dist.cdf[i], 1e-20f);
  du = rclamp(du, 0.0f, 1.0f);
  float s = std::sqrt(du);
  V3 p = (v0[i] + (s * (1.0f - u.y)) * e1[i]) + (s * u.y) * e2[i];
  V3 to_light = p - from;
  float dist2 = length_squared(to_light);
  if (dist2 <= 0.0f) return false;
  float dl = std::sqrt(dist2);
  V3 wi = to_light / dl;
  float cos_light = -dot(wi, normal[i]);
  if (cos_light <= 0.0f) return false;
  *pdf = dist2 / (cos_light * area);
  *pos = p;
  *wi_out = wi;
  *dist_out = dl;
  *li = emission;
  return true;
}

float MeshLight::pdf(V3 from, V3 wi) const {
  float t;
  uint32_t i;
  if (!intersect(from, wi, 1e-4f, INFINITY, &t, &i)) return 0.0f;
  float cos_light = -dot(wi, normal[i]);
  if (cos_light <= 0.0f) return 0.0f;
  return t * t / (cos_light * area);
}

V3 MeshLight::le(V3, V3 n, V3 wo) const { return dot(n, wo) > 0.0f ? emission : splat(0.0f); }

// light.rs:214-242
Hdri::Hdri(uint32_t w, uint32_t h, std::vector<V3> px, float inten)
    : width(w), height(h), pixels(std::move(px)), intensity(inten) {
  std::vector<float> weights;
  weights.reserve((size_t)w * h);
  for (uint32_t y = 0; y < h; ++y) {
    float theta = ((float)y + 0.5f) / (float)h * PI;
    float sin_theta = rmax(std::sin(theta), 0.0f);
    for (uint32_t x = 0; x < w; ++x) {
      V3 p = pixels[(size_t)y * w + x];
      float lum = 0.21267f * p.x + 0.71516f * p.y + 0.07217f * p.z;
      weights.push_back(rmax(lum, 0.0f) * sin_theta);
    }
  }
  dist = Distribution2D(weights.data(), w, h);
}

// light.rs:252-272
std::vector<V3> hdri_gradient_pixels(V3 zenith, V3 horizon, uint32_t* w_out, uint32_t* h_out) {
  const uint32_t h = 64, w = 128;
  std::vector<V3> px;
  px.reserve(w * h);
  for (uint32_t y = 0; y < h; ++y) {
    float v = ((float)y + 0.5f) / (float)h;
    float cos_theta = 1.0f - 2.0f * v;
    V3 color = cos_theta >= 0.0f ? lerp(horizon, zenith, cos_theta)
                                 : horizon * rmax(1.0f + cos_theta * 0.5f, 0.0f);
    for (uint32_t x = 0; x < w; ++x) px.push_back(color);
  }
  *w_out = w;
  *h_out = h;
  return px;
}

V2 Hdri::dir_to_uv(V3 wi) {  // :274-280
  float theta = std::acos(rclamp(wi.y, -1.0f, 1.0f));
  float phi = std::atan2(wi.z, wi.x);
  float u = (phi + PI) * FRAC_1_PI * 0.5f;
  float v = theta * FRAC_1_PI;
  return V2{u, v};
}

V3 Hdri::uv_to_dir(V2 uv) {  // :282-287
  float phi = uv.x * TAU - PI;
  float theta = uv.y * PI;
  float sin_theta = std::sin(theta);
  return v3(sin_theta * std::cos(phi), std::cos(theta), sin_theta * std::sin(phi));
}

static inline float rfract(float x) { return x - std::trunc(x); }
static inline int32_t rem_euclid(int32_t a, int32_t b) {
  int32_t r = a % b;
  return r < 0 ? r + b : r;
}

V3 Hdri::sample_texture(V2 uv) const {  // :290-314
  int32_t w = (int32_t)width, h = (int32_t)height;
  float u = rfract(rfract(uv.x) + 1.0f);
  float v = rclamp(uv.y, 0.0f, 1.0f);
  float fx = u * (float)w - 0.5f;
  float fy = rclamp(v * (float)h - 0.5f, 0.0f, (float)h - 1.0f);
  int32_t x0i = f2i32(std::floor(fx));
  int32_t y0i = f2i32(std::floor(fy));
  float dx = fx - (float)x0i;
  float dy = fy - (float)y0i;
  int32_t x0 = rem_euclid(x0i, w);
  int32_t x1 = rem_euclid(x0 + 1, w);
  int32_t y0 = std::min(std::max(y0i, 0), h - 1);
  int32_t y1 = std::min(y0 + 1, h - 1);
  auto at = [&](int32_t x, int32_t y) { return pixels[(size_t)y * width + (size_t)x]; };
  V3 p00 = at(x0, y0), p10 = at(x1, y0), p01 = at(x0, y1), p11 = at(x1, y1);
  return p00 * ((1.0f - dx) * (1.0f - dy)) + p10 * (dx * (1.0f - dy)) + p01 * ((1.0f - dx) * dy) +
         p11 * (dx * dy);
}

V3 Hdri::le(V3 wi) const { return sample_texture(dir_to_uv(wi)) * intensity; }  // :318-320

void Hdri::sample(V2 u, V3* wi, V3* li, float* pdf) const {  // :322-338
  V2 uv;
  float pdf_uv;
  dist.sample_continuous(u, &uv, &pdf_uv);
  *wi = uv_to_dir(uv);
  float theta = uv.y * PI;
  float sin_theta = std::sin(theta);
  *pdf = (sin_theta > 1e-6f && pdf_uv > 0.0f) ? pdf_uv / (2.0f * PI * PI * sin_theta) : 0.0f;
  *li = sample_texture(uv) * intensity;
}

float Hdri::pdf(V3 wi) const {  // :340-349
  V2 uv = dir_to_uv(wi);
  float theta = uv.y * PI;
  float sin_theta = std::sin(theta);
  if (sin_theta < 1e-6f) return 0.0f;
  float pdf_uv = dist.pdf_at(uv);
  return pdf_uv / (2.0f * PI * PI * sin_theta);
}

// ----------------------------------------------------------------- camera.rs
// camera.rs:46-61 (pinhole) and :101-124 (thin lens share the same basis).
void Camera::look_at(V3 o, V3 target, V3 world_up, float vfov_deg, float aspect) {
  origin = o;
  forward = normalize(target - o);
  right = normalize(cross(forward, world_up));
  up = cross(right, forward);
  const float RADS_PER_DEG = PI / 180.0f;
  half_height = std::tan(vfov_deg * RADS_PER_DEG / 2.0f);
  half_width = aspect * half_height;
}

Ray Camera::generate_ray(V2 ndc, V2 lens, float time) const {
  if (!thin_lens) {  // camera.rs:65-73
    V3 dir = forward + ndc.x * half_width * right + ndc.y * half_height * up;
    Ray r = ray_new(origin, dir);
    r.time = time;
    return r;
  }
  // camera.rs:125-143
  V3 pinhole_dir = normalize(forward + ndc.x * half_width * right + ndc.y * half_height * up);
  float focal_t = focus_distance / dot(pinhole_dir, forward);
  V3 focal_point = origin + focal_t * pinhole_dir;
  V2 disk = concentric_disk(lens);
  V3 lens_offset = aperture_radius * (disk.x * right + disk.y * up);
  V3 lens_origin = origin + lens_offset;
  V3 dir = normalize(focal_point - lens_origin);
  Ray r = ray_new(lens_origin, dir);
  r.time = time;
  return r;
}

// ------------------------------------------------------------------ scene.rs
Scene::~Scene() { delete env; }

// scene.rs:57-97. Returns 0 miss, 1 surface (id = primitive), 2 light (id = light).
thread_local RenderStats* g_render_stats = nullptr;

int Scene::intersect(const Ray& ray, SurfaceInteraction* it_out, uint32_t* id_out) const {
  int kind = 0;
  float t_max = ray.t_max;
  TraversalStats* st = g_render_stats ? &g_render_stats->ext_st : nullptr;
  if (g_render_stats) g_render_stats->ext++;
  for (size_t id = 0; id < prims.size(); ++id) {
    Ray r = ray;
    r.t_max = t_max;
    SurfaceInteraction it;
    if (prims[id].intersect(r, &it, st)) {
      if (it.t < t_max) {
        t_max = it.t;
        *it_out = it;
        *id_out = (uint32_t)id;
        kind = 1;
      }
    }
  }
  for (size_t id = 0; id < lights.size(); ++id) {
    float t;
    V3 p;
    if (lights[id].intersect_plane(ray.o, ray.d, ray.t_min, t_max, &t, &p)) {
      if (t < t_max) {
        t_max = t;
        it_out->t = t;
        it_out->position = p;
        it_out->normal = lights[id].normal;
        it_out->tri = 0;
        it_out->u = it_out->v = 0.0f;
        *id_out = (uint32_t)id;
        kind = 2;
      }
    }
  }
  // emissive meshes after the quads, with the same strict t < t_max
  for (size_t id = 0; id < mesh_lights.size(); ++id) {
    float t;
    uint32_t tri;
    if (mesh_lights[id].intersect(ray.o, ray.d, ray.t_min, t_max, &t, &tri) && t < t_max) {
      t_max = t;
      it_out->t = t;
      it_out->position = ray_at(ray, t);
      it_out->normal = mesh_lights[id].normal[tri];
      it_out->tri = 0;
      it_out->u = it_out->v = 0.0f;
      *id_out = (uint32_t)(lights.size() + id);
      kind = 2;
    }
  }
  if (g_render_stats) {
    if (kind == 1) g_render_stats->ext_surface++;
    if (g_render_stats->first && kind != 0) g_render_stats->first_hit++;
    g_render_stats->first = false;
  }
  return kind;
}

bool Scene::occluded_dir(V3 from, V3 dir, float time) const {  // scene.rs:102-116
  Ray ray{from, dir, 1e-3f, INFINITY, time};
  SurfaceInteraction it;
  TraversalStats* st = g_render_stats ? &g_render_stats->shadow_st : nullptr;
  if (g_render_stats) g_render_stats->shadow++;
  for (const auto& p : prims)
    if (p.intersect(ray, &it, st)) return true;
  return false;
}

bool Scene::occluded(V3 from, V3 to, float time) const {  // scene.rs:120-140
  V3 to_from = to - from;
  float dist = length(to_from);
  if (dist <= 0.0f) return false;
  V3 dir = to_from / dist;
  Ray ray{from, dir, 1e-3f, dist - 1e-3f, time};
  SurfaceInteraction it;
  TraversalStats* st = g_render_stats ? &g_render_stats->shadow_st : nullptr;
  if (g_render_stats) g_render_stats->shadow++;
  for (const auto& p : prims)
    if (p.intersect(ray, &it, st)) return true;
  return false;
}

Bounds3 Scene::bounds() const {  // scene.rs:146-152
  Bounds3 b = bounds_empty();
  for (const auto& p : prims) b = bunion(b, p.world_bounds());
  return b;
}

// ------------------------------------------------------------- integrator.rs
float power_heuristic(float pa, float pb) {  // integrator.rs:234-240
  float a2 = pa * pa;
  float b2 = pb * pb;
  float denom = a2 + b2;
  return denom > 0.0f ? a2 / denom : 0.0f;
}

static const float RAY_EPS = 1e-3f;  // integrator.rs:34

// integrator.rs:57-231
PixelSample integrate(const Scene& sc, Ray ray, Sampler& sampler, uint32_t max_depth, uint32_t min_rr) {
  V3 l = splat(0.0f), beta = splat(1.0f);
  float last_pdf_bsdf = 0.0f;
  bool last_was_specular = true;
  V3 aov_albedo = splat(0.0f), aov_normal = splat(0.0f);
  float aov_depth = 0.0f;
  bool aovs_captured = false;
  const V3 ZERO = splat(0.0f);
  for (uint32_t depth = 0; depth < max_depth; ++depth) {
    SurfaceInteraction hit;
    uint32_t id = 0;
    int kind = sc.intersect(ray, &hit, &id);
    if (kind == 0) {
      if (sc.env) {
        V3 le = sc.env->le(ray.d);
        if (le != ZERO) {
          if (last_was_specular) {
            l = l + beta * le;
          } else {
            float pdf_env = sc.env->pdf(ray.d);
            float w_bsdf = power_heuristic(last_pdf_bsdf, pdf_env);
            l = l + beta * le * w_bsdf;
          }
        }
      }
      break;
    }
    if (!aovs_captured) {
      aovs_captured = true;
      aov_normal = hit.normal;
      aov_depth = hit.t;
      aov_albedo = kind == 1 ? bsdf_albedo(sc.materials[sc.material_id[id]]) : splat(1.0f);
    }
    if (kind == 2) {
      const bool quad = id < sc.lights.size();
      V3 le = quad ? sc.lights[id].le(hit.position, hit.normal, -ray.d)
                   : sc.mesh_lights[id - sc.lights.size()].le(hit.position, hit.normal, -ray.d);
      if (last_was_specular) {
        l = l + beta * le;
      } else {
        float pdf_light = quad ? sc.lights[id].pdf(ray.o, ray.d) : sc.mesh_lights[id - sc.lights.size()].pdf(ray.o, ray.d);
        float w_bsdf = power_heuristic(last_pdf_bsdf, pdf_light);
        l = l + beta * le * w_bsdf;
      }
      break;
    }
    const Material& mat = sc.materials[sc.material_id[id]];
    V3 p = hit.position;
    V3 ns = hit.normal;
    Frame frame = frame_from_normal(ns);
    V3 wo_local = to_local(frame, -ray.d);
    if (wo_local.z <= 0.0f) break;
    for (const auto& light : sc.lights) {
      V2 u = sampler.next_vec2();
      V3 lpos, lwi, lli;
      float ldist, lpdf;
      if (!light.sample(p, u, &lpos, &lwi, &ldist, &lli, &lpdf)) continue;
      if (lpdf <= 0.0f || lli == ZERO) continue;
      V3 wi_local = to_local(frame, lwi);
      if (wi_local.z <= 0.0f) continue;
      V3 f = bsdf_eval(mat, wo_local, wi_local);
      if (f == ZERO) continue;
      V3 p_offset = p + RAY_EPS * ns;
      V3 target = lpos - RAY_EPS * lwi;
      if (sc.occluded(p_offset, target, ray.time)) continue;
      float pdf_bsdf = bsdf_pdf(mat, wo_local, wi_local);
      float w_light = power_heuristic(lpdf, pdf_bsdf);
      l = l + beta * f * lli * wi_local.z * w_light / lpdf;
    }
    for (const auto& light : sc.mesh_lights) {  // the same arm for each emissive mesh
      V2 u = sampler.next_vec2();
      V3 lpos, lwi, lli;
      float ldist, lpdf;
      if (!light.sample(p, u, &lpos, &lwi, &ldist, &lli, &lpdf)) continue;
      if (lpdf <= 0.0f || lli == ZERO) continue;
      V3 wi_local = to_local(frame, lwi);
      if (wi_local.z <= 0.0f) continue;
      V3 f = bsdf_eval(mat, wo_local, wi_local);
      if (f == ZERO) continue;
      V3 p_offset = p + RAY_EPS * ns;
      V3 target = lpos - RAY_EPS * lwi;
      if (sc.occluded(p_offset, target, ray.time)) continue;
      float pdf_bsdf = bsdf_pdf(mat, wo_local, wi_local);
      float w_light = power_heuristic(lpdf, pdf_bsdf);
      l = l + beta * f * lli * wi_local.z * w_light / lpdf;
    }
    if (sc.env) {
      V2 u = sampler.next_vec2();
      V3 ewi, eli;
      float epdf;
      sc.env->sample(u, &ewi, &eli, &epdf);
      if (epdf > 0.0f && eli != ZERO) {
        V3 wi_local = to_local(frame, ewi);
        if (wi_local.z > 0.0f) {
          V3 f = bsdf_eval(mat, wo_local, wi_local);
          if (f != ZERO) {
            V3 p_offset = p + RAY_EPS * ns;
            if (!sc.occluded_dir(p_offset, ewi, ray.time)) {
              float pdf_bsdf = bsdf_pdf(mat, wo_local, wi_local);
              float w_light = power_heuristic(epdf, pdf_bsdf);
              l = l + beta * f * eli * wi_local.z * w_light / epdf;
            }
          }
        }
      }
    }
    BsdfSample bs;
    if (!bsdf_sample(mat, wo_local, sampler.next_vec2(), &bs)) break;
    if (bs.pdf <= 0.0f) break;
    float cos_wi = std::fabs(bs.wi.z);
    if (cos_wi == 0.0f) break;
    beta = beta * (bs.f * cos_wi / bs.pdf);
    if (depth >= min_rr) {
      float q = rclamp(1.0f - max_element(beta), 0.05f, 0.95f);
      if (sampler.next_f32() < q) break;
      beta = beta / (1.0f - q);
    }
    last_pdf_bsdf = bs.pdf;
    last_was_specular = false;
    V3 wi_world = to_world(frame, bs.wi);
    ray = Ray{p + RAY_EPS * ns, wi_world, 1e-4f, INFINITY, ray.time};
  }
  return PixelSample{l, aov_albedo, aov_normal, aov_depth};
}

}  // namespace orc

// ============================================================================
// C API for the Python tests (ctypes). Plain pointers, no ownership transfer
// except the opaque handles.
// ============================================================================
using namespace orc;

namespace {
struct Render {
  uint32_t width, height, spp, spp_offset, max_depth, min_rr;
};

// crates/pyre-cli/src/main.rs:353-360
TriangleMesh quad_mesh(const V3 c[4], V3 n) {
  TriangleMesh m;
  m.positions.assign(c, c + 4);
  m.indices = {0, 1, 2, 0, 2, 3};
  m.normals.assign(4, n);
  m.has_normals = true;
  return m;
}

// crates/pyre-cli/src/main.rs:565-599
TriangleMesh make_uv_sphere(V3 center, float radius, uint32_t us, uint32_t vs) {
  TriangleMesh m;
  for (uint32_t v = 0; v <= vs; ++v) {
    float phi = PI * (float)v / (float)vs;
    float sp = std::sin(phi), cp = std::cos(phi);
    for (uint32_t u = 0; u <= us; ++u) {
      float theta = TAU * (float)u / (float)us;
      V3 n = v3(sp * std::cos(theta), cp, sp * std::sin(theta));
      m.positions.push_back(center + radius * n);
      m.normals.push_back(n);
    }
  }
  uint32_t row = us + 1;
  for (uint32_t v = 0; v < vs; ++v)
    for (uint32_t u = 0; u < us; ++u) {
      uint32_t i00 = v * row + u, i01 = v * row + u + 1, i10 = (v + 1) * row + u, i11 = (v + 1) * row + u + 1;
      uint32_t t[6] = {i00, i10, i11, i00, i11, i01};
      m.indices.insert(m.indices.end(), t, t + 6);
    }
  m.has_normals = true;
  return m;
}

void add_prim(Scene* sc, TriangleMesh mesh, uint32_t mat, const BvhParams& p) {
  MeshInstance inst;
  inst.mesh = std::move(mesh);
  inst.build(p);
  sc->prims.push_back(std::move(inst));
  sc->material_id.push_back(mat);
}

// crates/pyre-cli/src/main.rs:365-496
Scene* cornell_box(bool motion_blur, const BvhParams& bp) {
  Scene* sc = new Scene();
  uint32_t white = 0, red = 1, green = 2;
  sc->materials.push_back(Material{0, splat(0.73f), 0, 0, 0});
  sc->materials.push_back(Material{0, v3(0.65f, 0.05f, 0.05f), 0, 0, 0});
  sc->materials.push_back(Material{0, v3(0.12f, 0.45f, 0.15f), 0, 0, 0});
  const float s = 1.0f;
  V3 floor_c[4] = {v3(-s, -s, -s), v3(s, -s, -s), v3(s, -s, s), v3(-s, -s, s)};
  V3 ceil_c[4] = {v3(-s, s, -s), v3(s, s, -s), v3(s, s, s), v3(-s, s, s)};
  V3 back_c[4] = {v3(-s, -s, -s), v3(s, -s, -s), v3(s, s, -s), v3(-s, s, -s)};
  V3 left_c[4] = {v3(-s, -s, -s), v3(-s, -s, s), v3(-s, s, s), v3(-s, s, -s)};
  V3 right_c[4] = {v3(s, -s, -s), v3(s, s, -s), v3(s, s, s), v3(s, -s, s)};
  add_prim(sc, quad_mesh(floor_c, v3(0, 1, 0)), white, bp);
  add_prim(sc, quad_mesh(ceil_c, -v3(0, 1, 0)), white, bp);
  add_prim(sc, quad_mesh(back_c, v3(0, 0, 1)), white, bp);
  add_prim(sc, quad_mesh(left_c, v3(1, 0, 0)), red, bp);
  add_prim(sc, quad_mesh(right_c, -v3(1, 0, 0)), green, bp);
  uint32_t gold = (uint32_t)sc->materials.size();
  sc->materials.push_back(Material{1, v3(1.0f, 0.766f, 0.336f), 1.0f, 0.1f, 0.5f});
  add_prim(sc, make_uv_sphere(v3(0.35f, -0.65f, 0.25f), 0.35f, 48, 24), gold, bp);
  if (motion_blur) {
    sc->prims.back().has_motion = true;
    sc->prims.back().off0 = splat(0.0f);
    sc->prims.back().off1 = v3(0.0f, 0.25f, 0.0f);
  }
  uint32_t plastic = (uint32_t)sc->materials.size();
  sc->materials.push_back(Material{1, splat(0.72f), 0.0f, 0.18f, 0.5f});
  add_prim(sc, make_uv_sphere(v3(-0.35f, -0.65f, -0.3f), 0.35f, 48, 24), plastic, bp);
  sc->lights.emplace_back(v3(-0.3f, 0.999f, -0.3f), v3(0.6f, 0.0f, 0.0f), v3(0.0f, 0.0f, 0.6f), splat(15.0f));
  return sc;
}

template <class F>
void parallel_for(size_t n, uint32_t threads, size_t chunk, F f) {
  if (threads <= 1 || n <= chunk) {
    f(0, n);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> pool;
  for (uint32_t t = 0; t < threads; ++t)
    pool.emplace_back([&]() {
      for (;;) {
        size_t a = next.fetch_add(chunk);
        if (a >= n) break;
        f(a, std::min(n, a + chunk));
      }
    });
  for (auto& th : pool) th.join();
}
inline V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }
inline void st3(float* p, V3 v) {
  p[0] = v.x;
  p[1] = v.y;
  p[2] = v.z;
}
}  // namespace

extern "C" {

uint64_t orc_pixel_seed(uint32_t x, uint32_t y, uint32_t s) { return pixel_seed(x, y, s); }

void orc_rng_state(uint64_t seed, uint64_t* out4) {
  Sampler s(seed);
  std::memcpy(out4, s.s, 32);
}

void orc_rng_f32(uint64_t seed, uint32_t n, float* out) {
  Sampler s(seed);
  for (uint32_t i = 0; i < n; ++i) out[i] = s.next_f32();
}

void orc_rng_u64(uint64_t seed, uint32_t n, uint64_t* out) {
  Sampler s(seed);
  for (uint32_t i = 0; i < n; ++i) out[i] = s.next_u64();
}

// xoshiro256++ outputs from an explicit state (for published test vectors).
void orc_rng_u64_from_state(const uint64_t* state4, uint32_t n, uint64_t* out) {
  Sampler s(0);
  std::memcpy(s.s, state4, 32);
  for (uint32_t i = 0; i < n; ++i) out[i] = s.next_u64();
}

void orc_concentric_disk(const float* u, uint32_t n, float* out) {
  for (uint32_t i = 0; i < n; ++i) {
    V2 d = concentric_disk(V2{u[2 * i], u[2 * i + 1]});
    out[2 * i] = d.x;
    out[2 * i + 1] = d.y;
  }
}

// --- Distribution1D / 2D
void* orc_dist1d_new(const float* w, uint32_t n) { return new Distribution1D(w, n); }
void orc_dist1d_free(void* d) { delete (Distribution1D*)d; }
void orc_dist1d_cdf(void* d, float* out) {
  auto* p = (Distribution1D*)d;
  std::memcpy(out, p->cdf.data(), p->cdf.size() * 4);
}
float orc_dist1d_integral(void* d) { return ((Distribution1D*)d)->integral; }
void orc_dist1d_sample(void* d, const float* u, uint32_t n, float* t, float* pdf, uint32_t* bin) {
  for (uint32_t i = 0; i < n; ++i) {
    size_t b;
    ((Distribution1D*)d)->sample_continuous(u[i], &t[i], &pdf[i], &b);
    bin[i] = (uint32_t)b;
  }
}
void orc_dist1d_pdf(void* d, const float* t, uint32_t n, float* pdf) {
  for (uint32_t i = 0; i < n; ++i) pdf[i] = ((Distribution1D*)d)->pdf_at(t[i]);
}
void* orc_dist2d_new(const float* w, uint32_t width, uint32_t height) {
  return new Distribution2D(w, width, height);
}
void orc_dist2d_free(void* d) { delete (Distribution2D*)d; }
void orc_dist2d_sample(void* d, const float* u, uint32_t n, float* uv, float* pdf) {
  for (uint32_t i = 0; i < n; ++i) {
    V2 o;
    ((Distribution2D*)d)->sample_continuous(V2{u[2 * i], u[2 * i + 1]}, &o, &pdf[i]);
    uv[2 * i] = o.x;
    uv[2 * i + 1] = o.y;
  }
}
void orc_dist2d_pdf(void* d, const float* uv, uint32_t n, float* pdf) {
  for (uint32_t i = 0; i < n; ++i) pdf[i] = ((Distribution2D*)d)->pdf_at(V2{uv[2 * i], uv[2 * i + 1]});
}

// --- BSDFs. mat = {kind, base.rgb, metallic, roughness, specular} as 7 floats.
static Material mat_from(const float* m) {
  return Material{(int)m[0], v3(m[1], m[2], m[3]), m[4], m[5], m[6]};
}
void orc_bsdf_eval(const float* mat, const float* wo, const float* wi, uint32_t n, float* f, float* pdf) {
  Material m = mat_from(mat);
  for (uint32_t i = 0; i < n; ++i) {
    st3(f + 3 * i, bsdf_eval(m, ld3(wo + 3 * i), ld3(wi + 3 * i)));
    pdf[i] = bsdf_pdf(m, ld3(wo + 3 * i), ld3(wi + 3 * i));
  }
}
void orc_bsdf_sample(const float* mat, const float* wo, const float* u, uint32_t n, float* wi, float* f,
                     float* pdf, uint8_t* ok) {
  Material m = mat_from(mat);
  for (uint32_t i = 0; i < n; ++i) {
    BsdfSample bs{};
    ok[i] = bsdf_sample(m, ld3(wo + 3 * i), V2{u[2 * i], u[2 * i + 1]}, &bs) ? 1 : 0;
    st3(wi + 3 * i, bs.wi);
    st3(f + 3 * i, bs.f);
    pdf[i] = bs.pdf;
  }
}

// --- Quad light: light = p0, eu, ev, emission (12 floats)
void orc_quad_sample(const float* L, const float* from, const float* u, uint32_t n, float* pos, float* wi,
                     float* dist, float* pdf, uint8_t* ok) {
  QuadLight ql(ld3(L), ld3(L + 3), ld3(L + 6), ld3(L + 9));
  for (uint32_t i = 0; i < n; ++i) {
    V3 p{}, w{}, li{};
    float d = 0, pd = 0;
    ok[i] = ql.sample(ld3(from + 3 * i), V2{u[2 * i], u[2 * i + 1]}, &p, &w, &d, &li, &pd) ? 1 : 0;
    st3(pos + 3 * i, p);
    st3(wi + 3 * i, w);
    dist[i] = d;
    pdf[i] = pd;
  }
}
void orc_quad_pdf(const float* L, const float* from, const float* wi, uint32_t n, float* pdf) {
  QuadLight ql(ld3(L), ld3(L + 3), ld3(L + 6), ld3(L + 9));
  for (uint32_t i = 0; i < n; ++i) pdf[i] = ql.pdf(ld3(from + 3 * i), ld3(wi + 3 * i));
}

// --- HDRI environment
void* orc_hdri_new(uint32_t w, uint32_t h, const float* rgb, float intensity) {
  std::vector<V3> px((size_t)w * h);
  for (size_t i = 0; i < px.size(); ++i) px[i] = ld3(rgb + 3 * i);
  return new Hdri(w, h, std::move(px), intensity);
}
void* orc_hdri_gradient(const float* zenith, const float* horizon, float intensity) {
  uint32_t w, h;
  auto px = hdri_gradient_pixels(ld3(zenith), ld3(horizon), &w, &h);
  return new Hdri(w, h, std::move(px), intensity);
}
void orc_hdri_free(void* e) { delete (Hdri*)e; }
// The map's size (w, h) and, when out != NULL, its texels (3 floats each, row-major).
void orc_hdri_texels(void* e, uint32_t* wh, float* out) {
  const Hdri* H = (const Hdri*)e;
  wh[0] = H->width;
  wh[1] = H->height;
  if (out)
    for (size_t i = 0; i < H->pixels.size(); ++i) st3(out + 3 * i, H->pixels[i]);
}
void orc_hdri_sample(void* e, const float* u, uint32_t n, float* wi, float* li, float* pdf) {
  for (uint32_t i = 0; i < n; ++i) {
    V3 w, l;
    ((Hdri*)e)->sample(V2{u[2 * i], u[2 * i + 1]}, &w, &l, &pdf[i]);
    st3(wi + 3 * i, w);
    st3(li + 3 * i, l);
  }
}
void orc_hdri_pdf(void* e, const float* wi, uint32_t n, float* pdf) {
  for (uint32_t i = 0; i < n; ++i) pdf[i] = ((Hdri*)e)->pdf(ld3(wi + 3 * i));
}
void orc_hdri_le(void* e, const float* wi, uint32_t n, float* le) {
  for (uint32_t i = 0; i < n; ++i) st3(le + 3 * i, ((Hdri*)e)->le(ld3(wi + 3 * i)));
}
void orc_hdri_uv(const float* wi, uint32_t n, float* uv) {
  for (uint32_t i = 0; i < n; ++i) {
    V2 o = Hdri::dir_to_uv(ld3(wi + 3 * i));
    uv[2 * i] = o.x;
    uv[2 * i + 1] = o.y;
  }
}
void orc_hdri_dir(const float* uv, uint32_t n, float* wi) {
  for (uint32_t i = 0; i < n; ++i) st3(wi + 3 * i, Hdri::uv_to_dir(V2{uv[2 * i], uv[2 * i + 1]}));
}

// --- Scenes
void* orc_scene_new() { return new Scene(); }
void orc_scene_free(void* s) { delete (Scene*)s; }
void* orc_scene_cornell(int motion_blur, uint32_t bins, uint32_t max_leaf) {
  BvhParams bp;
  bp.bins = bins;
  bp.max_leaf = max_leaf;
  return cornell_box(motion_blur != 0, bp);
}
uint32_t orc_scene_add_material(void* s, const float* mat) {
  auto* sc = (Scene*)s;
  sc->materials.push_back(mat_from(mat));
  return (uint32_t)sc->materials.size() - 1;
}
// normals may be null. Returns the primitive id.
uint32_t orc_scene_add_mesh(void* s, const float* pos, uint32_t nv, const uint32_t* idx, uint32_t nt,
                            const float* nrm, uint32_t material_id, uint32_t bins, uint32_t max_leaf) {
  TriangleMesh m;
  m.positions.resize(nv);
  for (uint32_t i = 0; i < nv; ++i) m.positions[i] = ld3(pos + 3 * i);
  m.indices.assign(idx, idx + 3 * (size_t)nt);
  if (nrm) {
    m.has_normals = true;
    m.normals.resize(nv);
    for (uint32_t i = 0; i < nv; ++i) m.normals[i] = ld3(nrm + 3 * i);
  }
  BvhParams bp;
  bp.bins = bins;
  bp.max_leaf = max_leaf;
  add_prim((Scene*)s, std::move(m), material_id, bp);
  return (uint32_t)((Scene*)s)->prims.size() - 1;
}
void orc_scene_set_motion(void* s, uint32_t prim, const float* off0, const float* off1) {
  auto& p = ((Scene*)s)->prims[prim];
  p.has_motion = true;
  p.off0 = ld3(off0);
  p.off1 = ld3(off1);
}
void orc_scene_add_light(void* s, const float* L) {
  ((Scene*)s)->lights.emplace_back(ld3(L), ld3(L + 3), ld3(L + 6), ld3(L + 9));
}
void orc_scene_add_mesh_light(void* s, const float* pos, uint32_t nv, const uint32_t* idx, uint32_t nt,
                              const float* em) {
  ((Scene*)s)->mesh_lights.emplace_back(pos, nv, idx, nt, ld3(em));
}
// --- Emissive mesh probes: sample / pdf / intersect of a free-standing mesh light
void* orc_mesh_light_new(const float* pos, uint32_t nv, const uint32_t* idx, uint32_t nt, const float* em) {
  return new MeshLight(pos, nv, idx, nt, ld3(em));
}
void orc_mesh_light_free(void* m) { delete (MeshLight*)m; }
float orc_mesh_light_area(void* m) { return ((MeshLight*)m)->area; }
// per query: ok, pos[3], wi[3], dist, pdf, tri
void orc_mesh_light_sample(void* m, const float* from, const float* u, uint32_t n, uint8_t* ok, float* pos,
                           float* wi, float* dist, float* pdf, uint32_t* tri) {
  const MeshLight& L = *(MeshLight*)m;
  for (uint32_t i = 0; i < n; ++i) {
    V3 p, w, li;
    float d = 0, pd = 0;
    uint32_t t = 0;
    ok[i] = L.sample(ld3(from + 3 * i), V2{u[2 * i], u[2 * i + 1]}, &p, &w, &d, &li, &pd, &t);
    st3(pos + 3 * i, ok[i] ? p : splat(0));
    st3(wi + 3 * i, ok[i] ? w : splat(0));
    dist[i] = ok[i] ? d : 0;
    pdf[i] = ok[i] ? pd : 0;
    tri[i] = t;
  }
}
void orc_mesh_light_pdf(void* m, const float* from, const float* wi, uint32_t n, float* pdf) {
  const MeshLight& L = *(MeshLight*)m;
  for (uint32_t i = 0; i < n; ++i) pdf[i] = L.pdf(ld3(from + 3 * i), ld3(wi + 3 * i));
}
void orc_scene_set_env(void* s, void* hdri) {
  auto* sc = (Scene*)s;
  delete sc->env;
  sc->env = (Hdri*)hdri;  // the scene takes ownership
}
uint32_t orc_scene_num_prims(void* s) { return (uint32_t)((Scene*)s)->prims.size(); }
uint32_t orc_scene_num_materials(void* s) { return (uint32_t)((Scene*)s)->materials.size(); }
void orc_scene_material(void* s, uint32_t i, float* out7) {
  const Material& m = ((Scene*)s)->materials[i];
  out7[0] = (float)m.kind;
  st3(out7 + 1, m.base);
  out7[4] = m.metallic;
  out7[5] = m.roughness;
  out7[6] = m.specular;
}
uint32_t orc_scene_num_lights(void* s) { return (uint32_t)((Scene*)s)->lights.size(); }
void orc_scene_light(void* s, uint32_t i, float* out12) {
  const QuadLight& l = ((Scene*)s)->lights[i];
  st3(out12, l.p0);
  st3(out12 + 3, l.eu);
  st3(out12 + 6, l.ev);
  st3(out12 + 9, l.emission);
}
// mesh info: nv, nt, has_normals, material
void orc_scene_mesh_info(void* s, uint32_t i, uint32_t* out4) {
  auto* sc = (Scene*)s;
  const auto& m = sc->prims[i].mesh;
  out4[0] = (uint32_t)m.positions.size();
  out4[1] = (uint32_t)m.triangle_count();
  out4[2] = m.has_normals ? 1 : 0;
  out4[3] = sc->material_id[i];
}
void orc_scene_mesh_data(void* s, uint32_t i, float* pos, uint32_t* idx, float* nrm) {
  const auto& m = ((Scene*)s)->prims[i].mesh;
  for (size_t k = 0; k < m.positions.size(); ++k) st3(pos + 3 * k, m.positions[k]);
  std::memcpy(idx, m.indices.data(), m.indices.size() * 4);
  if (nrm && m.has_normals)
    for (size_t k = 0; k < m.normals.size(); ++k) st3(nrm + 3 * k, m.normals[k]);
}
void orc_scene_bounds(void* s, float* out6) {
  Bounds3 b = ((Scene*)s)->bounds();
  st3(out6, b.min);
  st3(out6 + 3, b.max);
}
// BVH export of primitive i: sizes, then nodes as (bounds6 | offset, n, axis).
void orc_scene_bvh_info(void* s, uint32_t i, uint32_t* n_nodes, uint32_t* n_prims, uint32_t* depth) {
  const auto& b = ((Scene*)s)->prims[i].bvh;
  *n_nodes = (uint32_t)b.nodes.size();
  *n_prims = (uint32_t)b.prim.size();
  *depth = (uint32_t)b.max_depth;
}
void orc_scene_bvh_export(void* s, uint32_t i, float* bounds6, uint32_t* offset, uint32_t* nprim,
                          uint32_t* axis, uint32_t* prim) {
  const auto& b = ((Scene*)s)->prims[i].bvh;
  for (size_t k = 0; k < b.nodes.size(); ++k) {
    st3(bounds6 + 6 * k, b.nodes[k].bounds.min);
    st3(bounds6 + 6 * k + 3, b.nodes[k].bounds.max);
    offset[k] = b.nodes[k].offset;
    nprim[k] = b.nodes[k].n_primitives;
    axis[k] = b.nodes[k].axis;
  }
  std::memcpy(prim, b.prim.data(), b.prim.size() * 4);
}

// rays: 9 floats (o, d, t_min, t_max, time). Scene::intersect; out per ray:
// kind (0 miss / 1 surface / 2 light), id, tri, and t, u, v, position, normal.
static void scene_intersect_range(Scene* sc, const float* rays, size_t a, size_t b, int32_t* kind,
                                  uint32_t* id, uint32_t* tri, float* tuv, float* pn) {
  for (size_t i = a; i < b; ++i) {
    const float* r = rays + 9 * i;
    Ray ray{ld3(r), ld3(r + 3), r[6], r[7], r[8]};
    SurfaceInteraction it{};
    uint32_t pid = UINT32_MAX;
    int k = sc->intersect(ray, &it, &pid);
    kind[i] = k;
    id[i] = k ? pid : UINT32_MAX;
    tri[i] = k == 1 ? it.tri : UINT32_MAX;
    tuv[3 * i] = k ? it.t : 0.0f;
    tuv[3 * i + 1] = k ? it.u : 0.0f;
    tuv[3 * i + 2] = k ? it.v : 0.0f;
    if (pn) {
      st3(pn + 6 * i, k ? it.position : splat(0.0f));
      st3(pn + 6 * i + 3, k ? it.normal : splat(0.0f));
    }
  }
}


void orc_scene_intersect(void* s, const float* rays, uint32_t n, uint32_t threads, int32_t* kind,
                         uint32_t* id, uint32_t* tri, float* tuv, float* pn) {
  parallel_for(n, threads, 4096, [&](size_t a, size_t b) {
    scene_intersect_range((Scene*)s, rays, a, b, kind, id, tri, tuv, pn);
  });
}

// Scene::occluded semantics generalised to a ray with explicit t range:
// any primitive with a hit in (t_min, t_max) (scene.rs:102-140).
void orc_scene_any_hit(void* s, const float* rays, uint32_t n, uint32_t threads, uint8_t* occ) {
  auto* sc = (Scene*)s;
  parallel_for(n, threads, 4096, [&](size_t a, size_t b) {
    for (size_t i = a; i < b; ++i) {
      const float* r = rays + 9 * i;
      Ray ray{ld3(r), ld3(r + 3), r[6], r[7], r[8]};
      SurfaceInteraction it;
      uint8_t o = 0;
      for (const auto& p : sc->prims)
        if (p.intersect(ray, &it, nullptr)) {
          o = 1;
          break;
        }
      occ[i] = o;
    }
  });
}

void orc_scene_occluded(void* s, const float* from, const float* to, const float* time, uint32_t n, uint8_t* occ) {
  for (uint32_t i = 0; i < n; ++i) occ[i] = ((Scene*)s)->occluded(ld3(from + 3 * i), ld3(to + 3 * i), time[i]);
}

// Reference traversal (bvh.rs:86-143) statistics on primitive 0: node fetches
// and triangle tests per ray, summed. any != 0 stops at the first hit, like
// Scene::occluded's per-primitive early exit.
void orc_mesh_traversal_stats(void* s, const float* rays, uint32_t n, uint32_t threads, uint64_t* nodes,
                              uint64_t* tris) {
  auto* sc = (Scene*)s;
  std::atomic<uint64_t> an{0}, at{0};
  parallel_for(n, threads, 1024, [&](size_t a, size_t b) {
    TraversalStats st;
    for (size_t i = a; i < b; ++i) {
      const float* r = rays + 9 * i;
      Ray ray{ld3(r), ld3(r + 3), r[6], r[7], r[8]};
      SurfaceInteraction it;
      sc->prims[0].intersect(ray, &it, &st);
    }
    an += st.nodes;
    at += st.tris;
  });
  *nodes = an;
  *tris = at;
}

// Brute force over every triangle of primitive 0: the minimum accepted t and the
// count of triangles achieving it (for tie detection).
void orc_mesh_brute_closest(void* s, const float* rays, uint32_t n, uint32_t threads, uint32_t* tri,
                            float* t_out, uint32_t* nties) {
  auto* sc = (Scene*)s;
  const auto& m = sc->prims[0].mesh;
  size_t nt = m.triangle_count();
  parallel_for(n, threads, 64, [&](size_t a, size_t b) {
    for (size_t i = a; i < b; ++i) {
      const float* r = rays + 9 * i;
      Ray ray{ld3(r), ld3(r + 3), r[6], r[7], r[8]};
      float best = INFINITY;
      uint32_t bt = UINT32_MAX, ties = 0;
      for (size_t k = 0; k < nt; ++k) {
        V3 v0 = m.positions[m.indices[3 * k]], v1 = m.positions[m.indices[3 * k + 1]],
           v2 = m.positions[m.indices[3 * k + 2]];
        float t, u, v;
        Ray rr = ray;
        if (!moller_trumbore(rr, v0, v1, v2, &t, &u, &v)) continue;
        if (t < best) {
          best = t;
          bt = (uint32_t)k;
          ties = 1;
        } else if (t == best) {
          ties++;
        }
      }
      tri[i] = bt;
      t_out[i] = bt == UINT32_MAX ? 0.0f : best;
      nties[i] = ties;
    }
  });
}

// --- Camera: cam = origin3, target3, up3, vfov, aspect, aperture, focus (13 floats)
static Camera cam_from(const float* c) {
  Camera cam;
  cam.look_at(ld3(c), ld3(c + 3), ld3(c + 6), c[9], c[10]);
  cam.aperture_radius = c[11];
  cam.focus_distance = c[12];
  cam.thin_lens = c[11] > 0.0f;
  return cam;
}
void orc_camera_rays(const float* cam, const float* samples, uint32_t n, float* rays) {
  Camera c = cam_from(cam);
  for (uint32_t i = 0; i < n; ++i) {
    const float* sm = samples + 5 * i;
    Ray r = c.generate_ray(V2{sm[0], sm[1]}, V2{sm[2], sm[3]}, sm[4]);
    st3(rays + 9 * i, r.o);
    st3(rays + 9 * i + 3, r.d);
    rays[9 * i + 6] = r.t_min;
    rays[9 * i + 7] = r.t_max;
    rays[9 * i + 8] = r.time;
  }
}

// main.rs:337-351
void orc_auto_frame(const float* b6, float vfov_deg, float aspect, float* origin, float* target) {
  Bounds3 b{ld3(b6), ld3(b6 + 3)};
  V3 center = bcentroid(b);
  float radius = length(b.max - b.min) * 0.5f;
  radius = rmax(radius, 1e-3f);
  const float RADS_PER_DEG = PI / 180.0f;
  float half_vfov = vfov_deg * RADS_PER_DEG * 0.5f;
  float half_hfov = std::atan(aspect * std::tan(half_vfov));
  float dist_v = radius / std::sin(half_vfov);
  float dist_h = radius / std::sin(half_hfov);
  float distance = rmax(dist_v, dist_h) * 1.15f;
  V3 dir = normalize(v3(0.0f, 0.3f, 1.0f));
  st3(origin, center + distance * dir);
  st3(target, center);
}

// One full path sample (main.rs:261-274) for pixel (x, y), sample s.
static PixelSample render_one(const Scene& sc, const Camera& cam, const Render& R, uint32_t x, uint32_t y,
                              uint32_t s) {
  Sampler smp(pixel_seed(x, y, s));
  float jx = smp.next_f32();
  float jy = smp.next_f32();
  float ndc_x = 2.0f * ((float)x + jx) / (float)R.width - 1.0f;
  float ndc_y = 1.0f - 2.0f * ((float)y + jy) / (float)R.height;
  V2 lens = smp.next_vec2();
  float time = smp.next_f32();
  Ray ray = cam.generate_ray(V2{ndc_x, ndc_y}, lens, time);
  return integrate(sc, ray, smp, R.max_depth, R.min_rr);
}

// Film::render with the main.rs:254-286 closure over rows [y0, y1) (a crop for
// bounded baselines). out: 10 floats per pixel of the crop (PixelSample).
// If per_sample != null it receives radiance (3 floats) of every sample, laid
// out [pixel][s] over the crop, for correlated-sampling checks.
void orc_render(void* s, const float* cam, uint32_t width, uint32_t height, uint32_t spp, uint32_t spp_offset,
                uint32_t max_depth, uint32_t min_rr, uint32_t y0, uint32_t y1, uint32_t threads, float* out,
                float* per_sample) {
  const Scene& sc = *(Scene*)s;
  Camera c = cam_from(cam);
  Render R{width, height, spp, spp_offset, max_depth, min_rr};
  size_t rows = y1 - y0;
  parallel_for(rows, threads, 1, [&](size_t a, size_t b) {
    for (size_t ry = a; ry < b; ++ry) {
      uint32_t y = y0 + (uint32_t)ry;
      for (uint32_t x = 0; x < width; ++x) {
        float inv_spp = 1.0f / (float)spp;
        V3 rad = splat(0), alb = splat(0), nrm = splat(0);
        float dep = 0.0f;
        size_t pix = ry * width + x;
        for (uint32_t si = 0; si < spp; ++si) {
          PixelSample ps = render_one(sc, c, R, x, y, spp_offset + si);
          rad = rad + ps.radiance;
          alb = alb + ps.albedo;
          nrm = nrm + ps.normal;
          dep += ps.depth;
          if (per_sample) st3(per_sample + 3 * (pix * spp + si), ps.radiance);
        }
        float* o = out + 10 * pix;
        st3(o, rad * inv_spp);
        st3(o + 3, alb * inv_spp);
        st3(o + 6, nrm * inv_spp);
        o[9] = dep * inv_spp;
      }
    }
  });
}

// The reference traversal's work in a render of the pixels (x, y) with
// x % stride == 0 and y % stride == 0, spp samples each: out = {extension
// queries, their node fetches, their triangle tests, shadow queries, their
// node fetches, their triangle tests, extension queries that hit a surface,
// samples whose camera ray hit a surface or a light}.
void orc_render_stats(void* s, const float* cam, uint32_t width, uint32_t height, uint32_t spp, uint32_t spp_offset,
                      uint32_t max_depth, uint32_t min_rr, uint32_t stride, uint32_t threads, uint64_t* out) {
  const Scene& sc = *(Scene*)s;
  Camera c = cam_from(cam);
  Render R{width, height, spp, spp_offset, max_depth, min_rr};
  const uint32_t rows = (height + stride - 1) / stride;
  std::mutex mu;
  RenderStats total;
  parallel_for(rows, threads, 1, [&](size_t a, size_t b) {
    RenderStats local;
    g_render_stats = &local;
    for (size_t ry = a; ry < b; ++ry) {
      const uint32_t y = (uint32_t)ry * stride;
      for (uint32_t x = 0; x < width; x += stride)
        for (uint32_t si = 0; si < spp; ++si) {
          local.first = true;
          (void)render_one(sc, c, R, x, y, spp_offset + si);
        }
    }
    g_render_stats = nullptr;
    std::lock_guard<std::mutex> lk(mu);
    total.ext += local.ext;
    total.shadow += local.shadow;
    total.ext_st.nodes += local.ext_st.nodes;
    total.ext_st.tris += local.ext_st.tris;
    total.shadow_st.nodes += local.shadow_st.nodes;
    total.shadow_st.tris += local.shadow_st.tris;
    total.ext_surface += local.ext_surface;
    total.first_hit += local.first_hit;
  });
  out[0] = total.ext;
  out[1] = total.ext_st.nodes;
  out[2] = total.ext_st.tris;
  out[3] = total.shadow;
  out[4] = total.shadow_st.nodes;
  out[5] = total.shadow_st.tris;
  out[6] = total.ext_surface;
  out[7] = total.first_hit;
}

// film.rs:181-198 (beauty), :125-134 (albedo), :136-147 (normal), :149-173 (depth)
void orc_encode_png(const float* samples, uint32_t w, uint32_t h, uint32_t channel, uint8_t* out) {
  size_t n = (size_t)w * h;
  const float g = 1.0f / 2.2f;
  if (channel == 0 || channel == 1) {
    for (size_t i = 0; i < n; ++i) {
      const float* v = samples + 10 * i + (channel == 0 ? 0 : 3);
      for (int c = 0; c < 3; ++c) {
        float x = rmin(rmax(v[c], 0.0f), 1.0f);  // glam clamp = max(min).min(max)
        out[3 * i + c] = f2u8(std::pow(x, g) * 255.0f);
      }
    }
  } else if (channel == 2) {
    for (size_t i = 0; i < n; ++i) {
      const float* v = samples + 10 * i + 6;
      for (int c = 0; c < 3; ++c) {
        float e = rmin(rmax(v[c] * 0.5f + 0.5f, 0.0f), 1.0f) * 255.0f;
        out[3 * i + c] = f2u8(e);
      }
    }
  } else {
    float lo = INFINITY, hi = 0.0f;
    for (size_t i = 0; i < n; ++i) {
      float d = samples[10 * i + 9];
      if (d > 0.0f) {
        lo = rmin(lo, d);
        hi = rmax(hi, d);
      }
    }
    float range = rmax(hi - lo, 1e-6f);
    for (size_t i = 0; i < n; ++i) {
      float d = samples[10 * i + 9];
      float v = d > 0.0f ? rclamp(1.0f - (d - lo) / range, 0.0f, 1.0f) : 0.0f;
      out[i] = f2u8(v * 255.0f);
    }
  }
}

// Frame helpers exposed for tests (math.rs:81-105).
void orc_frame(const float* n, uint32_t cnt, float* out9) {
  for (uint32_t i = 0; i < cnt; ++i) {
    Frame f = frame_from_normal(ld3(n + 3 * i));
    st3(out9 + 9 * i, f.t);
    st3(out9 + 9 * i + 3, f.bt);
    st3(out9 + 9 * i + 6, f.n);
  }
}

}  // extern "C"
