// Shade-stage functions on the device: BSDFs (crates/pyre/src/material.rs),
// quad and environment lights (light.rs), inverse-CDF sampling
// (distribution.rs) and camera rays (camera.rs). Same f32 operation order as
// the reference. The BSDF samplers' sin/cos reproduce the host libm's sinf /
// cosf bit for bit (PhiTable below); elsewhere (thin-lens disk, environment
// directions) sin/cos/acos/atan2 are the device math library's and agree with
// the reference to ~1 ulp.
#pragma once

#include "internal.hpp"

namespace pyre {

// sinf / cosf of phi = 2 pi u on the sampler's grid u = k 2^-24 (k < 2^24:
// every u a sample stream yields, sampler.rs:24-33), as the host's libm
// computes them: the angle of cosine_sample_hemisphere (material.rs:239-246)
// and sample_ggx_vndf (:291-317). The reference calls libm's sinf / cosf,
// which are not correctly rounded; the device evaluates sin / cos in double
// and rounds once (the correctly rounded float wherever the double lies far
// from a float rounding boundary), and the grid points where libm's float
// differs from that, or where the double lies within 2^-40 of a boundary,
// are exceptions listed with libm's values. rank[k / 64] = (exception bits of
// grid points 64(k/64) .. +63 (low, high word), exceptions before them, 0);
// val = (sinf, cosf) of every exception in grid order. Off the grid (the
// shade-stage probes' arbitrary u) the double evaluation alone is used.
struct PhiTable {
  const uint4* rank = nullptr;
  const float2* val = nullptr;
};
constexpr uint32_t kPhiGrid = 1u << 24;

// sin and cos in double for moderate |x| (the samplers' [0, 2 pi)): x reduced
// by the nearest multiple of pi/2 (pi/2 as a double plus its remainder, each
// product exact inside the fma), then the Taylor series of sin to x^15 and of
// cos to x^16 on |r| <= pi/4 (truncation below 2^-53 relative). Relative
// error a few double ulps: far inside the 2^-40 margin the exception table
// keeps from every float rounding boundary.
PH void sincos_f64(double x, double* s_out, double* c_out) {
  const double kPio2Hi = 1.5707963267948966, kPio2Lo = 6.123233995736766e-17;
  const double k = rint(x * 0.63661977236758134);
  double r = fma(-k, kPio2Hi, x);
  r = fma(-k, kPio2Lo, r);
  const double r2 = r * r;
  double ps = 1.0 / 1307674368000.0;           // 1/15!
  ps = fma(ps, -r2, 1.0 / 6227020800.0);       // 1/13!
  ps = fma(ps, -r2, 1.0 / 39916800.0);         // 1/11!
  ps = fma(ps, -r2, 1.0 / 362880.0);           // 1/9!
  ps = fma(ps, -r2, 1.0 / 5040.0);             // 1/7!
  ps = fma(ps, -r2, 1.0 / 120.0);              // 1/5!
  ps = fma(ps, -r2, 1.0 / 6.0);                // 1/3!
  const double sr = fma(-r * r2, ps, r);
  double pc = 1.0 / 20922789888000.0;          // 1/16!
  pc = fma(pc, -r2, 1.0 / 87178291200.0);      // 1/14!
  pc = fma(pc, -r2, 1.0 / 479001600.0);        // 1/12!
  pc = fma(pc, -r2, 1.0 / 3628800.0);          // 1/10!
  pc = fma(pc, -r2, 1.0 / 40320.0);            // 1/8!
  pc = fma(pc, -r2, 1.0 / 720.0);              // 1/6!
  pc = fma(pc, -r2, 1.0 / 24.0);               // 1/4!
  pc = fma(pc, -r2, 0.5);                      // 1/2!
  const double cr = fma(-r2, pc, 1.0);
  const int q = (int)((int64_t)k & 3);
  *s_out = q == 0 ? sr : q == 1 ? cr : q == 2 ? -sr : -cr;
  *c_out = q == 0 ? cr : q == 1 ? -sr : q == 2 ? -cr : sr;
}

PH void sincos_phi(float uy, float phi, const PhiTable& T, float* s, float* c) {
  const float ks = uy * 16777216.0f;  // exact: a power-of-two scale
  if (T.rank && ks >= 0.0f && ks < 16777216.0f) {
    const uint32_t k = (uint32_t)ks;
    if ((float)k == ks) {
      const uint4 r = T.rank[k >> 6];
      const uint64_t bits = (uint64_t)r.x | ((uint64_t)r.y << 32);
      const uint32_t b = k & 63u;
      if ((bits >> b) & 1ull) {
        const float2 v = T.val[r.z + (uint32_t)__builtin_popcountll(bits & ((1ull << b) - 1ull))];
        *s = v.x;
        *c = v.y;
        return;
      }
    }
  }
  double ds, dc;
  sincos_f64((double)phi, &ds, &dc);
  *s = (float)ds;
  *c = (float)dc;
}

// ---------------------------------------------------------------- material.rs
PH float luminance(V3 c) { return dot(c, v3(0.2126f, 0.7152f, 0.0722f)); }  // :232-235
PH float powi5(float x) { return x * ((x * x) * (x * x)); }  // llvm.powi(x, 5) as emitted for f32::powi(5)

PH V3 cosine_sample_hemisphere(V2 u, const PhiTable& T) {  // :239-246
  float r = sqrtf(u.x);
  float phi = 2.0f * kPi * u.y;
  float sp, cp;
  sincos_phi(u.y, phi, T, &sp, &cp);
  float x = r * cp;
  float y = r * sp;
  float z = sqrtf(rmax(1.0f - u.x, 0.0f));
  return v3(x, y, z);
}

PH float ggx_d(V3 h, float alpha) {  // :249-256
  if (h.z <= 0.0f) return 0.0f;
  float a2 = alpha * alpha;
  float denom = h.z * h.z * (a2 - 1.0f) + 1.0f;
  return a2 / (kPi * denom * denom);
}

PH V3 schlick_fresnel(float cos_d, V3 f0) {  // :259-262
  float c = rclamp(cos_d, 0.0f, 1.0f);
  return f0 + (splat(1.0f) - f0) * powi5(1.0f - c);
}

PH float ggx_lambda(V3 v, float alpha) {  // :264-272
  float cos2 = v.z * v.z;
  if (cos2 >= 1.0f) return 0.0f;
  float a2 = alpha * alpha;
  float tan2 = (1.0f - cos2) / rmax(cos2, 1e-8f);
  return 0.5f * (-1.0f + sqrtf(1.0f + a2 * tan2));
}
PH float ggx_g1(V3 v, float alpha) { return 1.0f / (1.0f + ggx_lambda(v, alpha)); }  // :274-276
PH float ggx_g2(V3 wo, V3 wi, float alpha) {                                        // :279-281
  return 1.0f / (1.0f + ggx_lambda(wo, alpha) + ggx_lambda(wi, alpha));
}
PH V3 reflect(V3 v, V3 n) { return 2.0f * dot(v, n) * n - v; }  // :285-287

PH V3 sample_ggx_vndf(V3 wo, float alpha, V2 u, const PhiTable& T) {  // :291-317
  V3 vh = normalize(v3(alpha * wo.x, alpha * wo.y, wo.z));
  float lensq = vh.x * vh.x + vh.y * vh.y;
  V3 t1 = lensq > 0.0f ? v3(-vh.y, vh.x, 0.0f) / sqrtf(lensq) : v3(1.0f, 0.0f, 0.0f);
  V3 t2 = cross(vh, t1);
  float r = sqrtf(u.x);
  float phi = 2.0f * kPi * u.y;
  float sp, cp;
  sincos_phi(u.y, phi, T, &sp, &cp);
  float t1c = r * cp;
  float t2c = r * sp;
  float s = 0.5f * (1.0f + vh.z);
  t2c = (1.0f - s) * sqrtf(rmax(1.0f - t1c * t1c, 0.0f)) + s * t2c;
  V3 nh = t1c * t1 + t2c * t2 + sqrtf(rmax(1.0f - t1c * t1c - t2c * t2c, 0.0f)) * vh;
  return normalize(v3(alpha * nh.x, alpha * nh.y, rmax(nh.z, 0.0f)));
}

PH float disney_alpha(const MatDev& m) { return rmax(m.roughness * m.roughness, 1e-3f); }  // :97-102
PH V3 disney_f0(const MatDev& m) { return lerp(splat(0.08f * m.specular), m.base, m.metallic); }  // :104-107
PH void disney_lobe_weights(const MatDev& m, float* pd, float* ps) {  // :112-121
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
PH V3 disney_eval_diffuse(const MatDev& m, V3 wo, V3 wi) {  // :125-138
  if (wo.z <= 0.0f || wi.z <= 0.0f || m.metallic >= 1.0f) return splat(0.0f);
  V3 h = normalize_or_zero(wo + wi);
  if (length_squared(h) < 1e-10f) return splat(0.0f);
  float cos_d = rclamp(dot(wi, h), 0.0f, 1.0f);
  float fd90 = 0.5f + 2.0f * m.roughness * cos_d * cos_d;
  float fd_v = 1.0f + (fd90 - 1.0f) * powi5(rmax(1.0f - wo.z, 0.0f));
  float fd_l = 1.0f + (fd90 - 1.0f) * powi5(rmax(1.0f - wi.z, 0.0f));
  return m.base * (1.0f - m.metallic) / kPi * fd_v * fd_l;
}
PH V3 disney_eval_specular(const MatDev& m, V3 wo, V3 wi) {  // :142-155
  if (wo.z <= 0.0f || wi.z <= 0.0f) return splat(0.0f);
  V3 h = normalize_or_zero(wo + wi);
  if (length_squared(h) < 1e-10f || h.z <= 0.0f) return splat(0.0f);
  float alpha = disney_alpha(m);
  float d = ggx_d(h, alpha);
  V3 f = schlick_fresnel(rmax(dot(wi, h), 0.0f), disney_f0(m));
  float g = ggx_g2(wo, wi, alpha);
  return d * f * g / (4.0f * wo.z * wi.z);
}
PH float disney_pdf_specular(const MatDev& m, V3 wo, V3 wi) {  // :168-180
  if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
  V3 h = normalize_or_zero(wo + wi);
  if (length_squared(h) < 1e-10f || h.z <= 0.0f) return 0.0f;
  float alpha = disney_alpha(m);
  float d = ggx_d(h, alpha);
  float g1 = ggx_g1(wo, alpha);
  return g1 * d / (4.0f * wo.z);
}

PH V3 bsdf_eval(const MatDev& m, V3 wo, V3 wi) {
  if (m.kind == PYRE_MAT_LAMBERTIAN) {  // :52-58
    if (wo.z <= 0.0f || wi.z <= 0.0f) return splat(0.0f);
    return m.base / kPi;
  }
  return disney_eval_diffuse(m, wo, wi) + disney_eval_specular(m, wo, wi);  // :188-190
}

PH float bsdf_pdf(const MatDev& m, V3 wo, V3 wi) {
  if (m.kind == PYRE_MAT_LAMBERTIAN) {  // :60-66
    if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
    return wi.z / kPi;
  }
  float pd, ps;  // :192-195 (pdf_diffuse :157-163 inlined)
  disney_lobe_weights(m, &pd, &ps);
  float pdf_d = wi.z <= 0.0f ? 0.0f : wi.z / kPi;
  return pd * pdf_d + ps * disney_pdf_specular(m, wo, wi);
}

// Bsdf::sample; returns false where the reference returns None.
PH bool bsdf_sample(const MatDev& m, V3 wo, V2 u, const PhiTable& T, V3* wi_out, V3* f_out, float* pdf_out) {
  if (m.kind == PYRE_MAT_LAMBERTIAN) {  // :68-76
    if (wo.z <= 0.0f) return false;
    V3 wi = cosine_sample_hemisphere(u, T);
    *wi_out = wi;
    *pdf_out = wi.z / kPi;
    *f_out = m.base / kPi;
    return true;
  }
  if (wo.z <= 0.0f) return false;  // :197-225
  float pd, ps;
  disney_lobe_weights(m, &pd, &ps);
  if (pd + ps <= 0.0f) return false;
  V3 wi;
  if (u.x < pd) {
    V2 ur{rmin(u.x / pd, 1.0f - 1e-7f), u.y};
    wi = cosine_sample_hemisphere(ur, T);
  } else {
    V2 ur{rmin((u.x - pd) / ps, 1.0f - 1e-7f), u.y};
    V3 h = sample_ggx_vndf(wo, disney_alpha(m), ur, T);
    wi = reflect(wo, h);
  }
  if (wi.z <= 0.0f) return false;
  float pdf = bsdf_pdf(m, wo, wi);
  if (pdf <= 0.0f) return false;
  *wi_out = wi;
  *f_out = bsdf_eval(m, wo, wi);
  *pdf_out = pdf;
  return true;
}

// ------------------------------------------------------------------ light.rs
inline QuadLightDev make_quad_light(V3 p0, V3 eu, V3 ev, V3 em) {  // :61-76 (host)
  QuadLightDev l;
  l.p0 = p0;
  l.eu = eu;
  l.ev = ev;
  l.emission = em;
  l.cross = cross(eu, ev);
  l.area = length(l.cross);
  l.inv_cross_len_sq = 1.0f / length_squared(l.cross);
  l.normal = normalize(l.cross);
  return l;
}

PH bool quad_intersect_plane(const QuadLightDev& L, V3 o, V3 d, float tmin, float tmax, float* t_out,
                             V3* p_out) {  // :86-107
  float denom = dot(d, L.normal);
  if (fabsf(denom) < 1e-9f) return false;
  float t = dot(L.p0 - o, L.normal) / denom;
  if (t < tmin || t > tmax) return false;
  V3 p = o + t * d;
  V3 dd = p - L.p0;
  float alpha = dot(cross(dd, L.ev), L.cross) * L.inv_cross_len_sq;
  if (!(alpha >= 0.0f && alpha <= 1.0f)) return false;
  float beta = dot(cross(L.eu, dd), L.cross) * L.inv_cross_len_sq;
  if (!(beta >= 0.0f && beta <= 1.0f)) return false;
  *t_out = t;
  *p_out = p;
  return true;
}

PH bool quad_sample(const QuadLightDev& L, V3 from, V2 u, V3* pos, V3* wi_out, float* dist_out,
                    float* pdf) {  // :111-133
  V3 p = L.p0 + u.x * L.eu + u.y * L.ev;
  V3 to_light = p - from;
  float dist2 = length_squared(to_light);
  if (dist2 <= 0.0f) return false;
  float dist = sqrtf(dist2);
  V3 wi = to_light / dist;
  float cos_light = -dot(wi, L.normal);
  if (cos_light <= 0.0f) return false;
  *pdf = dist2 / (cos_light * L.area);
  *pos = p;
  *wi_out = wi;
  *dist_out = dist;
  return true;
}

PH float quad_pdf(const QuadLightDev& L, V3 from, V3 wi) {  // :135-144
  float t;
  V3 p;
  if (!quad_intersect_plane(L, from, wi, 1e-4f, kInf, &t, &p)) return 0.0f;
  float cos_light = -dot(wi, L.normal);
  if (cos_light <= 0.0f) return 0.0f;
  return t * t / (cos_light * L.area);
}

// Emissive triangle mesh (pyre_mesh_light; no reference counterpart, the
// design is stated in pyre_hip.h). et holds 4 float4 per emitter triangle
// (v0, e1, e2, normal). The nearest Moller-Trumbore hit over the mesh's
// triangles in order, t in (t_min, t_max) tightened by each hit, so the
// first of equal t wins.
// Whether the line o + t d, t >= 0, meets the mesh's padded box: a cull test
// that no ray Moller-Trumbore could accept fails. The interval starts
// unbounded, and an axis whose terms are NaN (o on a plane of the box, d = 0
// there) leaves it so (fmin / fmax drop NaN).
PH bool mlight_box(const MeshLightDev& L, V3 o, V3 d) {
  const float ix = 1.0f / d.x, iy = 1.0f / d.y, iz = 1.0f / d.z;
  const float x0 = (L.bmin.x - o.x) * ix, x1 = (L.bmax.x - o.x) * ix;
  const float y0 = (L.bmin.y - o.y) * iy, y1 = (L.bmax.y - o.y) * iy;
  const float z0 = (L.bmin.z - o.z) * iz, z1 = (L.bmax.z - o.z) * iz;
  float tn = -kInf, tf = kInf;
  tn = fmaxf(tn, fminf(x0, x1));
  tf = fminf(tf, fmaxf(x0, x1));
  tn = fmaxf(tn, fminf(y0, y1));
  tf = fminf(tf, fmaxf(y0, y1));
  tn = fmaxf(tn, fminf(z0, z1));
  tf = fminf(tf, fmaxf(z0, z1));
  return tn <= tf && tf >= 0.0f;
}

PH bool mlight_intersect(const MeshLightDev& L, const float4* et, V3 o, V3 d, float tmin, float tmax, float* t_out,
                         uint32_t* tri_out) {
  bool any = false;
  if (!mlight_box(L, o, d)) return false;
  for (uint32_t i = 0; i < L.ntri; ++i) {
    const float4* q = et + 4 * (size_t)(L.tri0 + i);
    const float4 a = q[0], b = q[1], c = q[2];
    float t, u, v;
    if (mt_intersect(o, d, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), tmin, tmax, false, &t, &u, &v)) {
      tmax = t;
      *t_out = t;
      *tri_out = i;
      any = true;
    }
  }
  return any;
}

PH V3 mlight_normal(const MeshLightDev& L, const float4* et, uint32_t tri) {
  const float4 n = et[4 * (size_t)(L.tri0 + tri) + 3];
  return v3(n.x, n.y, n.z);
}

// A triangle by the bin search of Distribution1D::sample_continuous on u.x
// (distribution.rs:78-101), a uniform point on it from the bin's remapped
// u.x and u.y, then DiffuseAreaQuadLight::sample's steps (light.rs:111-133)
// with the mesh's total area.
PH bool mlight_sample(const MeshLightDev& L, const float4* et, const float* ecdf, V3 from, V2 u, V3* pos,
                      V3* wi_out, float* dist_out, float* pdf) {
  const float* cdf = ecdf + L.cdf0;
  uint32_t lo = 1, hi = L.ntri;
  while (lo < hi) {
    uint32_t mid = (lo + hi) / 2;
    if (cdf[mid] <= u.x)
      lo = mid + 1;
    else
      hi = mid;
  }
  const uint32_t i = lo - 1 < L.ntri - 1 ? lo - 1 : L.ntri - 1;
  float du = (u.x - cdf[i]) / rmax(cdf[i + 1] - cdf[i], 1e-20f);
  du = rclamp(du, 0.0f, 1.0f);
  const float s = sqrtf(du);
  const float4* q = et + 4 * (size_t)(L.tri0 + i);
  const float4 a = q[0], b = q[1], c = q[2], nn = q[3];
  V3 p = (v3(a.x, a.y, a.z) + (s * (1.0f - u.y)) * v3(b.x, b.y, b.z)) + (s * u.y) * v3(c.x, c.y, c.z);
  V3 to_light = p - from;
  float dist2 = length_squared(to_light);
  if (dist2 <= 0.0f) return false;
  float dist = sqrtf(dist2);
  V3 wi = to_light / dist;
  float cos_light = -dot(wi, v3(nn.x, nn.y, nn.z));
  if (cos_light <= 0.0f) return false;
  *pdf = dist2 / (cos_light * L.area);
  *pos = p;
  *wi_out = wi;
  *dist_out = dist;
  return true;
}

PH float mlight_pdf(const MeshLightDev& L, const float4* et, V3 from, V3 wi) {
  float t;
  uint32_t i;
  if (!mlight_intersect(L, et, from, wi, 1e-4f, kInf, &t, &i)) return 0.0f;
  float cos_light = -dot(wi, mlight_normal(L, et, i));
  if (cos_light <= 0.0f) return 0.0f;
  return t * t / (cos_light * L.area);
}

// ----------------------------------------------------------- distribution.rs
// Distribution1D::sample_continuous over a cdf of n + 1 entries (:78-101).
// The reference's binary search finds the first i in [1, n) with cdf[i] > u
// (n if none); with a guide table the same search runs over the bracket
// [guide[k], guide[k + 1]] that holds that index (k = floor(u kEnvGuide),
// exact for a power-of-two table), so the bin is the same.
PH void dist1d_sample(const float* cdf, uint32_t n, float u, float* t, float* pdf, uint32_t* bin_out,
                      const uint32_t* guide = nullptr) {
  uint32_t lo = 1, hi = n;
  if (guide) {
    uint32_t k = (uint32_t)(u * (float)kEnvGuide);
    k = k < kEnvGuide ? k : kEnvGuide - 1u;
    lo = guide[k];
    hi = guide[k + 1];
  }
  while (lo < hi) {
    uint32_t mid = (lo + hi) / 2;
    if (cdf[mid] <= u)
      lo = mid + 1;
    else
      hi = mid;
  }
  uint32_t bin = lo - 1 < n - 1 ? lo - 1 : n - 1;
  float cdf_lo = cdf[bin];
  float cdf_hi = cdf[bin + 1];
  float bin_size = rmax(cdf_hi - cdf_lo, 1e-20f);
  float du = (u - cdf_lo) / bin_size;
  *t = ((float)bin + du) / (float)n;
  *pdf = bin_size * (float)n;
  *bin_out = bin;
}

PH float dist1d_pdf(const float* cdf, uint32_t n, float t) {  // :104-112
  if (t < 0.0f || t >= 1.0f || n == 0) return 0.0f;
  uint32_t bin = f2u32(t * (float)n);
  if (bin > n - 1) bin = n - 1;
  return (cdf[bin + 1] - cdf[bin]) * (float)n;
}

// concentric_disk (distribution.rs:12-23)
PH V2 concentric_disk(V2 u) {
  V2 off{2.0f * u.x - 1.0f, 2.0f * u.y - 1.0f};
  if (off.x == 0.0f && off.y == 0.0f) return V2{0.0f, 0.0f};
  float r, theta;
  if (fabsf(off.x) > fabsf(off.y)) {
    r = off.x;
    theta = kFracPi4 * (off.y / off.x);
  } else {
    r = off.y;
    theta = kFracPi2 - kFracPi4 * (off.x / off.y);
  }
  return V2{r * cosf(theta), r * sinf(theta)};
}

// -------------------------------------------------------- HDRI (light.rs:205-350)
PH V2 env_dir_to_uv(V3 wi) {  // :274-280
  float theta = acosf(rclamp(wi.y, -1.0f, 1.0f));
  float phi = atan2f(wi.z, wi.x);
  float u = (phi + kPi) * kFrac1Pi * 0.5f;
  float v = theta * kFrac1Pi;
  return V2{u, v};
}

PH V3 env_uv_to_dir(V2 uv) {  // :282-287
  float phi = uv.x * kTau - kPi;
  float theta = uv.y * kPi;
  float sin_theta = sinf(theta);
  return v3(sin_theta * cosf(phi), cosf(theta), sin_theta * sinf(phi));
}

PH float rfract(float x) { return x - truncf(x); }
PH int32_t rem_euclid(int32_t a, int32_t b) {
  int32_t r = a % b;
  return r < 0 ? r + b : r;
}

__device__ __forceinline__ V3 env_texel(const EnvDev& e, int32_t x, int32_t y) {
  float4 t = e.texels[(size_t)y * e.width + (size_t)x];
  return v3(t.x, t.y, t.z);
}

__device__ __forceinline__ V3 env_sample_texture(const EnvDev& e, V2 uv) {  // :290-314
  int32_t w = (int32_t)e.width, h = (int32_t)e.height;
  float u = rfract(rfract(uv.x) + 1.0f);
  float v = rclamp(uv.y, 0.0f, 1.0f);
  float fx = u * (float)w - 0.5f;
  float fy = rclamp(v * (float)h - 0.5f, 0.0f, (float)h - 1.0f);
  int32_t x0i = f2i32(floorf(fx));
  int32_t y0i = f2i32(floorf(fy));
  float dx = fx - (float)x0i;
  float dy = fy - (float)y0i;
  int32_t x0 = rem_euclid(x0i, w);
  int32_t x1 = rem_euclid(x0 + 1, w);
  int32_t y0 = min(max(y0i, 0), h - 1);
  int32_t y1 = min(y0 + 1, h - 1);
  V3 p00 = env_texel(e, x0, y0), p10 = env_texel(e, x1, y0), p01 = env_texel(e, x0, y1), p11 = env_texel(e, x1, y1);
  return p00 * ((1.0f - dx) * (1.0f - dy)) + p10 * (dx * (1.0f - dy)) + p01 * ((1.0f - dx) * dy) + p11 * (dx * dy);
}

__device__ __forceinline__ V3 env_le(const EnvDev& e, V3 wi) {  // :318-320
  return env_sample_texture(e, env_dir_to_uv(wi)) * e.intensity;
}

__device__ __forceinline__ void env_sample(const EnvDev& e, V2 u, V3* wi, V3* li, float* pdf) {  // :322-338
  float v, pdf_v, uu, pdf_u;
  uint32_t row, col;
  dist1d_sample(e.marg_cdf, e.height, u.y, &v, &pdf_v, &row, e.marg_guide);  // distribution.rs:151-155
  dist1d_sample(e.cond_cdf + (size_t)row * (e.width + 1), e.width, u.x, &uu, &pdf_u, &col,
                e.cond_guide ? e.cond_guide + (size_t)row * (kEnvGuide + 1) : nullptr);
  V2 uv{uu, v};
  float pdf_uv = pdf_u * pdf_v;
  *wi = env_uv_to_dir(uv);
  float theta = uv.y * kPi;
  float sin_theta = sinf(theta);
  *pdf = (sin_theta > 1e-6f && pdf_uv > 0.0f) ? pdf_uv / (2.0f * kPi * kPi * sin_theta) : 0.0f;
  *li = env_sample_texture(e, uv) * e.intensity;
}

__device__ __forceinline__ float env_pdf(const EnvDev& e, V3 wi) {  // :340-349
  V2 uv = env_dir_to_uv(wi);
  float theta = uv.y * kPi;
  float sin_theta = sinf(theta);
  if (sin_theta < 1e-6f) return 0.0f;
  // Distribution2D::pdf_at (distribution.rs:157-163)
  float pdf_uv = 0.0f;
  if (!(uv.x < 0.0f || uv.x >= 1.0f || uv.y < 0.0f || uv.y >= 1.0f)) {
    uint32_t row = f2u32(uv.y * (float)e.height);
    if (row > e.height - 1) row = e.height - 1;
    pdf_uv = dist1d_pdf(e.cond_cdf + (size_t)row * (e.width + 1), e.width, uv.x) *
             dist1d_pdf(e.marg_cdf, e.height, uv.y);
  }
  return pdf_uv / (2.0f * kPi * kPi * sin_theta);
}

// ----------------------------------------------------------------- camera.rs
struct CamDev {
  V3 origin, forward, right, up;
  float half_width, half_height, aperture_radius, focus_distance;
  uint32_t thin_lens;
};

// look_at (camera.rs:46-61, :101-124), evaluated on the host.
inline CamDev make_camera(const pyre_camera_desc& c) {
  CamDev k;
  V3 o = v3(c.origin[0], c.origin[1], c.origin[2]);
  V3 t = v3(c.target[0], c.target[1], c.target[2]);
  V3 up = v3(c.up[0], c.up[1], c.up[2]);
  k.origin = o;
  k.forward = normalize(t - o);
  k.right = normalize(cross(k.forward, up));
  k.up = cross(k.right, k.forward);
  const float rads_per_deg = kPi / 180.0f;
  k.half_height = tanf(c.vfov_deg * rads_per_deg / 2.0f);
  k.half_width = c.aspect * k.half_height;
  k.aperture_radius = c.aperture_radius;
  k.focus_distance = c.focus_distance > 0.0f ? c.focus_distance : length(t - o);  // main.rs:203-205
  k.thin_lens = c.aperture_radius > 0.0f ? 1u : 0u;                              // main.rs:206
  return k;
}

// generate_ray (camera.rs:65-73 pinhole, :125-143 thin lens) including
// Ray::new's normalisation (math.rs:16-24).
PH void camera_ray(const CamDev& c, V2 ndc, V2 lens, V3* o, V3* d) {
  if (!c.thin_lens) {
    V3 dir = c.forward + ndc.x * c.half_width * c.right + ndc.y * c.half_height * c.up;
    *o = c.origin;
    *d = normalize(dir);
    return;
  }
  V3 pinhole_dir = normalize(c.forward + ndc.x * c.half_width * c.right + ndc.y * c.half_height * c.up);
  float focal_t = c.focus_distance / dot(pinhole_dir, c.forward);
  V3 focal_point = c.origin + focal_t * pinhole_dir;
  V2 disk = concentric_disk(lens);
  V3 lens_offset = c.aperture_radius * (disk.x * c.right + disk.y * c.up);
  V3 lens_origin = c.origin + lens_offset;
  V3 dir = normalize(focal_point - lens_origin);
  *o = lens_origin;
  *d = normalize(dir);
}

PH float power_heuristic(float pa, float pb) {  // integrator.rs:234-240
  float a2 = pa * pa;
  float b2 = pb * pb;
  float denom = a2 + b2;
  return denom > 0.0f ? a2 / denom : 0.0f;
}

}  // namespace pyre
