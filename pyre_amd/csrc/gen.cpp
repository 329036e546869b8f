// Host-side scene and workload builders (include/pyre_gen.h).
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <vector>

#include "../../include/pyre_gen.h"
#include "pmath.hpp"
#include "preset.hpp"

using namespace pyre;

namespace {

// xoshiro256++ stream for workload generation (same generator as the
// per-sample sampler, seeded through SplitMix64).
struct HostRng {
  Rng r;
  explicit HostRng(uint64_t seed) : r(rng_seed(seed)) {}
  double f64() { return (double)(rng_next_u64(r) >> 11) * (1.0 / 9007199254740992.0); }
  double uniform(double a, double b) { return a + (b - a) * f64(); }
  double gauss() {  // Box-Muller, u1 in (0, 1]
    double u1 = 1.0 - f64();
    double u2 = f64();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

using Mesh = PresetMesh;

}  // namespace

namespace {

void push3(std::vector<float>& v, V3 a) {
  v.push_back(a.x);
  v.push_back(a.y);
  v.push_back(a.z);
}

// main.rs:353-360
Mesh quad_mesh(const V3 c[4], V3 n, uint32_t mat) {
  Mesh m;
  for (int i = 0; i < 4; ++i) {
    push3(m.pos, c[i]);
    push3(m.nrm, n);
  }
  m.idx = {0, 1, 2, 0, 2, 3};
  m.material = mat;
  return m;
}

// main.rs:565-599
Mesh make_uv_sphere(V3 center, float radius, uint32_t us, uint32_t vs, uint32_t mat) {
  Mesh m;
  for (uint32_t v = 0; v <= vs; ++v) {
    float phi = kPi * (float)v / (float)vs;
    float sp = sinf(phi), cp = cosf(phi);
    for (uint32_t u = 0; u <= us; ++u) {
      float theta = kTau * (float)u / (float)us;
      V3 n = v3(sp * cosf(theta), cp, sp * sinf(theta));
      push3(m.pos, center + radius * n);
      push3(m.nrm, n);
    }
  }
  uint32_t row = us + 1;
  for (uint32_t v = 0; v < vs; ++v)
    for (uint32_t u = 0; u < us; ++u) {
      uint32_t i00 = v * row + u, i01 = v * row + u + 1, i10 = (v + 1) * row + u, i11 = (v + 1) * row + u + 1;
      uint32_t t[6] = {i00, i10, i11, i00, i11, i01};
      m.idx.insert(m.idx.end(), t, t + 6);
    }
  m.material = mat;
  return m;
}

pyre_material lambert(V3 c) { return pyre_material{PYRE_MAT_LAMBERTIAN, {c.x, c.y, c.z}, 0.0f, 0.0f, 0.0f}; }
pyre_material disney(V3 c, float m, float r, float s) {
  return pyre_material{PYRE_MAT_DISNEY, {c.x, c.y, c.z}, m, r, s};
}

// Unit icosphere subdivided `level` times (midpoints projected to the sphere
// in double precision).
void icosphere(uint32_t level, std::vector<double>& P, std::vector<uint32_t>& F) {
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  double base[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                        {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  P.clear();
  for (auto& b : base) {
    double l = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    P.push_back(b[0] / l);
    P.push_back(b[1] / l);
    P.push_back(b[2] / l);
  }
  F = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
       3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
  for (uint32_t l = 0; l < level; ++l) {
    std::map<uint64_t, uint32_t> cache;
    auto mid = [&](uint32_t a, uint32_t b) {
      uint64_t key = a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
      auto it = cache.find(key);
      if (it != cache.end()) return it->second;
      double x = (P[3 * a] + P[3 * b]) * 0.5, y = (P[3 * a + 1] + P[3 * b + 1]) * 0.5,
             z = (P[3 * a + 2] + P[3 * b + 2]) * 0.5;
      double len = std::sqrt(x * x + y * y + z * z);
      uint32_t id = (uint32_t)(P.size() / 3);
      P.push_back(x / len);
      P.push_back(y / len);
      P.push_back(z / len);
      cache.emplace(key, id);
      return id;
    };
    std::vector<uint32_t> nf;
    nf.reserve(F.size() * 4);
    for (size_t f = 0; f < F.size(); f += 3) {
      uint32_t a = F[f], b = F[f + 1], c = F[f + 2];
      uint32_t ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
      uint32_t q[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
      nf.insert(nf.end(), q, q + 12);
    }
    F.swap(nf);
  }
}

}  // namespace

void preset_finish(pyre_preset* p) {
  p->mdesc.clear();
  for (auto& m : p->meshes) {
    pyre_mesh_desc d{};
    d.positions = m.pos.data();
    d.indices = m.idx.data();
    d.normals = m.nrm.empty() ? nullptr : m.nrm.data();
    d.n_vertices = (uint32_t)(m.pos.size() / 3);
    d.n_triangles = (uint32_t)(m.idx.size() / 3);
    d.material_id = m.material;
    d.motion = m.motion.size() == 6 ? m.motion.data() : nullptr;
    p->mdesc.push_back(d);
  }
  p->mldesc.clear();
  for (auto& m : p->mesh_lights) {
    pyre_mesh_light d{};
    d.positions = m.pos.data();
    d.indices = m.idx.data();
    d.n_vertices = (uint32_t)(m.pos.size() / 3);
    d.n_triangles = (uint32_t)(m.idx.size() / 3);
    for (int k = 0; k < 3; ++k) d.emission[k] = m.emission[k];
    p->mldesc.push_back(d);
  }
  p->edesc.width = p->env_w;
  p->edesc.height = p->env_h;
  p->edesc.rgb = p->env_rgb.data();
  p->edesc.intensity = p->env_intensity;
}

void preset_scene_bounds(pyre_preset* p) {
  // Bounds3::EMPTY grown by every triangle's vertices (math.rs:37-77)
  float lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
  for (const auto& m : p->meshes)
    for (uint32_t v : m.idx)
      for (int a = 0; a < 3; ++a) {
        lo[a] = rmin(lo[a], m.pos[3 * (size_t)v + a]);
        hi[a] = rmax(hi[a], m.pos[3 * (size_t)v + a]);
      }
  for (int a = 0; a < 3; ++a) {
    p->frame_min[a] = lo[a];
    p->frame_max[a] = hi[a];
  }
}

extern "C" {

int pyre_gen_random_triangles(uint64_t seed, uint32_t n, float sigma, float* verts) {
  if (n && !verts) return PYRE_EINVAL;
  HostRng g(seed);
  for (uint32_t i = 0; i < n; ++i) {
    double c[3] = {g.uniform(-1, 1), g.uniform(-1, 1), g.uniform(-1, 1)};
    for (int k = 0; k < 3; ++k)
      for (int a = 0; a < 3; ++a) verts[9 * (size_t)i + 3 * k + a] = (float)(c[a] + (double)sigma * g.gauss());
  }
  return PYRE_OK;
}

int pyre_gen_random_rays(uint64_t seed, uint32_t n, uint32_t n_closest, pyre_ray* rays, pyre_ray32* rays32) {
  HostRng g(seed);
  for (uint32_t i = 0; i < n; ++i) {
    float o[3];
    for (int a = 0; a < 3; ++a) o[a] = (float)g.uniform(-1.5, 1.5);
    double z = 1.0 - 2.0 * g.f64();
    double phi = 2.0 * M_PI * g.f64();
    double r = std::sqrt(std::fmax(0.0, 1.0 - z * z));
    V3 d = normalize(v3((float)(r * std::cos(phi)), (float)(r * std::sin(phi)), (float)z));  // math.rs:19
    float tmax = i < n_closest ? kInf : (float)g.uniform(0.1, 3.0);
    if (rays) {
      pyre_ray& R = rays[i];
      R.o[0] = o[0];
      R.o[1] = o[1];
      R.o[2] = o[2];
      R.d[0] = d.x;
      R.d[1] = d.y;
      R.d[2] = d.z;
      R.t_min = 1e-4f;
      R.t_max = tmax;
      R.time = 0.0f;
    }
    if (rays32) {
      pyre_ray32& R = rays32[i];
      R.o[0] = o[0];
      R.o[1] = o[1];
      R.o[2] = o[2];
      R.t_min = 1e-4f;
      R.d[0] = d.x;
      R.d[1] = d.y;
      R.d[2] = d.z;
      R.t_max = tmax;
    }
  }
  return PYRE_OK;
}

// main.rs:365-496
int pyre_preset_cornell(int motion_blur, pyre_preset** out) {
  if (!out) return PYRE_EINVAL;
  pyre_preset* p = new (std::nothrow) pyre_preset();
  if (!p) return PYRE_ENOMEM;
  p->kind = PRESET_CORNELL;
  p->materials.push_back(lambert(splat(0.73f)));
  p->materials.push_back(lambert(v3(0.65f, 0.05f, 0.05f)));
  p->materials.push_back(lambert(v3(0.12f, 0.45f, 0.15f)));
  const float s = 1.0f;
  V3 floor_c[4] = {v3(-s, -s, -s), v3(s, -s, -s), v3(s, -s, s), v3(-s, -s, s)};
  V3 ceil_c[4] = {v3(-s, s, -s), v3(s, s, -s), v3(s, s, s), v3(-s, s, s)};
  V3 back_c[4] = {v3(-s, -s, -s), v3(s, -s, -s), v3(s, s, -s), v3(-s, s, -s)};
  V3 left_c[4] = {v3(-s, -s, -s), v3(-s, -s, s), v3(-s, s, s), v3(-s, s, -s)};
  V3 right_c[4] = {v3(s, -s, -s), v3(s, s, -s), v3(s, s, s), v3(s, -s, s)};
  p->meshes.push_back(quad_mesh(floor_c, v3(0, 1, 0), 0));
  p->meshes.push_back(quad_mesh(ceil_c, -v3(0, 1, 0), 0));
  p->meshes.push_back(quad_mesh(back_c, v3(0, 0, 1), 0));
  p->meshes.push_back(quad_mesh(left_c, v3(1, 0, 0), 1));
  p->meshes.push_back(quad_mesh(right_c, -v3(1, 0, 0), 2));
  p->materials.push_back(disney(v3(1.0f, 0.766f, 0.336f), 1.0f, 0.1f, 0.5f));
  p->meshes.push_back(make_uv_sphere(v3(0.35f, -0.65f, 0.25f), 0.35f, 48, 24, 3));
  if (motion_blur) p->meshes.back().motion = {0.0f, 0.0f, 0.0f, 0.0f, 0.25f, 0.0f};  // main.rs:455-465
  p->materials.push_back(disney(splat(0.72f), 0.0f, 0.18f, 0.5f));
  p->meshes.push_back(make_uv_sphere(v3(-0.35f, -0.65f, -0.3f), 0.35f, 48, 24, 4));
  p->lights.push_back(pyre_quad_light{{-0.3f, 0.999f, -0.3f}, {0.6f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.6f}, {15, 15, 15}});
  preset_finish(p);
  *out = p;
  return PYRE_OK;
}

int pyre_preset_studio(int motion_blur, float env_intensity, pyre_preset** out) {
  if (!out) return PYRE_EINVAL;
  pyre_preset* p = new (std::nothrow) pyre_preset();
  if (!p) return PYRE_ENOMEM;
  p->kind = PRESET_STUDIO;
  // studio_scene (main.rs:502-562)
  p->materials.push_back(lambert(splat(0.55f)));
  const float s = 50.0f;
  V3 ground[4] = {v3(-s, 0.0f, -s), v3(s, 0.0f, -s), v3(s, 0.0f, s), v3(-s, 0.0f, s)};
  p->meshes.push_back(quad_mesh(ground, v3(0, 1, 0), 0));
  p->materials.push_back(disney(v3(1.0f, 0.766f, 0.336f), 1.0f, 0.1f, 0.5f));
  p->meshes.push_back(make_uv_sphere(v3(0.4f, 0.35f, 0.0f), 0.35f, 64, 32, 1));
  if (motion_blur) p->meshes.back().motion = {-0.15f, 0.0f, 0.0f, 0.15f, 0.0f, 0.0f};
  p->materials.push_back(disney(splat(0.72f), 0.0f, 0.2f, 0.5f));
  p->meshes.push_back(make_uv_sphere(v3(-0.4f, 0.35f, 0.0f), 0.35f, 64, 32, 2));
  // the CLI's sky for the studio without an HDRI (main.rs:167-174):
  // HdriEnvironmentLight::gradient (light.rs:252-272), 128 x 64
  const V3 zenith = v3(0.20f, 0.35f, 0.65f), horizon = v3(0.95f, 0.85f, 0.70f);
  p->env_w = 128;
  p->env_h = 64;
  p->env_intensity = env_intensity;
  p->env_rgb.resize(3 * 128 * 64);
  for (uint32_t y = 0; y < 64; ++y) {
    const float v = ((float)y + 0.5f) / 64.0f;
    const float cos_theta = 1.0f - 2.0f * v;
    const V3 c = cos_theta >= 0.0f ? lerp(horizon, zenith, cos_theta) : horizon * rmax(1.0f + cos_theta * 0.5f, 0.0f);
    for (uint32_t x = 0; x < 128; ++x) {
      float* q = p->env_rgb.data() + 3 * ((size_t)y * 128 + x);
      q[0] = c.x;
      q[1] = c.y;
      q[2] = c.z;
    }
  }
  preset_finish(p);
  *out = p;
  return PYRE_OK;
}

int pyre_preset_production(uint64_t seed, uint32_t n_instances, uint32_t level, uint32_t env_w, uint32_t env_h,
                           pyre_preset** out) {
  return pyre_preset_instances(seed, n_instances, level, env_w, env_h, 0.0f, out);
}

int pyre_preset_instances(uint64_t seed, uint32_t n_instances, uint32_t level, uint32_t env_w, uint32_t env_h,
                          float lambert_fraction, pyre_preset** out) {
  if (!out || n_instances == 0 || level > 9 || env_w == 0 || env_h == 0) return PYRE_EINVAL;
  if (!(lambert_fraction >= 0.0f && lambert_fraction <= 1.0f)) return PYRE_EINVAL;
  pyre_preset* p = new (std::nothrow) pyre_preset();
  if (!p) return PYRE_ENOMEM;
  HostRng g(seed);
  std::vector<double> P;
  std::vector<uint32_t> F;
  icosphere(level, P, F);
  const size_t nv = P.size() / 3;
  double bmin[3] = {1e300, 1e300, 1e300}, bmax[3] = {-1e300, -1e300, -1e300};
  for (uint32_t i = 0; i < n_instances; ++i) {
    double s = g.uniform(0.2, 1.0);
    // uniform rotation (Shoemake 1992)
    double u1 = g.f64(), u2 = g.f64(), u3 = g.f64();
    double qx = std::sqrt(1 - u1) * std::sin(2 * M_PI * u2), qy = std::sqrt(1 - u1) * std::cos(2 * M_PI * u2);
    double qz = std::sqrt(u1) * std::sin(2 * M_PI * u3), qw = std::sqrt(u1) * std::cos(2 * M_PI * u3);
    double R[3][3] = {{1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)},
                      {2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)},
                      {2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)}};
    double tr[3] = {g.uniform(-10, 10), g.uniform(-10, 10), g.uniform(-10, 10)};
    double base[3] = {g.f64(), g.f64(), g.f64()};
    double rough = g.uniform(0.05, 1.0);
    double metal = g.f64() < 0.3 ? 1.0 : 0.0;
    const V3 bc = v3((float)base[0], (float)base[1], (float)base[2]);
    // the Lambertian draw is taken only when asked for, so the Disney-only
    // scene keeps its random stream
    if (lambert_fraction > 0.0f && g.f64() < lambert_fraction)
      p->materials.push_back(lambert(bc));
    else
      p->materials.push_back(disney(bc, (float)metal, (float)rough, 0.5f));
    Mesh m;
    m.material = i;
    m.pos.resize(3 * nv);
    m.nrm.resize(3 * nv);
    for (size_t v = 0; v < nv; ++v) {
      const double* q = &P[3 * v];
      for (int a = 0; a < 3; ++a) {
        double rq = R[a][0] * q[0] + R[a][1] * q[1] + R[a][2] * q[2];
        double w = s * rq + tr[a];
        m.pos[3 * v + a] = (float)w;
        m.nrm[3 * v + a] = (float)rq;  // inverse-transpose of a similarity = rotation
        bmin[a] = std::fmin(bmin[a], (double)m.pos[3 * v + a]);
        bmax[a] = std::fmax(bmax[a], (double)m.pos[3 * v + a]);
      }
    }
    m.idx = F;
    p->meshes.push_back(std::move(m));
  }
  for (int a = 0; a < 3; ++a) {
    p->frame_min[a] = (float)bmin[a];
    p->frame_max[a] = (float)bmax[a];
  }
  // ground quad under the instances, twice their horizontal extent
  uint32_t gmat = (uint32_t)p->materials.size();
  p->materials.push_back(lambert(splat(0.55f)));
  double cx = 0.5 * (bmin[0] + bmax[0]), cz = 0.5 * (bmin[2] + bmax[2]);
  double ext = std::fmax(bmax[0] - bmin[0], bmax[2] - bmin[2]);
  float y = (float)bmin[1];
  float x0 = (float)(cx - ext), x1 = (float)(cx + ext), z0 = (float)(cz - ext), z1 = (float)(cz + ext);
  V3 gc[4] = {v3(x0, y, z0), v3(x1, y, z0), v3(x1, y, z1), v3(x0, y, z1)};
  p->meshes.push_back(quad_mesh(gc, v3(0, 1, 0), gmat));
  // sky gradient (light.rs:252-272 with the studio colours of main.rs:170-174)
  // plus a Gaussian sun lobe
  const V3 zenith = v3(0.20f, 0.35f, 0.65f), horizon = v3(0.95f, 0.85f, 0.70f);
  const double zl = 0.2126 * zenith.x + 0.7152 * zenith.y + 0.0722 * zenith.z;
  const double sun_peak = 1e4 * zl, sigma = 0.5 * M_PI / 180.0;
  double sel = g.uniform(20.0, 60.0) * M_PI / 180.0, saz = g.uniform(0.0, 2.0 * M_PI);
  double sd[3] = {std::cos(sel) * std::cos(saz), std::sin(sel), std::cos(sel) * std::sin(saz)};
  p->env_w = env_w;
  p->env_h = env_h;
  p->env_rgb.resize(3 * (size_t)env_w * env_h);
  for (uint32_t yy = 0; yy < env_h; ++yy) {
    float v = ((float)yy + 0.5f) / (float)env_h;
    float cos_theta = 1.0f - 2.0f * v;
    V3 color = cos_theta >= 0.0f ? lerp(horizon, zenith, cos_theta) : horizon * rmax(1.0f + cos_theta * 0.5f, 0.0f);
    const double dth = M_PI / env_h, dph = 2.0 * M_PI / env_w;
    for (uint32_t xx = 0; xx < env_w; ++xx) {
      // sun lobe averaged over the texel (8x8 subsamples near the sun, so the
      // sun's energy does not depend on the map resolution)
      auto lobe = [&](double th, double ph) {
        double d[3] = {std::sin(th) * std::cos(ph), std::cos(th), std::sin(th) * std::sin(ph)};
        double c = d[0] * sd[0] + d[1] * sd[1] + d[2] * sd[2];
        double ang = std::acos(std::fmax(-1.0, std::fmin(1.0, c)));
        return sun_peak * std::exp(-0.5 * (ang / sigma) * (ang / sigma));
      };
      double th0 = (double)yy * dth, ph0 = (double)xx * dph - M_PI;  // light.rs:282-287 convention
      double centre_d[3] = {std::sin(th0 + 0.5 * dth) * std::cos(ph0 + 0.5 * dph), std::cos(th0 + 0.5 * dth),
                            std::sin(th0 + 0.5 * dth) * std::sin(ph0 + 0.5 * dph)};
      double cang = std::acos(std::fmax(-1.0, std::fmin(1.0, centre_d[0] * sd[0] + centre_d[1] * sd[1] +
                                                                 centre_d[2] * sd[2])));
      double sun = 0.0;
      if (cang < 6.0 * sigma + dth + dph) {
        double wsum = 0.0;
        for (int a = 0; a < 8; ++a)
          for (int b = 0; b < 8; ++b) {
            double th = th0 + (a + 0.5) / 8.0 * dth, ph = ph0 + (b + 0.5) / 8.0 * dph;
            double w = std::sin(th);
            sun += w * lobe(th, ph);
            wsum += w;
          }
        sun /= wsum;
      }
      float* o = &p->env_rgb[3 * ((size_t)yy * env_w + xx)];
      o[0] = (float)(color.x + sun);
      o[1] = (float)(color.y + sun * 0.95);
      o[2] = (float)(color.z + sun * 0.85);
    }
  }
  preset_finish(p);
  *out = p;
  return PYRE_OK;
}

int pyre_preset_desc(const pyre_preset* p, pyre_scene_desc* d) {
  if (!p || !d) return PYRE_EINVAL;
  d->meshes = p->mdesc.data();
  d->n_meshes = (uint32_t)p->mdesc.size();
  d->materials = p->materials.data();
  d->n_materials = (uint32_t)p->materials.size();
  d->lights = p->lights.data();
  d->n_lights = (uint32_t)p->lights.size();
  d->mesh_lights = p->mldesc.empty() ? nullptr : p->mldesc.data();
  d->n_mesh_lights = (uint32_t)p->mldesc.size();
  d->env = p->env_w ? &p->edesc : nullptr;
  return PYRE_OK;
}

void pyre_auto_frame(const float bmin[3], const float bmax[3], float vfov_deg, float aspect, float origin[3],
                     float target[3]) {
  V3 lo = v3(bmin[0], bmin[1], bmin[2]), hi = v3(bmax[0], bmax[1], bmax[2]);
  V3 center = 0.5f * (lo + hi);  // math.rs:74-76
  float radius = length(hi - lo) * 0.5f;
  radius = rmax(radius, 1e-3f);
  const float rads_per_deg = kPi / 180.0f;
  float half_vfov = vfov_deg * rads_per_deg * 0.5f;
  float half_hfov = atanf(aspect * tanf(half_vfov));
  float dist_v = radius / sinf(half_vfov);
  float dist_h = radius / sinf(half_hfov);
  float distance = rmax(dist_v, dist_h) * 1.15f;
  V3 dir = normalize(v3(0.0f, 0.3f, 1.0f));
  V3 o = center + distance * dir;
  origin[0] = o.x;
  origin[1] = o.y;
  origin[2] = o.z;
  target[0] = center.x;
  target[1] = center.y;
  target[2] = center.z;
}

int pyre_preset_camera(const pyre_preset* p, float aspect, float aperture, pyre_camera_desc* cam) {
  if (!p || !cam) return PYRE_EINVAL;
  std::memset(cam, 0, sizeof(*cam));
  cam->up[1] = 1.0f;
  cam->vfov_deg = 45.0f;
  cam->aspect = aspect;
  cam->aperture_radius = aperture;
  cam->focus_distance = 0.0f;  // |target - origin|
  if (p->kind == PRESET_CORNELL) {
    cam->origin[2] = 3.0f;
  } else if (p->kind == PRESET_STUDIO) {  // main.rs:195-198
    const float o[3] = {1.6f, 0.5f, 2.6f}, t[3] = {0.0f, 0.05f, 0.0f};
    for (int a = 0; a < 3; ++a) {
      cam->origin[a] = o[a];
      cam->target[a] = t[a];
    }
    cam->vfov_deg = 38.0f;
  } else if (p->kind == PRESET_INSTANCES) {
    pyre_auto_frame(p->frame_min, p->frame_max, 45.0f, aspect, cam->origin, cam->target);
  } else if (p->view.has_camera) {  // main.rs:184-185
    for (int a = 0; a < 3; ++a) {
      cam->origin[a] = p->view.origin[a];
      cam->target[a] = p->view.target[a];
      cam->up[a] = p->view.up[a];
    }
    cam->vfov_deg = p->view.vfov_deg;
  } else {  // main.rs:186-194: framed over the scene bounds, or (0,0,3) -> 0
    const V3 d = v3(p->frame_max[0], p->frame_max[1], p->frame_max[2]) -
                 v3(p->frame_min[0], p->frame_min[1], p->frame_min[2]);
    const float diag = length(d);
    if (std::isfinite(diag) && diag > 0.0f)
      pyre_auto_frame(p->frame_min, p->frame_max, 45.0f, aspect, cam->origin, cam->target);
    else
      cam->origin[2] = 3.0f;
  }
  return PYRE_OK;
}

int pyre_preset_view(const pyre_preset* p, pyre_view* view) {
  if (!p || !view) return PYRE_EINVAL;
  *view = p->view;
  return PYRE_OK;
}

void pyre_preset_destroy(pyre_preset* p) { delete p; }

}  // extern "C"
