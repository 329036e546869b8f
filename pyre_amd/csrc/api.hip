// C ABI of the gfx950 backend: context, scene upload, BVH build/export,
// batched tracing and the shade-stage probes. The integrator's entry point
// (pyre_render) lives in integrator.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>

#include "internal.hpp"
#include "shade.hpp"

using namespace pyre;

namespace {

template <class F>
int guarded(pyre_ctx* ctx, F&& f) {
  try {
    if (ctx) PYRE_HIP_CHECK(hipSetDevice(ctx->device));
    f();
    if (ctx) ctx->err.clear();
    return PYRE_OK;
  } catch (const Error& e) {
    if (ctx) ctx->err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    if (ctx) ctx->err = "host allocation failed";
    return PYRE_ENOMEM;
  } catch (const std::exception& e) {
    if (ctx) ctx->err = e.what();
    return PYRE_EINVAL;
  }
}

inline unsigned blocks(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }
inline V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }

// Gather one mesh's indexed vertices (and normals) into scene-global
// per-triangle records.
__global__ void k_gather(const float* pos, const uint32_t* idx, const float* nrm, uint32_t nt, uint32_t tri0,
                         float* verts, float* normals) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  size_t g = (size_t)tri0 + t;
  for (int k = 0; k < 3; ++k) {
    uint32_t vi = idx[3 * (size_t)t + k];
    for (int c = 0; c < 3; ++c) verts[9 * g + 3 * k + c] = pos[3 * (size_t)vi + c];
    if (normals) {
      for (int c = 0; c < 3; ++c) normals[9 * g + 3 * k + c] = nrm ? nrm[3 * (size_t)vi + c] : 0.0f;
    }
  }
}

// Scene bounds over nv vertices: per-thread grid-stride min/max, then per
// wave and per block, one set of atomics per block.
__global__ __launch_bounds__(256) void k_bounds(const float* verts, size_t nv, int32_t* keys) {
  __shared__ int32_t part[4][6];
  int32_t lo[3] = {f2key(kInf), f2key(kInf), f2key(kInf)};
  int32_t hi[3] = {f2key(-kInf), f2key(-kInf), f2key(-kInf)};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (size_t)gridDim.x * blockDim.x)
    for (int c = 0; c < 3; ++c) {
      const int32_t k = f2key(verts[3 * i + c]);
      lo[c] = min(lo[c], k);
      hi[c] = max(hi[c], k);
    }
  for (int c = 0; c < 3; ++c) {
    for (int o = 32; o > 0; o >>= 1) {
      lo[c] = min(lo[c], __shfl_xor(lo[c], o));
      hi[c] = max(hi[c], __shfl_xor(hi[c], o));
    }
  }
  const int wave = threadIdx.x >> 6;
  if (__lane_id() == 0)
    for (int c = 0; c < 3; ++c) {
      part[wave][c] = lo[c];
      part[wave][3 + c] = hi[c];
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
      for (int c = 0; c < 3; ++c) {
        lo[c] = min(lo[c], part[w][c]);
        hi[c] = max(hi[c], part[w][3 + c]);
      }
    for (int c = 0; c < 3; ++c) {
      atomicMin(&keys[c], lo[c]);
      atomicMax(&keys[3 + c], hi[c]);
    }
  }
}

__global__ void k_bounds_init(int32_t* keys) {
  for (int c = 0; c < 3; ++c) {
    keys[c] = f2key(kInf);
    keys[3 + c] = f2key(-kInf);
  }
}

// Conditional CDFs, one thread per row, in the reference's sequential f32
// order (distribution.rs:39-62 over the weights of light.rs:223-233).
__global__ void k_env_rows(const float4* tex, const float* sin_theta, uint32_t w, uint32_t h, float* cond,
                           float* row_integral) {
  uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= h) return;
  float* cdf = cond + (size_t)y * (w + 1);
  float st = sin_theta[y];
  cdf[0] = 0.0f;
  float acc = 0.0f;
  for (uint32_t x = 0; x < w; ++x) {
    float4 p = tex[(size_t)y * w + x];
    float lum = 0.21267f * p.x + 0.71516f * p.y + 0.07217f * p.z;
    float wgt = rmax(lum, 0.0f) * st;
    acc += rmax(wgt, 0.0f);
    cdf[x + 1] = acc;
  }
  row_integral[y] = acc;
  if (acc > 0.0f) {
    for (uint32_t x = 0; x <= w; ++x) cdf[x] /= acc;
  } else {
    for (uint32_t x = 0; x <= w; ++x) cdf[x] = (float)x / (float)w;
  }
}

__global__ void k_env_marginal(const float* row_integral, uint32_t h, float* marg) {
  marg[0] = 0.0f;
  float acc = 0.0f;
  for (uint32_t y = 0; y < h; ++y) {
    acc += rmax(row_integral[y], 0.0f);
    marg[y + 1] = acc;
  }
  if (acc > 0.0f) {
    for (uint32_t y = 0; y <= h; ++y) marg[y] /= acc;
  } else {
    for (uint32_t y = 0; y <= h; ++y) marg[y] = (float)y / (float)h;
  }
}

// Guide table of each CDF (rows of n + 1 entries, `stride` apart): one
// thread per (row, k) finds the first i in [1, n) with cdf[i] > k / kEnvGuide.
__global__ void k_env_guide(const float* cdf, uint32_t n, size_t stride, uint32_t rows, uint32_t* guide) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * (kEnvGuide + 1)) return;
  const uint32_t row = t / (kEnvGuide + 1), k = t % (kEnvGuide + 1);
  const float* c = cdf + (size_t)row * stride;
  const float x = (float)k / (float)kEnvGuide;
  uint32_t lo = 1, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) / 2;
    if (c[mid] <= x)
      lo = mid + 1;
    else
      hi = mid;
  }
  guide[t] = lo;
}

void scene_free(pyre_scene* s) {
  bvh_release(s);
  dev_free(s->d_verts);
  dev_free(s->d_normals);
  dev_free(s->d_motion);
  dev_free(s->d_tri_offset);
  dev_free(s->d_mesh_blk);
  dev_free(s->d_mesh_info);
  dev_free(s->d_mats);
  dev_free(s->d_lights);
  dev_free(s->d_mlights);
  dev_free(s->d_etris);
  dev_free(s->d_ecdf);
  dev_free(s->d_env_texels);
  dev_free(s->d_env_cond);
  dev_free(s->d_env_marg);
  dev_free(s->d_env_guide);
}

void upload_env(pyre_scene* s, const pyre_env_desc& e, hipStream_t st) {
  require(e.width > 0 && e.height > 0 && e.rgb, "environment needs width, height and pixels");
  const size_t n = (size_t)e.width * e.height;
  std::vector<float4> tex(n);
  for (size_t i = 0; i < n; ++i) tex[i] = make_float4(e.rgb[3 * i], e.rgb[3 * i + 1], e.rgb[3 * i + 2], 0.0f);
  // theta at the row centre and sin(theta).max(0) (light.rs:225-226), on the
  // host with the same libm call as the reference.
  std::vector<float> sin_theta(e.height);
  for (uint32_t y = 0; y < e.height; ++y) {
    float theta = ((float)y + 0.5f) / (float)e.height * kPi;
    sin_theta[y] = rmax(sinf(theta), 0.0f);
  }
  s->d_env_texels = dev_alloc<float4>(n);
  s->d_env_cond = dev_alloc<float>((size_t)e.height * (e.width + 1));
  s->d_env_marg = dev_alloc<float>(e.height + 1);
  float* d_sin = dev_alloc<float>(e.height);
  float* d_rows = dev_alloc<float>(e.height);
  PYRE_HIP_CHECK(hipMemcpyAsync(s->d_env_texels, tex.data(), n * sizeof(float4), hipMemcpyHostToDevice, st));
  PYRE_HIP_CHECK(hipMemcpyAsync(d_sin, sin_theta.data(), 4 * (size_t)e.height, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_env_rows, dim3(blocks(e.height, 64)), dim3(64), 0, st, s->d_env_texels, d_sin, e.width,
                     e.height, s->d_env_cond, d_rows);
  hipLaunchKernelGGL(k_env_marginal, dim3(1), dim3(1), 0, st, d_rows, e.height, s->d_env_marg);
  s->d_env_guide = dev_alloc<uint32_t>((size_t)(e.height + 1) * (kEnvGuide + 1));
  uint32_t* marg_guide = s->d_env_guide + (size_t)e.height * (kEnvGuide + 1);
  hipLaunchKernelGGL(k_env_guide, dim3(blocks((size_t)e.height * (kEnvGuide + 1), 256)), dim3(256), 0, st,
                     s->d_env_cond, e.width, (size_t)e.width + 1, e.height, s->d_env_guide);
  hipLaunchKernelGGL(k_env_guide, dim3(blocks(kEnvGuide + 1, 256)), dim3(256), 0, st, s->d_env_marg, e.height,
                     (size_t)0, 1u, marg_guide);
  PYRE_HIP_CHECK(hipStreamSynchronize(st));
  dev_free(d_sin);
  dev_free(d_rows);
  s->has_env = true;
  s->env.width = e.width;
  s->env.height = e.height;
  s->env.intensity = e.intensity;
  s->env.texels = s->d_env_texels;
  s->env.cond_cdf = s->d_env_cond;
  s->env.marg_cdf = s->d_env_marg;
  s->env.cond_guide = s->d_env_guide;
  s->env.marg_guide = marg_guide;
}

// ---------------------------------------------------------- probe kernels
__global__ void k_bsdf_eval(MatDev m, const float* wo, const float* wi, uint32_t n, float* f, float* pdf) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V3 o = v3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), w = v3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
  V3 r = bsdf_eval(m, o, w);
  f[3 * i] = r.x;
  f[3 * i + 1] = r.y;
  f[3 * i + 2] = r.z;
  pdf[i] = bsdf_pdf(m, o, w);
}

__global__ void k_bsdf_sample(MatDev m, const float* wo, const float* u, uint32_t n, PhiTable T, float* wi, float* f,
                              float* pdf, uint32_t* ok) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V3 o = v3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
  V3 w = splat(0.0f), fv = splat(0.0f);
  float p = 0.0f;
  bool good = bsdf_sample(m, o, V2{u[2 * i], u[2 * i + 1]}, T, &w, &fv, &p);
  wi[3 * i] = w.x;
  wi[3 * i + 1] = w.y;
  wi[3 * i + 2] = w.z;
  f[3 * i] = fv.x;
  f[3 * i + 1] = fv.y;
  f[3 * i + 2] = fv.z;
  pdf[i] = p;
  ok[i] = good ? 1u : 0u;
}

__global__ void k_env_sample(EnvDev e, const float* u, uint32_t n, float* wi, float* li, float* pdf) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V3 w, l;
  float p;
  env_sample(e, V2{u[2 * i], u[2 * i + 1]}, &w, &l, &p);
  wi[3 * i] = w.x;
  wi[3 * i + 1] = w.y;
  wi[3 * i + 2] = w.z;
  li[3 * i] = l.x;
  li[3 * i + 1] = l.y;
  li[3 * i + 2] = l.z;
  pdf[i] = p;
}

__global__ void k_env_pdf(EnvDev e, const float* wi, uint32_t n, float* pdf) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pdf[i] = env_pdf(e, v3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]));
}

__global__ void k_env_le(EnvDev e, const float* wi, uint32_t n, float* le) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V3 l = env_le(e, v3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]));
  le[3 * i] = l.x;
  le[3 * i + 1] = l.y;
  le[3 * i + 2] = l.z;
}

__global__ void k_light_sample(QuadLightDev L, const float* from, const float* u, uint32_t n, float* pos, float* wi,
                               float* dist, float* pdf, uint32_t* ok) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V3 p = splat(0.0f), w = splat(0.0f);
  float d = 0.0f, pd = 0.0f;
  bool good = quad_sample(L, v3(from[3 * i], from[3 * i + 1], from[3 * i + 2]), V2{u[2 * i], u[2 * i + 1]}, &p, &w,
                          &d, &pd);
  pos[3 * i] = p.x;
  pos[3 * i + 1] = p.y;
  pos[3 * i + 2] = p.z;
  wi[3 * i] = w.x;
  wi[3 * i + 1] = w.y;
  wi[3 * i + 2] = w.z;
  dist[i] = d;
  pdf[i] = pd;
  ok[i] = good ? 1u : 0u;
}

__global__ void k_light_pdf(QuadLightDev L, const float* from, const float* wi, uint32_t n, float* pdf) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    pdf[i] = quad_pdf(L, v3(from[3 * i], from[3 * i + 1], from[3 * i + 2]), v3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]));
}

// Converts host rays (36 B) to the device layout (32 B).
__global__ void k_pack_rays(const float* in, uint32_t n, float4* out, float* times) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* r = in + 9 * (size_t)i;
  out[2 * (size_t)i] = make_float4(r[0], r[1], r[2], r[6]);
  out[2 * (size_t)i + 1] = make_float4(r[3], r[4], r[5], r[7]);
  if (times) times[i] = r[8];
}

// Scene-global triangle id -> (primitive, local triangle).
__global__ void k_unpack_hits(const pyre_hit8* h8, const float2* uv, uint32_t n, const uint32_t* tri_offset,
                              uint32_t n_meshes, pyre_hit* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  pyre_hit o;
  pyre_hit8 h = h8[i];
  if (h.prim == PYRE_MISS) {
    o.inst = PYRE_MISS;
    o.tri = PYRE_MISS;
    o.t = 0.0f;
    o.u = o.v = 0.0f;
  } else {
    uint32_t lo = 0, hi = n_meshes;  // largest m with tri_offset[m] <= prim
    while (hi - lo > 1) {
      uint32_t mid = (lo + hi) / 2;
      if (tri_offset[mid] <= h.prim)
        lo = mid;
      else
        hi = mid;
    }
    o.inst = lo;
    o.tri = h.prim - tri_offset[lo];
    o.t = h.t;
    o.u = uv[i].x;
    o.v = uv[i].y;
  }
  out[i] = o;
}

template <class T>
struct DevVec {
  T* p = nullptr;
  explicit DevVec(size_t n) { p = dev_alloc<T>(n); }
  ~DevVec() { dev_free(p); }
};

template <class T>
void h2d(T* d, const T* h, size_t n, hipStream_t st) {
  if (n) PYRE_HIP_CHECK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, st));
}
template <class T>
void d2h(T* h, const T* d, size_t n, hipStream_t st) {
  if (n) PYRE_HIP_CHECK(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, st));
}

MatDev to_dev(const pyre_material& m) {
  require(m.kind == PYRE_MAT_LAMBERTIAN || m.kind == PYRE_MAT_DISNEY, "unknown material kind");
  MatDev d;
  d.kind = m.kind;
  d.base = ld3(m.base_color);
  d.metallic = m.metallic;
  d.roughness = m.roughness;
  d.specular = m.specular;
  return d;
}

}  // namespace

// The exception table of PhiTable (shade.hpp), built once per process from
// the host's libm: for every grid point k < 2^24, phi = (2 pi) * (k 2^-24) in
// f32 as the samplers form it; an exception where libm's sinf or cosf differs
// from the double evaluation rounded to f32, or where that double lies within
// 2^-40 (relative) of a float rounding boundary (so a device double that is
// off by a few ulps could round the other way).
struct HostPhiTable {
  std::vector<uint4> rank;
  std::vector<float2> val;
};

static bool near_boundary(double d, float f) {
  if (d == 0.0) return false;
  const double up = 0.5 * ((double)f + (double)std::nextafter(f, INFINITY));
  const double dn = 0.5 * ((double)f + (double)std::nextafter(f, -INFINITY));
  return std::min(std::fabs(d - up), std::fabs(d - dn)) < std::fabs(d) * 0x1p-40;
}

static bool phi_exception(uint32_t k, float* s_out, float* c_out) {
  const float phi = 2.0f * kPi * ((float)k * (1.0f / 16777216.0f));
  const float s = ::sinf(phi), c = ::cosf(phi);
  const double ds = ::sin((double)phi), dc = ::cos((double)phi);
  *s_out = s;
  *c_out = c;
  return __builtin_bit_cast(uint32_t, s) != __builtin_bit_cast(uint32_t, (float)ds) ||
         __builtin_bit_cast(uint32_t, c) != __builtin_bit_cast(uint32_t, (float)dc) ||
         near_boundary(ds, (float)ds) || near_boundary(dc, (float)dc);
}

static const HostPhiTable& host_phi_table() {
  static const HostPhiTable table = [] {
    constexpr uint32_t words = kPhiGrid / 64;
    std::vector<uint64_t> bits(words, 0);
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto par = [&](auto&& body) {
      std::vector<std::thread> th;
      for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
          for (uint32_t w = t; w < words; w += nt) body(w);
        });
      for (auto& x : th) x.join();
    };
    par([&](uint32_t w) {
      uint64_t b = 0;
      float s, c;
      for (uint32_t j = 0; j < 64; ++j)
        if (phi_exception(w * 64 + j, &s, &c)) b |= 1ull << j;
      bits[w] = b;
    });
    HostPhiTable T;
    T.rank.resize(words);
    uint32_t before = 0;
    for (uint32_t w = 0; w < words; ++w) {
      T.rank[w] = make_uint4((uint32_t)bits[w], (uint32_t)(bits[w] >> 32), before, 0u);
      before += (uint32_t)__builtin_popcountll(bits[w]);
    }
    T.val.resize(std::max(1u, before));
    par([&](uint32_t w) {
      uint32_t i = T.rank[w].z;
      for (uint32_t j = 0; j < 64; ++j)
        if ((bits[w] >> j) & 1ull) {
          float s, c;
          phi_exception(w * 64 + j, &s, &c);
          T.val[i++] = make_float2(s, c);
        }
    });
    return T;
  }();
  return table;
}

// Emissive meshes (pyre_mesh_light): per triangle v0, e1 = v1 - v0,
// e2 = v2 - v0 (the f32 operations of mesh.rs:154-155), normalize(e1 x e2)
// and the weight 0.5 |e1 x e2|; per mesh Distribution1D::new over the
// weights (distribution.rs:39-62). A mesh whose total area is not a positive
// finite number is refused (its pdf would divide by it).
void upload_mesh_lights(pyre_scene* s, const pyre_scene_desc& d, hipStream_t st) {
  s->n_mlights = d.n_mesh_lights;
  if (!d.n_mesh_lights) return;
  require(d.mesh_lights != nullptr, "mesh_lights pointer is null");
  std::vector<float4> tris;
  std::vector<float> cdf;
  size_t total = 0;
  for (uint32_t i = 0; i < d.n_mesh_lights; ++i) {
    const pyre_mesh_light& m = d.mesh_lights[i];
    require(m.n_triangles > 0 && m.positions && m.indices, "mesh light without triangles");
    total += m.n_triangles;
    require(total <= PYRE_MESH_LIGHT_MAX_TRIS, "mesh lights hold more than PYRE_MESH_LIGHT_MAX_TRIS triangles");
    MeshLightDev L{};
    L.tri0 = (uint32_t)(tris.size() / 4);
    L.ntri = m.n_triangles;
    L.cdf0 = (uint32_t)cdf.size();
    L.emission = v3(m.emission[0], m.emission[1], m.emission[2]);
    std::vector<float> w(m.n_triangles);
    V3 lo = v3(kInf, kInf, kInf), hi = v3(-kInf, -kInf, -kInf);
    for (uint32_t t = 0; t < m.n_triangles; ++t) {
      uint32_t ia = m.indices[3 * t], ib = m.indices[3 * t + 1], ic = m.indices[3 * t + 2];
      require(ia < m.n_vertices && ib < m.n_vertices && ic < m.n_vertices, "mesh light index out of range");
      V3 a = ld3(m.positions + 3 * (size_t)ia), b = ld3(m.positions + 3 * (size_t)ib),
         c = ld3(m.positions + 3 * (size_t)ic);
      V3 e1 = b - a, e2 = c - a;
      V3 cr = cross(e1, e2);
      V3 n = normalize(cr);
      for (const V3& q : {a, b, c}) {
        lo = v3(std::fmin(lo.x, q.x), std::fmin(lo.y, q.y), std::fmin(lo.z, q.z));
        hi = v3(std::fmax(hi.x, q.x), std::fmax(hi.y, q.y), std::fmax(hi.z, q.z));
      }
      tris.push_back(make_float4(a.x, a.y, a.z, 0.0f));
      tris.push_back(make_float4(e1.x, e1.y, e1.z, 0.0f));
      tris.push_back(make_float4(e2.x, e2.y, e2.z, 0.0f));
      tris.push_back(make_float4(n.x, n.y, n.z, 0.0f));
      w[t] = 0.5f * length(cr);
    }
    float acc = 0.0f;
    cdf.push_back(0.0f);
    for (float x : w) {
      acc += std::fmax(x, 0.0f);
      cdf.push_back(acc);
    }
    require(acc > 0.0f && std::isfinite(acc), "mesh light with no area");
    require(std::isfinite(lo.x) && std::isfinite(lo.y) && std::isfinite(lo.z) && std::isfinite(hi.x) &&
                std::isfinite(hi.y) && std::isfinite(hi.z),
            "mesh light with a non-finite vertex");
    // A hit that Moller-Trumbore accepts lies within rounding of a triangle;
    // the pad (1e-4 of the box's size and of its coordinates' magnitude) is
    // orders of magnitude wider than that rounding and than the slab test's.
    {
      const float ext = std::fmax(hi.x - lo.x, std::fmax(hi.y - lo.y, hi.z - lo.z));
      const float mag = std::fmax(std::fmax(std::fabs(lo.x), std::fabs(hi.x)),
                                  std::fmax(std::fmax(std::fabs(lo.y), std::fabs(hi.y)),
                                            std::fmax(std::fabs(lo.z), std::fabs(hi.z))));
      const float pad = 1e-4f * (ext + mag) + 1e-20f;
      L.bmin = v3(lo.x - pad, lo.y - pad, lo.z - pad);
      L.bmax = v3(hi.x + pad, hi.y + pad, hi.z + pad);
    }
    for (size_t k = L.cdf0; k < cdf.size(); ++k) cdf[k] /= acc;
    L.area = acc;
    s->h_mlights.push_back(L);
  }
  s->n_etris = (uint32_t)total;
  s->d_mlights = dev_alloc<MeshLightDev>(s->h_mlights.size());
  h2d(s->d_mlights, s->h_mlights.data(), s->h_mlights.size(), st);
  s->d_etris = dev_alloc<float4>(tris.size());
  h2d(s->d_etris, tris.data(), tris.size(), st);
  s->d_ecdf = dev_alloc<float>(cdf.size());
  h2d(s->d_ecdf, cdf.data(), cdf.size(), st);
}

extern "C" {

int pyre_ctx_create(int device, pyre_ctx** out) {
  if (!out) return PYRE_EINVAL;
  *out = nullptr;
  pyre_ctx* c = new (std::nothrow) pyre_ctx();
  if (!c) return PYRE_ENOMEM;
  c->device = device;
  int rc = guarded(c, [&] {
    int n = 0;
    PYRE_HIP_CHECK(hipGetDeviceCount(&n));
    require(device >= 0 && device < n, "no such HIP device");
    PYRE_HIP_CHECK(hipSetDevice(device));
    hipDeviceProp_t prop;
    PYRE_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    c->name = prop.gcnArchName;
    c->cus = prop.multiProcessorCount;
    PYRE_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    noalloc_scratch_init(c);
    const HostPhiTable& T = host_phi_table();
    c->d_phi_rank = dev_alloc<uint4>(T.rank.size());
    c->d_phi_val = dev_alloc<float2>(T.val.size());
    PYRE_HIP_CHECK(hipMemcpy(c->d_phi_rank, T.rank.data(), T.rank.size() * sizeof(uint4), hipMemcpyHostToDevice));
    PYRE_HIP_CHECK(hipMemcpy(c->d_phi_val, T.val.data(), T.val.size() * sizeof(float2), hipMemcpyHostToDevice));
  });
  if (rc != PYRE_OK) {
    pyre_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return PYRE_OK;
}

void pyre_ctx_destroy(pyre_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  ctx->render_cache.reset();
  trace_scratch_release(ctx);
  dev_free(ctx->batch_acc);
  dev_free(ctx->batch_acc64);
  dev_free(ctx->d_phi_rank);
  dev_free(ctx->d_phi_val);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* pyre_last_error(const pyre_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int pyre_ctx_set_tuning(pyre_ctx* ctx, const pyre_tuning* t) {
  if (!ctx || !t) return PYRE_EINVAL;
  const bool ok = t->trace_blocks_per_cu <= 8 && (t->stack_ring == 0 || t->stack_ring == 4 || t->stack_ring == 8 ||
                                                  t->stack_ring == 16) &&
                  t->shade_waves <= 4 && t->traversal <= 4 && t->eager <= 1 && t->memory_fraction >= 0.0f &&
                  t->memory_fraction <= 1.0f;
  if (!ok) {
    ctx->err = "pyre_ctx_set_tuning: a field is out of range";
    return PYRE_EINVAL;
  }
  ctx->tune = *t;
  ctx->err.clear();
  return PYRE_OK;
}

int pyre_ctx_get_tuning(const pyre_ctx* ctx, pyre_tuning* t) {
  if (!ctx || !t) return PYRE_EINVAL;
  *t = ctx->tune;
  return PYRE_OK;
}

int pyre_ctx_device_info(pyre_ctx* ctx, char* name, size_t name_len, int* cu_count) {
  if (!ctx) return PYRE_EINVAL;
  if (name && name_len) {
    std::strncpy(name, ctx->name.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  if (cu_count) *cu_count = ctx->cus;
  return PYRE_OK;
}

int pyre_scene_upload(pyre_ctx* ctx, const pyre_scene_desc* desc, pyre_scene** out) {
  if (!ctx || !desc || !out) return PYRE_EINVAL;
  *out = nullptr;
  pyre_scene* s = new (std::nothrow) pyre_scene();
  if (!s) return PYRE_ENOMEM;
  s->ctx = ctx;
  int rc = guarded(ctx, [&] {
    hipStream_t st = ctx->stream;
    const pyre_scene_desc& d = *desc;
    require(d.n_meshes == 0 || d.meshes, "meshes pointer is null");
    require(d.n_materials == 0 || d.materials, "materials pointer is null");
    require(d.n_lights == 0 || d.lights, "lights pointer is null");
    s->n_meshes = d.n_meshes;
    s->n_materials = d.n_materials;
    s->n_lights = d.n_lights;
    s->h_tri_offset.assign(d.n_meshes + 1, 0);
    bool any_normals = false;
    size_t tri_total = 0, max_v = 0, max_t = 0;
    for (uint32_t m = 0; m < d.n_meshes; ++m) {
      const pyre_mesh_desc& md = d.meshes[m];
      require(md.n_triangles == 0 || (md.positions && md.indices), "mesh without positions or indices");
      require(md.material_id < d.n_materials, "mesh material_id out of range");
      for (size_t k = 0; k < 3 * (size_t)md.n_triangles; ++k)
        require(md.indices[k] < md.n_vertices, "mesh index out of range");
      s->h_tri_offset[m] = (uint32_t)tri_total;
      tri_total += md.n_triangles;
      require(tri_total < 0xFFFFFFF0ull, "too many triangles");
      max_v = std::max(max_v, (size_t)md.n_vertices);
      max_t = std::max(max_t, (size_t)md.n_triangles);
      any_normals |= md.normals != nullptr;
      s->h_mesh_info.push_back(md.material_id | (md.normals ? 0x80000000u : 0u));
      const float* mo = md.motion;
      if (mo)
        for (int k = 0; k < 6; ++k) require(std::isfinite(mo[k]), "mesh motion offsets must be finite");
      s->h_moving.push_back(mo ? 1 : 0);
      s->h_motion.push_back(mo ? make_float4(mo[0], mo[1], mo[2], 0.0f) : make_float4(0, 0, 0, 0));
      s->h_motion.push_back(mo ? make_float4(mo[3], mo[4], mo[5], 0.0f) : make_float4(0, 0, 0, 0));
      s->has_motion |= mo != nullptr;
    }
    s->h_tri_offset[d.n_meshes] = (uint32_t)tri_total;
    s->n_tris = (uint32_t)tri_total;
    s->d_verts = dev_alloc<float>(9 * tri_total);
    if (any_normals) s->d_normals = dev_alloc<float>(9 * tri_total);
    {
      DevVec<float> pos(3 * max_v), nrm(any_normals ? 3 * max_v : 1);
      DevVec<uint32_t> idx(3 * max_t);
      for (uint32_t m = 0; m < d.n_meshes; ++m) {
        const pyre_mesh_desc& md = d.meshes[m];
        if (md.n_triangles == 0) continue;
        h2d(pos.p, md.positions, 3 * (size_t)md.n_vertices, st);
        h2d(idx.p, md.indices, 3 * (size_t)md.n_triangles, st);
        if (md.normals) h2d(nrm.p, md.normals, 3 * (size_t)md.n_vertices, st);
        hipLaunchKernelGGL(k_gather, dim3(blocks(md.n_triangles, 256)), dim3(256), 0, st, pos.p, idx.p,
                           md.normals ? nrm.p : nullptr, md.n_triangles, s->h_tri_offset[m], s->d_verts,
                           s->d_normals);
        PYRE_HIP_CHECK(hipStreamSynchronize(st));  // staging buffers are reused
      }
    }
    s->d_tri_offset = dev_alloc<uint32_t>(d.n_meshes + 1);
    h2d(s->d_tri_offset, s->h_tri_offset.data(), d.n_meshes + 1, st);
    {
      // entry k: the largest m with tri_offset[m] <= k 2^kMeshBlkShift, so the
      // mesh of triangle t lies in [entry t >> shift, entry (t >> shift) + 1]
      const size_t nblk = (tri_total >> kMeshBlkShift) + 2;
      std::vector<uint32_t> blk(nblk, 0);
      uint32_t m = 0;
      for (size_t k = 0; k < nblk; ++k) {
        const size_t t = k << kMeshBlkShift;
        while (m + 1 < d.n_meshes && s->h_tri_offset[m + 1] <= t) ++m;
        blk[k] = m;
      }
      s->d_mesh_blk = dev_alloc<uint32_t>(nblk);
      h2d(s->d_mesh_blk, blk.data(), nblk, st);
      PYRE_HIP_CHECK(hipStreamSynchronize(st));  // blk is freed at the end of this scope
    }
    s->d_mesh_info = dev_alloc<uint32_t>(std::max<size_t>(d.n_meshes, 1));
    h2d(s->d_mesh_info, s->h_mesh_info.data(), d.n_meshes, st);
    if (s->has_motion) {
      s->d_motion = dev_alloc<float4>(s->h_motion.size());
      h2d(s->d_motion, s->h_motion.data(), s->h_motion.size(), st);
    }
    for (uint32_t i = 0; i < d.n_materials; ++i) s->h_mats.push_back(to_dev(d.materials[i]));
    s->d_mats = dev_alloc<MatDev>(std::max<size_t>(d.n_materials, 1));
    h2d(s->d_mats, s->h_mats.data(), d.n_materials, st);
    for (uint32_t i = 0; i < d.n_lights; ++i) {
      const pyre_quad_light& l = d.lights[i];
      s->h_lights.push_back(make_quad_light(ld3(l.p0), ld3(l.edge_u), ld3(l.edge_v), ld3(l.emission)));
    }
    s->d_lights = dev_alloc<QuadLightDev>(std::max<size_t>(d.n_lights, 1));
    h2d(s->d_lights, s->h_lights.data(), d.n_lights, st);
    upload_mesh_lights(s, d, st);
    if (d.env) upload_env(s, *d.env, st);
    // Scene::bounds over the referenced vertices
    DevVec<int32_t> keys(6);
    hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(1), 0, st, keys.p);
    if (tri_total)
      hipLaunchKernelGGL(k_bounds, dim3(std::min<unsigned>(blocks(3 * tri_total, 256), (unsigned)std::max(1, ctx->cus) * 8)),
                         dim3(256), 0, st, s->d_verts, 3 * tri_total,
                         keys.p);
    int32_t hk[6];
    d2h(hk, keys.p, 6, st);
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
    for (int c = 0; c < 3; ++c) {
      s->bmin[c] = key2f(hk[c]);
      s->bmax[c] = key2f(hk[3 + c]);
    }
    if (s->has_motion) {
      // Scene::bounds over MeshInstance::world_bounds (mesh.rs:97-107): a
      // moving mesh contributes its box at both ends of the shutter
      V3 lo = splat(kInf), hi = splat(-kInf);
      for (uint32_t m = 0; m < d.n_meshes; ++m) {
        const pyre_mesh_desc& md = d.meshes[m];
        V3 blo = splat(kInf), bhi = splat(-kInf);
        for (size_t k = 0; k < 3 * (size_t)md.n_triangles; ++k) {
          const float* v = md.positions + 3 * (size_t)md.indices[k];
          blo = vmin(blo, v3(v[0], v[1], v[2]));
          bhi = vmax(bhi, v3(v[0], v[1], v[2]));
        }
        if (md.n_triangles == 0) continue;
        if (md.motion) {
          const V3 o0 = v3(md.motion[0], md.motion[1], md.motion[2]), o1 = v3(md.motion[3], md.motion[4], md.motion[5]);
          lo = vmin(lo, vmin(blo + o0, blo + o1));
          hi = vmax(hi, vmax(bhi + o0, bhi + o1));
        } else {
          lo = vmin(lo, blo);
          hi = vmax(hi, bhi);
        }
      }
      const float l3[3] = {lo.x, lo.y, lo.z}, h3[3] = {hi.x, hi.y, hi.z};
      for (int c = 0; c < 3; ++c) {
        s->bmin[c] = l3[c];
        s->bmax[c] = h3[c];
      }
    }
  });
  if (rc != PYRE_OK) {
    scene_free(s);
    delete s;
    return rc;
  }
  *out = s;
  return PYRE_OK;
}

void pyre_scene_destroy(pyre_scene* s) {
  if (!s) return;
  (void)hipSetDevice(s->ctx->device);
  scene_free(s);
  delete s;
}

int pyre_scene_bounds(pyre_ctx* ctx, pyre_scene* s, float out_min[3], float out_max[3]) {
  if (!ctx || !s || !out_min || !out_max) return PYRE_EINVAL;
  for (int c = 0; c < 3; ++c) {
    out_min[c] = s->bmin[c];
    out_max[c] = s->bmax[c];
  }
  return PYRE_OK;
}

int pyre_bvh_build(pyre_ctx* ctx, pyre_scene* s, const pyre_bvh_params* p) {
  if (!ctx || !s) return PYRE_EINVAL;
  pyre_bvh_params def{12, 8, 1.0f, 2.0f};  // bvh.rs:13-16
  return guarded(ctx, [&] {
    bvh_build(s, p ? *p : def, ctx->stream);
    trace_scratch_reserve(ctx, s->depth);
  });
}

int pyre_bvh_info(pyre_ctx* ctx, pyre_scene* s, uint32_t* n_nodes, uint32_t* n_leaves, uint32_t* depth,
                  double* build_ms) {
  if (!ctx || !s) return PYRE_EINVAL;
  if (!s->built) {
    ctx->err = "pyre_bvh_info before pyre_bvh_build";
    return PYRE_ESTATE;
  }
  if (n_nodes) *n_nodes = s->n_bnodes;
  if (n_leaves) *n_leaves = s->n_leaves;
  if (depth) *depth = s->depth;
  if (build_ms) *build_ms = s->build_ms;
  return PYRE_OK;
}

int pyre_bvh_export(pyre_ctx* ctx, pyre_scene* s, float* bounds6, uint32_t* a, uint32_t* b, uint32_t* is_leaf,
                    uint32_t* axis, uint32_t* prim_order) {
  if (!ctx || !s) return PYRE_EINVAL;
  if (!s->built) {
    ctx->err = "pyre_bvh_export before pyre_bvh_build";
    return PYRE_ESTATE;
  }
  return guarded(ctx, [&] {
    std::vector<BNode> nodes(s->n_bnodes);
    d2h(nodes.data(), s->d_bnodes, s->n_bnodes, ctx->stream);
    if (prim_order) d2h(prim_order, s->d_prim_order, s->n_tris, ctx->stream);
    PYRE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (uint32_t i = 0; i < s->n_bnodes; ++i) {
      const BNode& nd = nodes[i];
      bool used = (nd.flags & 2u) != 0, leaf = (nd.flags & 1u) != 0;
      if (bounds6)
        for (int k = 0; k < 6; ++k) bounds6[6 * (size_t)i + k] = used ? key2f(nd.bkey[k]) : 0.0f;
      if (a) a[i] = !used ? PYRE_MISS : (leaf ? nd.start : (uint32_t)nd.left);
      if (b) b[i] = !used ? PYRE_MISS : (leaf ? nd.count : (uint32_t)nd.right);
      if (is_leaf) is_leaf[i] = !used ? 2u : (leaf ? 1u : 0u);
      if (axis) axis[i] = leaf ? 0u : nd.axis;
    }
  });
}

int pyre_intersect_closest(pyre_ctx* ctx, pyre_scene* s, const pyre_ray* rays, size_t n, pyre_hit* hits) {
  if (!ctx || !s || (n && (!rays || !hits))) return PYRE_EINVAL;
  return guarded(ctx, [&] {
    if (n == 0) return;
    hipStream_t st = ctx->stream;
    DevVec<float> raw(9 * n);
    DevVec<float4> r32(2 * n);
    DevVec<pyre_hit8> h8(n);
    DevVec<float2> uv(n);
    DevVec<pyre_hit> out(n);
    DevVec<float> times(s->has_motion ? n : 1);
    h2d(raw.p, (const float*)rays, 9 * n, st);
    hipLaunchKernelGGL(k_pack_rays, dim3(blocks(n, 256)), dim3(256), 0, st, raw.p, (uint32_t)n, r32.p,
                       s->has_motion ? times.p : nullptr);
    trace_closest(s, r32.p, n, h8.p, uv.p, st, true, s->has_motion ? times.p : nullptr);
    hipLaunchKernelGGL(k_unpack_hits, dim3(blocks(n, 256)), dim3(256), 0, st, h8.p, uv.p, (uint32_t)n,
                       s->d_tri_offset, s->n_meshes, out.p);
    d2h(hits, out.p, n, st);
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
  });
}

int pyre_intersect_any(pyre_ctx* ctx, pyre_scene* s, const pyre_ray* rays, size_t n, uint32_t* occluded) {
  if (!ctx || !s || (n && (!rays || !occluded))) return PYRE_EINVAL;
  return guarded(ctx, [&] {
    if (n == 0) return;
    hipStream_t st = ctx->stream;
    DevVec<float> raw(9 * n);
    DevVec<float4> r32(2 * n);
    DevVec<uint32_t> occ(n);
    DevVec<float> times(s->has_motion ? n : 1);
    h2d(raw.p, (const float*)rays, 9 * n, st);
    hipLaunchKernelGGL(k_pack_rays, dim3(blocks(n, 256)), dim3(256), 0, st, raw.p, (uint32_t)n, r32.p,
                       s->has_motion ? times.p : nullptr);
    trace_any(s, r32.p, n, occ.p, st, true, s->has_motion ? times.p : nullptr);
    d2h(occluded, occ.p, n, st);
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
  });
}

int pyre_trace_closest_dev(pyre_ctx* ctx, pyre_scene* s, const void* rays, size_t n, void* hits, void* stream) {
  if (!ctx || !s || (n && (!rays || !hits))) return PYRE_EINVAL;
  return guarded(ctx, [&] {
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    trace_closest(s, (const float4*)rays, n, (pyre_hit8*)hits, nullptr, st, false);
  });
}

int pyre_trace_closest_dev_noalloc(pyre_ctx* ctx, pyre_scene* s, const void* rays, size_t n, void* hits,
                                   void* stream) {
  if (!ctx || !s || (n && (!rays || !hits))) return PYRE_EINVAL;
  return guarded(ctx, [&] {
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    trace_closest_noalloc(s, (const float4*)rays, n, (pyre_hit8*)hits, st);
  });
}

int pyre_trace_any_dev(pyre_ctx* ctx, pyre_scene* s, const void* rays, size_t n, void* occluded, void* stream) {
  if (!ctx || !s || (n && (!rays || !occluded))) return PYRE_EINVAL;
  return guarded(ctx, [&] {
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    trace_any(s, (const float4*)rays, n, (uint32_t*)occluded, st, false);
  });
}

int pyre_bsdf_eval_batch(pyre_ctx* ctx, const pyre_material* m, const float* wo, const float* wi, size_t n, float* f,
                         float* pdf) {
  if (!ctx || !m || (n && (!wo || !wi || !f || !pdf))) return PYRE_EINVAL;
  return guarded(ctx, [&] {
    if (!n) return;
    hipStream_t st = ctx->stream;
    MatDev md = to_dev(*m);
    DevVec<float> a(3 * n), b(3 * n), fo(3 * n), po(n);
    h2d(a.p, wo, 3 * n, st);
    h2d(b.p, wi, 3 * n, st);
    hipLaunchKernelGGL(k_bsdf_eval, dim3(blocks(n, 256)), dim3(256), 0, st, md, a.p, b.p, (uint32_t)n, fo.p, po.p);
    d2h(f, fo.p, 3 * n, st);
    d2h(pdf, po.p, n, st);
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
  });
}

int pyre_bsdf_sample_batch(pyre_ctx* ctx, const pyre_material* m, const float* wo, const float* u, size_t n,
                           float* wi, float* f, float* pdf, uint32_t* ok) {
  if (!ctx || !m || (n && (!wo || !u || !wi || !f || !pdf || !ok))) return PYRE_EINVAL;
  return guarded(ctx, [&] {
    if (!n) return;
    hipStream_t st = ctx->stream;
    MatDev md = to_dev(*m);
    DevVec<float> a(3 * n), uu(2 * n), wo_(3 * n), fo(3 * n), po(n);
    DevVec<uint32_t> okd(n);
    h2d(a.p, wo, 3 * n, st);
    h2d(uu.p, u, 2 * n, st);
    PhiTable T;
    T.rank = ctx->d_phi_rank;
    T.val = ctx->d_phi_val;
    hipLaunchKernelGGL(k_bsdf_sample, dim3(blocks(n, 256)), dim3(256), 0, st, md, a.p, uu.p, (uint32_t)n, T, wo_.p,
                       fo.p, po.p, okd.p);
    d2h(wi, wo_.p, 3 * n, st);
    d2h(f, fo.p, 3 * n, st);
    d2h(pdf, po.p, n, st);
    d2h(ok, okd.p, n, st);
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
  });
}

static int need_env(pyre_ctx* ctx, pyre_scene* s) {
  if (!s->has_env) {
    ctx->err = "scene has no environment light";
    return PYRE_ESTATE;
  }
  return PYRE_OK;
}

int pyre_env_sample_batch(pyre_ctx* ctx, pyre_scene* s, const float* u, size_t n, float* wi, float* li, float* pdf) {
  if (!ctx || !s || (n && (!u || !wi || !li || !pdf))) return PYRE_EINVAL;
  if (int rc = need_env(ctx, s)) return rc;
  return guarded(ctx, [&] {
    if (!n) return;
    hipStream_t st = ctx->stream;
    DevVec<float> uu(2 * n), w(3 * n), l(3 * n), p(n);
    h2d(uu.p, u, 2 * n, st);
    hipLaunchKernelGGL(k_env_sample, dim3(blocks(n, 256)), dim3(256), 0, st, s->env, uu.p, (uint32_t)n, w.p, l.p, p.p);
    d2h(wi, w.p, 3 * n, st);
    d2h(li, l.p, 3 * n, st);
    d2h(pdf, p.p, n, st);
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
  });
}

int pyre_env_pdf_batch(pyre_ctx* ctx, pyre_scene* s, const float* wi, size_t n, float* pdf) {
  if (!ctx || !s || (n && (!wi || !pdf))) return PYRE_EINVAL;
  if (int rc = need_env(ctx, s)) return rc;
  return guarded(ctx, [&] {
    if (!n) return;
    hipStream_t st = ctx->stream;
    DevVec<float> w(3 * n), p(n);
    h2d(w.p, wi, 3 * n, st);
    hipLaunchKernelGGL(k_env_pdf, dim3(blocks(n, 256)), dim3(256), 0, st, s->env, w.p, (uint32_t)n, p.p);
    d2h(pdf, p.p, n, st);
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
  });
}

int pyre_env_le_batch(pyre_ctx* ctx, pyre_scene* s, const float* wi, size_t n, float* le) {
  if (!ctx || !s || (n && (!wi || !le))) return PYRE_EINVAL;
  if (int rc = need_env(ctx, s)) return rc;
  return guarded(ctx, [&] {
    if (!n) return;
    hipStream_t st = ctx->stream;
    DevVec<float> w(3 * n), l(3 * n);
    h2d(w.p, wi, 3 * n, st);
    hipLaunchKernelGGL(k_env_le, dim3(blocks(n, 256)), dim3(256), 0, st, s->env, w.p, (uint32_t)n, l.p);
    d2h(le, l.p, 3 * n, st);
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
  });
}

int pyre_light_sample_batch(pyre_ctx* ctx, pyre_scene* s, uint32_t light, const float* from, const float* u,
                            size_t n, float* pos, float* wi, float* dist, float* pdf, uint32_t* ok) {
  if (!ctx || !s || (n && (!from || !u || !pos || !wi || !dist || !pdf || !ok))) return PYRE_EINVAL;
  if (light >= s->n_lights) {
    ctx->err = "light index out of range";
    return PYRE_EINVAL;
  }
  return guarded(ctx, [&] {
    if (!n) return;
    hipStream_t st = ctx->stream;
    DevVec<float> fr(3 * n), uu(2 * n), po(3 * n), w(3 * n), d(n), p(n);
    DevVec<uint32_t> okd(n);
    h2d(fr.p, from, 3 * n, st);
    h2d(uu.p, u, 2 * n, st);
    hipLaunchKernelGGL(k_light_sample, dim3(blocks(n, 256)), dim3(256), 0, st, s->h_lights[light], fr.p, uu.p,
                       (uint32_t)n, po.p, w.p, d.p, p.p, okd.p);
    d2h(pos, po.p, 3 * n, st);
    d2h(wi, w.p, 3 * n, st);
    d2h(dist, d.p, n, st);
    d2h(pdf, p.p, n, st);
    d2h(ok, okd.p, n, st);
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
  });
}

int pyre_light_pdf_batch(pyre_ctx* ctx, pyre_scene* s, uint32_t light, const float* from, const float* wi, size_t n,
                         float* pdf) {
  if (!ctx || !s || (n && (!from || !wi || !pdf))) return PYRE_EINVAL;
  if (light >= s->n_lights) {
    ctx->err = "light index out of range";
    return PYRE_EINVAL;
  }
  return guarded(ctx, [&] {
    if (!n) return;
    hipStream_t st = ctx->stream;
    DevVec<float> fr(3 * n), w(3 * n), p(n);
    h2d(fr.p, from, 3 * n, st);
    h2d(w.p, wi, 3 * n, st);
    hipLaunchKernelGGL(k_light_pdf, dim3(blocks(n, 256)), dim3(256), 0, st, s->h_lights[light], fr.p, w.p,
                       (uint32_t)n, p.p);
    d2h(pdf, p.p, n, st);
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
  });
}

}  // extern "C"
