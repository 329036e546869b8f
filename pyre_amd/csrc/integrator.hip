// Wavefront path integrator for gfx950: the reference's per-sample loop
// (PathIntegrator::integrate, crates/pyre/src/integrator.rs:57-231, driven by
// crates/pyre-cli/src/main.rs:253-287) split into stages over path queues:
//
//   generate    one path per (pixel, sample): pixel_seed stream, jitter,
//               lens and time draws, camera ray (main.rs:261-273)
//   extend      closest hit for every queued path (Scene::intersect minus
//               the light loop, scene.rs:57-85)
//   shade       light-loop hit test (scene.rs:87-95), miss/emission with MIS,
//               AOV capture, NEE toward every quad light and the environment
//               (each becomes a shadow-ray slot carrying its contribution),
//               BSDF sampling, Russian roulette, next ray -> next queue
//   shadow      any-hit per slot (Scene::occluded / occluded_dir,
//               scene.rs:102-140); unoccluded contributions are added to the
//               path radiance in the reference's order (lights, then env)
//   accumulate  per-pixel f32 sums over samples in sample order
//               (main.rs:275-285), optional FP64 radiance sums
//
// Queues are compacted with wave64 ballots: one atomic per wave. Every
// stage is a persistent grid-stride kernel whose item count lives in device
// memory, so a whole bounce needs no host synchronisation. RNG draws happen
// in exactly the reference's order per path (jitter x/y, lens x/y, time; per
// surface bounce: 2 per light, 2 for the environment, 2 for the BSDF, 1 for
// Russian roulette), so a GPU path consumes the same stream as the CPU one.
#include <algorithm>
#include <vector>

#include "shade.hpp"
#include "traverse.hpp"

using namespace pyre;

namespace {

constexpr int kThreads = 256;
constexpr uint32_t FL_SPECULAR = 1u, FL_AOV = 2u;
constexpr float kRayEps = 1e-3f;  // integrator.rs:34

struct SceneDev {
  const Node4* nodes4;
  const float4* tris;
  const float* verts;
  const float* normals;
  const uint32_t* tri_offset;
  const uint32_t* mesh_info;
  const MatDev* mats;
  const QuadLightDev* lights;
  uint32_t n_meshes, n_lights, has_env;
  EnvDev env;
};

struct Paths {
  float4* ray;       // 2 per path: (o, t_min), (d, t_max)
  ulonglong2* rng;   // 2 per path
  float4* beta_pdf;  // beta, last_pdf_bsdf
  float4* lum;       // L, flags
  float4* aov0;      // albedo, depth
  float4* aov1;      // normal
  float4* hit;       // gid bits, t, u, v
  float4* sray;      // 2 per slot
  float4* scontrib;  // contribution per slot
  uint32_t* scount;
  uint32_t* q_cur;
  uint32_t* q_next;
  uint32_t* q_shadow;
  uint32_t* counters;  // [0] current queue, [1] next queue, [2] shadow queue
  unsigned long long* stats;  // [0] extension rays, [1] shadow rays (whole render)
  uint32_t slots;      // shadow slots per path (lights + env)
};

struct PassParams {
  uint32_t width, height, npix, s_base, n_samples, paths, max_depth, min_rr;
};

__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }

// Wave-aggregated queue append: every lane with `push` gets a distinct slot.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool push) {
  uint64_t m = __ballot(push);
  if (m == 0) return 0;
  int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if ((int)__lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  return base + (uint32_t)__popcll(m & lanemask_lt());
}

__device__ __forceinline__ Rng load_rng(const Paths& P, uint32_t p) {
  ulonglong2 a = P.rng[2 * (size_t)p], b = P.rng[2 * (size_t)p + 1];
  return Rng{a.x, a.y, b.x, b.y};
}
__device__ __forceinline__ void store_rng(const Paths& P, uint32_t p, const Rng& r) {
  P.rng[2 * (size_t)p] = make_ulonglong2(r.s0, r.s1);
  P.rng[2 * (size_t)p + 1] = make_ulonglong2(r.s2, r.s3);
}
__device__ __forceinline__ V3 xyz(float4 a) { return v3(a.x, a.y, a.z); }
__device__ __forceinline__ float4 f4(V3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

__global__ void k_counters(uint32_t* c, uint32_t a, uint32_t b, uint32_t d) {
  c[0] = a;
  c[1] = b;
  c[2] = d;
}

// next queue becomes current; next and shadow queues are emptied
__global__ void k_rotate(uint32_t* c, unsigned long long* stats) {
  stats[0] += c[0];
  c[0] = c[1];
  c[1] = 0;
  c[2] = 0;
}

// main.rs:261-273 + camera.rs + integrator.rs:63-74 state
__global__ __launch_bounds__(kThreads) void k_generate(Paths P, PassParams pp, CamDev cam) {
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t p = blockIdx.x * kThreads + threadIdx.x; p < pp.paths; p += stride) {
    uint32_t pixel = p % pp.npix, sl = p / pp.npix;
    uint32_t x = pixel % pp.width, y = pixel / pp.width;
    Rng rng = rng_seed(pixel_seed(x, y, pp.s_base + sl));
    float jx = rng_f32(rng);
    float jy = rng_f32(rng);
    float ndc_x = 2.0f * ((float)x + jx) / (float)pp.width - 1.0f;
    float ndc_y = 1.0f - 2.0f * ((float)y + jy) / (float)pp.height;
    V2 lens = rng_vec2(rng);
    (void)rng_f32(rng);  // shutter time (static scenes)
    V3 o, d;
    camera_ray(cam, V2{ndc_x, ndc_y}, lens, &o, &d);
    P.ray[2 * (size_t)p] = f4(o, 1e-4f);
    P.ray[2 * (size_t)p + 1] = f4(d, kInf);
    store_rng(P, p, rng);
    P.beta_pdf[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    P.lum[p] = make_float4(0.0f, 0.0f, 0.0f, __builtin_bit_cast(float, FL_SPECULAR));
    P.aov0[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    P.aov1[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    P.q_cur[p] = p;
  }
}

constexpr int kRing = 8;

__global__ __launch_bounds__(kThreads) void k_extend(Paths P, SceneDev S, int32_t* gstack) {
  __shared__ int32_t ring[kRing * kThreads];
  const uint32_t n = P.counters[0];
  const uint32_t tid = blockIdx.x * kThreads + threadIdx.x;
  const uint32_t stride = gridDim.x * kThreads;
  HybridStack<kRing> st{ring + threadIdx.x, kThreads, gstack + tid, stride, 0, 0};
  for (uint32_t i = tid; i < n; i += stride) {
    uint32_t p = P.q_cur[i];
    float4 a = P.ray[2 * (size_t)p], b = P.ray[2 * (size_t)p + 1];
    TRay r = make_tray(xyz(a), xyz(b), a.w, b.w);
    THit h;
    traverse4<false, kRing>(S.nodes4, S.tris, r, st, &h);
    P.hit[p] = make_float4(__builtin_bit_cast(float, h.gid), h.t, h.u, h.v);
  }
}

__global__ __launch_bounds__(kThreads) void k_shadow(Paths P, SceneDev S, int32_t* gstack) {
  __shared__ int32_t ring[kRing * kThreads];
  const uint32_t n = P.counters[2];
  const uint32_t tid = blockIdx.x * kThreads + threadIdx.x;
  const uint32_t stride = gridDim.x * kThreads;
  HybridStack<kRing> st{ring + threadIdx.x, kThreads, gstack + tid, stride, 0, 0};
  for (uint32_t i = tid; i < n; i += stride) {
    uint32_t p = P.q_shadow[i];
    uint32_t cnt = P.scount[p];
    float4 L = P.lum[p];
    for (uint32_t k = 0; k < cnt; ++k) {
      size_t slot = (size_t)p * P.slots + k;
      float4 a = P.sray[2 * slot], b = P.sray[2 * slot + 1];
      bool occ = false;
      if (b.w > a.w) {  // empty range (distance <= 0): nothing can occlude (scene.rs:123-125)
        TRay r = make_tray(xyz(a), xyz(b), a.w, b.w);
        THit h;
        traverse4<true, kRing>(S.nodes4, S.tris, r, st, &h);
        occ = h.gid != PYRE_MISS;
      }
      if (!occ) {
        float4 c = P.scontrib[slot];
        L.x = L.x + c.x;
        L.y = L.y + c.y;
        L.z = L.z + c.z;
      }
    }
    P.lum[p] = L;
  }
}

// Surface normal of scene triangle gid at barycentrics (u, v) (mesh.rs:130-136).
__device__ __forceinline__ V3 surface_normal(const SceneDev& S, uint32_t gid, uint32_t mesh, float u, float v) {
  if (S.mesh_info[mesh] & 0x80000000u) {
    const float* nn = S.normals + 9 * (size_t)gid;
    float w = 1.0f - u - v;
    V3 na = v3(nn[0], nn[1], nn[2]), nb = v3(nn[3], nn[4], nn[5]), nc = v3(nn[6], nn[7], nn[8]);
    return normalize(w * na + u * nb + v * nc);
  }
  const float* vv = S.verts + 9 * (size_t)gid;
  V3 v0 = v3(vv[0], vv[1], vv[2]), v1 = v3(vv[3], vv[4], vv[5]), v2 = v3(vv[6], vv[7], vv[8]);
  return normalize(cross(v1 - v0, v2 - v0));
}

__device__ __forceinline__ uint32_t mesh_of(const SceneDev& S, uint32_t gid) {
  uint32_t lo = 0, hi = S.n_meshes;  // largest m with tri_offset[m] <= gid
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) / 2;
    if (S.tri_offset[mid] <= gid)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// One bounce of integrator.rs:76-222 for path p. Returns whether the path
// continues (next ray written) and whether it queued shadow rays.
__device__ void shade_path(const Paths& P, const SceneDev& S, const PassParams& pp, uint32_t depth, uint32_t p,
                           bool* cont, bool* shadow, uint32_t* n_shadow) {
  *cont = false;
  *shadow = false;
  *n_shadow = 0;
  float4 ra = P.ray[2 * (size_t)p], rb = P.ray[2 * (size_t)p + 1];
  V3 ro = xyz(ra), rd = xyz(rb);
  float r_tmin = ra.w;
  float4 hv = P.hit[p];
  uint32_t gid = __builtin_bit_cast(uint32_t, hv.x);
  float4 bp = P.beta_pdf[p];
  V3 beta = xyz(bp);
  float last_pdf = bp.w;
  float4 lf = P.lum[p];
  V3 l = xyz(lf);
  uint32_t flags = __builtin_bit_cast(uint32_t, lf.w);
  bool specular = (flags & FL_SPECULAR) != 0;

  // Scene::intersect light loop (scene.rs:87-95): strict t < t_max.
  int kind = gid != PYRE_MISS ? 1 : 0;
  float t_max = kind ? hv.y : rb.w;
  uint32_t light_id = 0;
  V3 lpos = splat(0.0f);
  for (uint32_t li = 0; li < S.n_lights; ++li) {
    float t;
    V3 pos;
    if (quad_intersect_plane(S.lights[li], ro, rd, r_tmin, t_max, &t, &pos) && t < t_max) {
      t_max = t;
      kind = 2;
      light_id = li;
      lpos = pos;
    }
  }

  if (kind == 0) {  // integrator.rs:77-94
    if (S.has_env) {
      V3 le = env_le(S.env, rd);
      if (!is_zero(le)) {
        if (specular) {
          l = l + beta * le;
        } else {
          float pdf_env = env_pdf(S.env, rd);
          float w = power_heuristic(last_pdf, pdf_env);
          l = l + beta * le * w;
        }
      }
    }
    P.lum[p] = f4(l, lf.w);
    return;
  }

  uint32_t mesh = 0;
  V3 pos, ns;
  if (kind == 1) {
    mesh = mesh_of(S, gid);
    pos = ro + rd * hv.y;  // Ray::at (math.rs:26-28)
    ns = surface_normal(S, gid, mesh, hv.z, hv.w);
  } else {
    pos = lpos;
    ns = S.lights[light_id].normal;
  }
  if (!(flags & FL_AOV)) {  // integrator.rs:96-110
    flags |= FL_AOV;
    V3 alb = kind == 1 ? S.mats[S.mesh_info[mesh] & 0x7FFFFFFFu].base : splat(1.0f);
    P.aov0[p] = f4(alb, t_max);
    P.aov1[p] = f4(ns, 0.0f);
  }
  if (kind == 2) {  // integrator.rs:113-129
    const QuadLightDev& L = S.lights[light_id];
    V3 le = dot(ns, -rd) > 0.0f ? L.emission : splat(0.0f);  // light.rs:146-152
    if (specular) {
      l = l + beta * le;
    } else {
      float pdf_light = quad_pdf(L, ro, rd);
      float w = power_heuristic(last_pdf, pdf_light);
      l = l + beta * le * w;
    }
    P.lum[p] = f4(l, __builtin_bit_cast(float, flags));
    return;
  }

  const MatDev mat = S.mats[S.mesh_info[mesh] & 0x7FFFFFFFu];
  Frame frame = frame_from_normal(ns);
  V3 wo = to_local(frame, -rd);
  if (wo.z <= 0.0f) {  // back face: integrator.rs:136-139
    P.lum[p] = f4(l, __builtin_bit_cast(float, flags));
    return;
  }
  Rng rng = load_rng(P, p);
  uint32_t ns_count = 0;
  const size_t slot0 = (size_t)p * P.slots;
  // NEE toward each quad light (integrator.rs:142-167)
  for (uint32_t li = 0; li < S.n_lights; ++li) {
    V2 u = rng_vec2(rng);
    const QuadLightDev& L = S.lights[li];
    V3 lp, lwi;
    float ldist, lpdf;
    if (!quad_sample(L, pos, u, &lp, &lwi, &ldist, &lpdf)) continue;
    if (lpdf <= 0.0f || is_zero(L.emission)) continue;
    V3 wi = to_local(frame, lwi);
    if (wi.z <= 0.0f) continue;
    V3 f = bsdf_eval(mat, wo, wi);
    if (is_zero(f)) continue;
    V3 p_off = pos + kRayEps * ns;
    V3 target = lp - kRayEps * lwi;
    // Scene::occluded ray (scene.rs:120-135)
    V3 to_from = target - p_off;
    float dist = length(to_from);
    V3 dir = to_from / dist;
    float pdf_bsdf = bsdf_pdf(mat, wo, wi);
    float w = power_heuristic(lpdf, pdf_bsdf);
    V3 c = beta * f * L.emission * wi.z * w / lpdf;
    size_t sl = slot0 + ns_count++;
    bool empty = !(dist > 0.0f);
    P.sray[2 * sl] = f4(p_off, 1e-3f);
    P.sray[2 * sl + 1] = f4(empty ? v3(0.0f, 0.0f, 1.0f) : dir, empty ? -1.0f : dist - 1e-3f);
    P.scontrib[sl] = f4(c, 0.0f);
  }
  // environment NEE arm (integrator.rs:172-188)
  if (S.has_env) {
    V2 u = rng_vec2(rng);
    V3 ewi, eli;
    float epdf;
    env_sample(S.env, u, &ewi, &eli, &epdf);
    if (epdf > 0.0f && !is_zero(eli)) {
      V3 wi = to_local(frame, ewi);
      if (wi.z > 0.0f) {
        V3 f = bsdf_eval(mat, wo, wi);
        if (!is_zero(f)) {
          V3 p_off = pos + kRayEps * ns;
          float pdf_bsdf = bsdf_pdf(mat, wo, wi);
          float w = power_heuristic(epdf, pdf_bsdf);
          V3 c = beta * f * eli * wi.z * w / epdf;
          size_t sl = slot0 + ns_count++;
          P.sray[2 * sl] = f4(p_off, 1e-3f);  // Scene::occluded_dir (scene.rs:102-116)
          P.sray[2 * sl + 1] = f4(ewi, kInf);
          P.scontrib[sl] = f4(c, 0.0f);
        }
      }
    }
  }
  P.scount[p] = ns_count;
  *shadow = ns_count > 0;
  *n_shadow = ns_count;
  // BSDF sample (integrator.rs:190-201)
  bool alive = true;
  V3 bwi, bf;
  float bpdf;
  if (!bsdf_sample(mat, wo, rng_vec2(rng), &bwi, &bf, &bpdf)) alive = false;
  if (alive && bpdf <= 0.0f) alive = false;
  float cos_wi = alive ? fabsf(bwi.z) : 0.0f;
  if (alive && cos_wi == 0.0f) alive = false;
  if (alive) {
    beta = beta * (bf * cos_wi / bpdf);
    if (depth >= pp.min_rr) {  // integrator.rs:203-209
      float q = rclamp(1.0f - max_element(beta), 0.05f, 0.95f);
      if (rng_f32(rng) < q)
        alive = false;
      else
        beta = beta / (1.0f - q);
    }
  }
  if (alive) {
    flags &= ~FL_SPECULAR;
    V3 wi_world = to_world(frame, bwi);
    P.ray[2 * (size_t)p] = f4(pos + kRayEps * ns, 1e-4f);
    P.ray[2 * (size_t)p + 1] = f4(wi_world, kInf);
    P.beta_pdf[p] = f4(beta, bpdf);
    *cont = depth + 1 < pp.max_depth;
  }
  store_rng(P, p, rng);
  P.lum[p] = f4(l, __builtin_bit_cast(float, flags));
}

__global__ __launch_bounds__(kThreads) void k_shade(Paths P, SceneDev S, PassParams pp, uint32_t depth) {
  const uint32_t n = P.counters[0];
  const uint32_t stride = gridDim.x * kThreads;
  // uniform trip count per wave so the ballots below see converged lanes
  const uint32_t base = blockIdx.x * kThreads + (threadIdx.x & ~63u);
  for (uint32_t i0 = base; i0 < n; i0 += stride) {
    uint32_t i = i0 + (threadIdx.x & 63u);
    bool cont = false, shadow = false;
    uint32_t p = 0, nsh = 0;
    if (i < n) {
      p = P.q_cur[i];
      shade_path(P, S, pp, depth, p, &cont, &shadow, &nsh);
    }
    for (int o = 32; o > 0; o >>= 1) nsh += __shfl_xor(nsh, o);
    if ((threadIdx.x & 63u) == 0 && nsh) atomicAdd(&P.stats[1], (unsigned long long)nsh);
    uint32_t qn = wave_append(&P.counters[1], cont);
    if (cont) P.q_next[qn] = p;
    uint32_t qs = wave_append(&P.counters[2], shadow);
    if (shadow) P.q_shadow[qs] = p;
  }
}

// main.rs:275-285 in sample order; optional FP64 radiance sums.
__global__ __launch_bounds__(kThreads) void k_accumulate(Paths P, PassParams pp, float* acc, double* acc64) {
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t px = blockIdx.x * kThreads + threadIdx.x; px < pp.npix; px += stride) {
    float a[10];
    for (int k = 0; k < 10; ++k) a[k] = acc[10 * (size_t)px + k];
    double d0 = 0, d1 = 0, d2 = 0;
    if (acc64) {
      d0 = acc64[3 * (size_t)px];
      d1 = acc64[3 * (size_t)px + 1];
      d2 = acc64[3 * (size_t)px + 2];
    }
    for (uint32_t sl = 0; sl < pp.n_samples; ++sl) {
      uint32_t p = sl * pp.npix + px;
      float4 L = P.lum[p], A = P.aov0[p], N = P.aov1[p];
      a[0] = a[0] + L.x;
      a[1] = a[1] + L.y;
      a[2] = a[2] + L.z;
      a[3] = a[3] + A.x;
      a[4] = a[4] + A.y;
      a[5] = a[5] + A.z;
      a[6] = a[6] + N.x;
      a[7] = a[7] + N.y;
      a[8] = a[8] + N.z;
      a[9] = a[9] + A.w;
      d0 += (double)L.x;
      d1 += (double)L.y;
      d2 += (double)L.z;
    }
    for (int k = 0; k < 10; ++k) acc[10 * (size_t)px + k] = a[k];
    if (acc64) {
      acc64[3 * (size_t)px] = d0;
      acc64[3 * (size_t)px + 1] = d1;
      acc64[3 * (size_t)px + 2] = d2;
    }
  }
}

__global__ void k_scale(const float* acc, size_t n, float inv_spp, float* out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = acc[i] * inv_spp;
}

__global__ void k_zero_f32(float* a, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0f;
}

template <class T>
struct Buf {
  T* p = nullptr;
  explicit Buf(size_t n) { p = dev_alloc<T>(n); }
  ~Buf() { dev_free(p); }
};

void trace_stage(pyre_scene* s, bool shadow, hipStream_t st, const Paths& P, const SceneDev& S) {
  const unsigned grid = trace_grid(s);
  if (!shadow)
    hipLaunchKernelGGL(k_extend, dim3(grid), dim3(kThreads), 0, st, P, S, trace_gstack(s));
  else
    hipLaunchKernelGGL(k_shadow, dim3(grid), dim3(kThreads), 0, st, P, S, trace_gstack(s));
}

}  // namespace

extern "C" int pyre_render_stats(pyre_ctx* ctx, pyre_scene* s, uint64_t* extension_rays, uint64_t* shadow_rays) {
  if (!ctx || !s) return PYRE_EINVAL;
  if (extension_rays) *extension_rays = s->last_extension_rays;
  if (shadow_rays) *shadow_rays = s->last_shadow_rays;
  return PYRE_OK;
}

extern "C" int pyre_render(pyre_ctx* ctx, pyre_scene* s, const pyre_camera_desc* camera,
                           const pyre_integrator_params* integrator, uint32_t width, uint32_t height, uint32_t spp,
                           uint32_t spp_offset, pyre_pixel_sample* out, double* accum64) {
  if (!ctx || !s || !camera || !integrator) return PYRE_EINVAL;
  try {
    PYRE_HIP_CHECK(hipSetDevice(ctx->device));
    require(width > 0 && height > 0, "width and height must be positive");
    require(s->built, "pyre_render before pyre_bvh_build");
    require((uint64_t)width * height < (1ull << 31), "image too large");
    hipStream_t st = ctx->stream;
    const uint32_t npix = width * height;
    const CamDev cam = make_camera(*camera);
    SceneDev S{s->d_nodes4, s->d_tris, s->d_verts, s->d_normals, s->d_tri_offset, s->d_mesh_info, s->d_mats,
               s->d_lights, s->n_meshes, s->n_lights, s->has_env ? 1u : 0u, s->env};
    const uint32_t slots = std::max(1u, s->n_lights + (s->has_env ? 1u : 0u));
    // paths per pass: whole samples of the image, about 2M paths
    const uint32_t target = 1u << 21;
    uint32_t per_pass = std::max(1u, std::min(spp ? spp : 1u, (target + npix - 1) / npix));
    const size_t cap = (size_t)per_pass * npix;
    Buf<float4> ray(2 * cap), bp(cap), lum(cap), aov0(cap), aov1(cap), hit(cap), sray(2 * cap * slots),
        scon(cap * slots);
    Buf<ulonglong2> rng(2 * cap);
    Buf<uint32_t> scount(cap), q0(cap), q1(cap), qs(cap), ctr(4);
    Buf<unsigned long long> stats(2);
    PYRE_HIP_CHECK(hipMemsetAsync(stats.p, 0, 16, st));
    Buf<float> acc(10 * (size_t)npix), outd(10 * (size_t)npix);
    Buf<double> acc64(accum64 ? 3 * (size_t)npix : 1);
    hipLaunchKernelGGL(k_zero_f32, dim3((unsigned)((10 * (size_t)npix + 255) / 256)), dim3(256), 0, st, acc.p,
                       10 * (size_t)npix);
    if (accum64)
      PYRE_HIP_CHECK(hipMemcpyAsync(acc64.p, accum64, 3 * sizeof(double) * npix, hipMemcpyHostToDevice, st));
    PYRE_HIP_CHECK(hipMemsetAsync(s->d_flag, 0, 4, st));
    Paths P{ray.p, rng.p, bp.p, lum.p, aov0.p, aov1.p, hit.p, sray.p, scon.p, scount.p, q0.p, q1.p, qs.p, ctr.p,
            stats.p, slots};
    const unsigned grid = (unsigned)std::max(1, ctx->cus) * 8;
    for (uint32_t s0 = 0; s0 < spp; s0 += per_pass) {
      uint32_t ns = std::min(per_pass, spp - s0);
      PassParams pp{width, height, npix, spp_offset + s0, ns, ns * npix, integrator->max_depth,
                    integrator->min_rr_depth};
      P.q_cur = q0.p;
      P.q_next = q1.p;
      hipLaunchKernelGGL(k_counters, dim3(1), dim3(1), 0, st, ctr.p, pp.paths, 0u, 0u);
      hipLaunchKernelGGL(k_generate, dim3(grid), dim3(kThreads), 0, st, P, pp, cam);
      for (uint32_t d = 0; d < integrator->max_depth; ++d) {
        trace_stage(s, false, st, P, S);
        hipLaunchKernelGGL(k_shade, dim3(grid), dim3(kThreads), 0, st, P, S, pp, d);
        if (s->n_lights + (s->has_env ? 1u : 0u) > 0) trace_stage(s, true, st, P, S);
        hipLaunchKernelGGL(k_rotate, dim3(1), dim3(1), 0, st, ctr.p, stats.p);
        std::swap(P.q_cur, P.q_next);
      }
      hipLaunchKernelGGL(k_accumulate, dim3(grid), dim3(kThreads), 0, st, P, pp, acc.p,
                         accum64 ? acc64.p : (double*)nullptr);
      PYRE_HIP_CHECK(hipGetLastError());
    }
    if (out) {
      const float inv_spp = 1.0f / (float)spp;  // main.rs:255
      hipLaunchKernelGGL(k_scale, dim3((unsigned)((10 * (size_t)npix + 255) / 256)), dim3(256), 0, st, acc.p,
                         10 * (size_t)npix, inv_spp, outd.p);
      PYRE_HIP_CHECK(hipMemcpyAsync(out, outd.p, 10 * sizeof(float) * npix, hipMemcpyDeviceToHost, st));
    }
    if (accum64)
      PYRE_HIP_CHECK(hipMemcpyAsync(accum64, acc64.p, 3 * sizeof(double) * npix, hipMemcpyDeviceToHost, st));
    uint32_t ovf = 0;
    unsigned long long hs[2] = {0, 0};
    PYRE_HIP_CHECK(hipMemcpyAsync(&ovf, s->d_flag, 4, hipMemcpyDeviceToHost, st));
    PYRE_HIP_CHECK(hipMemcpyAsync(hs, stats.p, 16, hipMemcpyDeviceToHost, st));
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
    if (ovf) throw Error(PYRE_EOVERFLOW, "traversal stack overflow");
    s->last_extension_rays = hs[0];
    s->last_shadow_rays = hs[1];
    ctx->err.clear();
    return PYRE_OK;
  } catch (const Error& e) {
    ctx->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return PYRE_EINVAL;
  }
}
