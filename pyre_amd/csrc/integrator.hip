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
//               (main.rs:275-285) and FP64 radiance sums
//
// Queues are compacted with wave64 ballots: one atomic per wave. Every
// stage is a persistent grid-stride kernel whose item count and pass
// parameters live in device memory, so a pass needs no host synchronisation
// and is captured once into a hipGraph and replayed. RNG draws happen in
// exactly the reference's order per path (jitter x/y, lens x/y, time; per
// surface bounce: 2 per light (quads, then emissive meshes), 2 for the environment, 2 for the BSDF, 1 for
// Russian roulette), so a GPU path consumes the same stream as the CPU one.
//
// Radiance cache (no counterpart in the reference): a CACHE pass ends a
// render path at the first surface vertex reached by a bounce off a diffuse
// material and adds beta x cache(query) there. The same pass may carry
// training paths after its render paths: full unbiased paths that record, at
// every surface vertex, the cache query and the radiance/throughput before
// the vertex, so that after the path ends the vertex's outgoing radiance
// estimate is (L_end - L_before) / beta.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <new>
#include <string>
#include <vector>

#include "nrc.hpp"
#include "shade.hpp"
#include "traverse.hpp"

using namespace pyre;

namespace {

constexpr int kThreads = 256;
constexpr uint32_t FL_SPECULAR = 1u, FL_AOV = 2u, FL_DIFFUSE = 4u, FL_TRAIN = 8u, FL_SPLIT = 16u;
// MODE_SPLIT: a path reaching the vertex where a CACHE pass would query the
// cache continues there on a stream of its own (the cache's best case)
constexpr uint32_t MODE_PLAIN = 0u, MODE_CACHE = 2u, MODE_SPLIT = 3u;
constexpr float kRayEps = 1e-3f;  // integrator.rs:34
constexpr uint32_t kVrec = 5;     // float4 per training vertex record
constexpr uint32_t kTrainStreamBase = 0x80000000u;  // sample indices of training paths
constexpr uint32_t kCounters = 32;

struct SceneDev {
  const Node4* nodes4;
  const float4* tris;
  const float* verts;
  const float* normals;
  const uint32_t* tri_offset;
  const uint32_t* mesh_blk;  // mesh lookup table (kMeshBlkShift)
  const uint32_t* mesh_info;
  const MatDev* mats;
  const QuadLightDev* lights;
  uint32_t n_meshes, n_lights, has_env;
  EnvDev env;
  const float4* motion;  // 2 per mesh (offset_t0, offset_t1) when some mesh moves, else null
  PhiTable phi;          // the host libm's sinf/cosf on the BSDF samplers' grid
  // emissive meshes (pyre_mesh_light): light ids n_lights .. n_lights + n_mlights - 1
  const MeshLightDev* mlights = nullptr;
  const float4* etris = nullptr;
  const float* ecdf = nullptr;
  uint32_t n_mlights = 0;
};

// One render job: scene, camera, image and its sums (film.rs:38-109).
struct JobDev {
  SceneDev S;
  CamDev cam;
  uint32_t width, height, npix, max_depth, min_rr;
  float* acc;      // 10 f32 sums per pixel (PixelSample order)
  double* acc64;   // 3 FP64 radiance sums per pixel
  float* welford;  // 3 f32 running-mean radiance per pixel (the viewer's accumulator) or null
};

// Samples [s_base, s_base + n_samples) of every pixel of one job, as the
// pool's paths [path0, path0 + n_samples * npix) (sample-major) and the
// accumulation items [pix0, pix0 + npix); n_before = samples already in the
// job's running mean.
struct SliceDev {
  uint32_t job, s_base, n_samples, path0, pix0, n_before;
};

// One pass over the path pool: slices [slice0, slice0 + n_slices) hold the
// render paths [0, train0) and accumulation items [0, pixels). Paths
// [train0, paths) are radiance-cache training paths of job slices[slice0].job:
// training path k is sample t_sbase + k / npix of pixel (k * pix_mul +
// pix_add) % npix.
struct PassDev {
  uint32_t slice0, n_slices, paths, pixels;
  uint32_t mode, pix_mul, pix_add, vmax;
  float diffuse_rough;
  uint32_t train0, t_sbase;
  uint32_t split;      // MODE_SPLIT: the continuation's number
  uint32_t cache_off;  // MODE_CACHE: 1 = image paths run to the end (plain), training paths still record
};

// What every stage reads from device memory: the pass being run is
// passes[*pass_idx], so captured passes replay with new contents.
struct Sched {
  const JobDev* jobs;
  const SliceDev* slices;
  const PassDev* passes;
  const uint32_t* pass_idx;
};

__device__ __forceinline__ const PassDev& cur_pass(const Sched& sc) { return sc.passes[*sc.pass_idx]; }

// Slice of pool index v: the last slice k of the pass with key(k) <= v.
template <class Key>
__device__ __forceinline__ uint32_t find_slice(const Sched& sc, const PassDev& pp, uint32_t v, Key key) {
  uint32_t lo = pp.slice0, hi = pp.slice0 + pp.n_slices;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) / 2;
    if (key(sc.slices[mid]) <= v)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

struct Paths {
  float4* ray;       // 2 per path: (o, t_min), (d, t_max)
  ulonglong2* rng;   // 2 per path
  float4* beta_pdf;  // beta, last_pdf_bsdf
  float4* lum;       // L, flags
  float4* aov0;      // albedo, depth
  float4* aov1;      // normal
  float4* hit;       // gid bits, t, u, v
  float* time;       // shutter time (written for jobs whose scene has moving meshes)
  float4* sray;      // 2 per slot
  float4* scontrib;  // contribution per slot
  uint32_t* scount;
  uint32_t* q_cur;
  uint32_t* q_next;
  uint32_t* q_shadow;
  uint32_t* q_hit;     // paths of the current queue whose ray hit a surface (k_shade_split -> k_shade)
  uint32_t* counters;  // [0] current, [1] next, [2] shadow, [3] cache queries, [4] training samples,
                       // [5] surface-hit queue, [6] BVH hits of the bounce, [8] extend work, [16] shadow work
  unsigned long long* stats;  // [0] extension rays, [1] shadow rays, [2] cache queries, [3] training samples
  uint32_t* pjob;    // job of each path
  float4* vrec;      // training paths: kVrec float4 per (training path, vertex)
  uint32_t* vcount;  // training paths: vertices recorded
  float* qfeat;      // CACHE: 13 floats per path (its query, if it ended in the cache)
  uint32_t* qpath;   // CACHE: path of each query
  float* qrgb;       // CACHE: cache output per query
  uint32_t slots;    // shadow slots per path (lights + env)
};

__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }

// Block-wide append to Q queues: one global atomic per queue per call for
// the whole block (contended atomics on a handful of counters are the cost
// that matters at 2^24 paths per pass, not the LDS traffic). Every thread of
// the block must call it; slot[q] is meaningful where flag[q] is set.
constexpr int kWaves = kThreads / 64;
template <int Q>
struct BlockQueues {
  uint32_t cnt[Q][kWaves];
  uint32_t base[Q][kWaves];
};

template <int Q>
__device__ __forceinline__ void block_append(BlockQueues<Q>& sh, uint32_t* const (&ctr)[Q], const bool (&flag)[Q],
                                             uint32_t (&slot)[Q]) {
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  uint64_t m[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) m[q] = __ballot(flag[q]);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < Q; ++q) sh.cnt[q][wave] = (uint32_t)__popcll(m[q]);
  }
  __syncthreads();
  if (threadIdx.x < (uint32_t)Q && ctr[threadIdx.x]) {
    const int q = threadIdx.x;
    uint32_t tot = 0;
    for (int w = 0; w < kWaves; ++w) tot += sh.cnt[q][w];
    uint32_t b = tot ? atomicAdd(ctr[q], tot) : 0u;
    for (int w = 0; w < kWaves; ++w) {
      sh.base[q][w] = b;
      b += sh.cnt[q][w];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < Q; ++q) slot[q] = sh.base[q][wave] + (uint32_t)__popcll(m[q] & lanemask_lt());
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

__global__ void k_set_pass(PassDev* pp, SliceDev* sl, PassDev v, SliceDev w) {
  *pp = v;
  *sl = w;
}

__global__ void k_set_u32(uint32_t* a, uint32_t v) { *a = v; }
__global__ void k_advance(uint32_t* pass_idx) { *pass_idx += 1u; }

// counters[8..15] / [16..23]: work counters of the extend / shadow slices
__global__ void k_counters(uint32_t* c, Sched sc) {
  const uint32_t i = threadIdx.x;
  if (i < kCounters) c[i] = i == 0 ? cur_pass(sc).paths : 0u;
}

// next queue becomes current; next, shadow and surface-hit queues, the hit
// count and the work counters are emptied
__global__ void k_rotate(uint32_t* c, unsigned long long* stats) {
  const uint32_t i = threadIdx.x;
  const uint32_t cur = c[0], next = c[1];
  __syncthreads();
  if (i == 0) {
    stats[0] += cur;
    c[0] = next;
  } else if (i == 1 || i == 2 || i == 5 || i == 6 || (i >= 8 && i < kCounters)) {
    c[i] = 0u;
  }
}

// Which sample path p of the current pass is: its job, its index within the
// job's slice (or among the training paths), its pixel and its sample
// number (main.rs:261: one stream per (x, y, sample)).
struct PathSample {
  uint32_t job, local, x, y, sample;
  bool train;
};

__device__ __forceinline__ PathSample path_sample(const Sched& sc, const PassDev& pp, uint32_t p) {
  PathSample r;
  r.train = p >= pp.train0;
  uint32_t s_base;
  if (r.train) {
    r.job = sc.slices[pp.slice0].job;
    r.local = p - pp.train0;
    s_base = pp.t_sbase;
  } else {
    const SliceDev slc = sc.slices[find_slice(sc, pp, p, [](const SliceDev& q) { return q.path0; })];
    r.job = slc.job;
    r.local = p - slc.path0;
    s_base = slc.s_base;
  }
  const JobDev& J = sc.jobs[r.job];
  const uint32_t npix = J.npix;
  const uint32_t pixel =
      r.train ? (uint32_t)(((uint64_t)r.local * pp.pix_mul + pp.pix_add) % npix) : r.local % npix;
  r.sample = s_base + r.local / npix;
  r.x = pixel % J.width;
  r.y = pixel / J.width;
  return r;
}

// main.rs:261-273 + camera.rs + integrator.rs:63-74 state. The stream's
// state after the camera's draws is not stored: the first bounce reseeds it
// (camera_stream) instead of reading 32 bytes per path back.
__global__ __launch_bounds__(kThreads) void k_generate(Paths P, Sched sc) {
  const PassDev pp = cur_pass(sc);
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t p = blockIdx.x * kThreads + threadIdx.x; p < pp.paths; p += stride) {
    const PathSample ps = path_sample(sc, pp, p);
    const JobDev& J = sc.jobs[ps.job];
    Rng rng = rng_seed(pixel_seed(ps.x, ps.y, ps.sample));
    float jx = rng_f32(rng);
    float jy = rng_f32(rng);
    float ndc_x = 2.0f * ((float)ps.x + jx) / (float)J.width - 1.0f;
    float ndc_y = 1.0f - 2.0f * ((float)ps.y + jy) / (float)J.height;
    V2 lens = rng_vec2(rng);
    const float shutter = rng_f32(rng);  // CameraSample::time (main.rs:268)
    if (J.S.motion) P.time[p] = shutter;
    V3 o, d;
    camera_ray(J.cam, V2{ndc_x, ndc_y}, lens, &o, &d);
    P.ray[2 * (size_t)p] = f4(o, 1e-4f);
    P.ray[2 * (size_t)p + 1] = f4(d, kInf);
    P.lum[p] = make_float4(0.0f, 0.0f, 0.0f, __builtin_bit_cast(float, FL_SPECULAR | (ps.train ? FL_TRAIN : 0u)));
    P.pjob[p] = ps.job;
    if (ps.train) P.vcount[ps.local] = 0;
  }
}

// MODE_SPLIT: the stream a path continues on from its split vertex,
// continuation number pp.split of the path's sample.
__device__ __forceinline__ Rng split_stream(const Sched& sc, const PassDev& pp, uint32_t p) {
  const PathSample ps = path_sample(sc, pp, p);
  return rng_seed(pixel_seed(ps.x, ps.y, ps.sample) ^ (0xD1B54A32D192ED03ull * ((uint64_t)pp.split + 1ull)));
}

// The stream of path p as the first bounce finds it: seeded by pixel_seed
// and past the camera's five draws (jitter x / y, lens x / y, time).
__device__ __forceinline__ Rng camera_stream(const Sched& sc, const PassDev& pp, uint32_t p) {
  const PathSample ps = path_sample(sc, pp, p);
  Rng rng = rng_seed(pixel_seed(ps.x, ps.y, ps.sample));
  for (int k = 0; k < 5; ++k) (void)rng_next_u64(rng);
  return rng;
}

// Paths a wave claims at a time from a stage's queue counter once its
// static share is used up (the stage queues shrink bounce by bounce, so one
// chunk size; more, smaller claims contend on the counter).
constexpr uint32_t kPathChunk = 256;

// Closest hit for every queued path, lanes refilled from the queue as their
// rays finish. MOTION: some job's scene has moving meshes (the path's shutter
// time goes with its ray). first: the first bounce, whose queue is every
// path of the pass in order (not stored).
template <int RING, bool MOTION = false>
__global__ __launch_bounds__(kThreads) void k_extend(Paths P, const JobDev* __restrict__ jobs, int2* gstack,
                                                   bool one_job, bool first) {
  __shared__ int2 ring[RING * kThreads];
  const uint32_t tid = blockIdx.x * kThreads + threadIdx.x;
  HybridStack<RING> st{ring + threadIdx.x, kThreads, gstack + tid, gridDim.x * kThreads, 0, 0};
  WaveFeed<kPathChunk, kPathChunk> feed(P.counters[0], P.counters + 8);
  Trav4<false, RING, MOTION> tv;
  uint32_t p = 0, item = 0;
  const Node4* nodes = one_job ? jobs[0].S.nodes4 : nullptr;
  const float4* tris = one_job ? jobs[0].S.tris : nullptr;
  const float4* motion = MOTION && one_job ? jobs[0].S.motion : nullptr;
  bool active = false;
  uint32_t nhit = 0;  // rays of this lane that hit the BVH (the shade stage's choice of form)
  for (;;) {
    const uint64_t idle = __ballot(!active);
    if (__popcll(idle) >= kRefillLanes) {
      if (feed.take(idle, !active, item)) {
        p = first ? item : P.q_cur[item];  // the first bounce's queue is every path in order
        if (!one_job) {
          const JobDev& J = jobs[P.pjob[p]];
          nodes = J.S.nodes4;
          tris = J.S.tris;
          if (MOTION) motion = J.S.motion;
        }
        const float4 a = P.ray[2 * (size_t)p], b = P.ray[2 * (size_t)p + 1];
        tv.start(make_tray(xyz(a), xyz(b), a.w, b.w, MOTION && motion ? P.time[p] : 0.0f), st);
        active = true;
      }
      if (!__any(active)) break;
    }
    if (active && tv.step(nodes, tris, st, motion)) {
      THit h;
      tv.result(&h);
      P.hit[p] = make_float4(__builtin_bit_cast(float, h.gid), h.t, h.u, h.v);
      nhit += h.gid != PYRE_MISS ? 1u : 0u;
      active = false;
    }
  }
  // the loop ends for the whole wave at once
  for (int k = 32; k > 0; k >>= 1) nhit += __shfl_xor(nhit, k);
  if (__lane_id() == 0 && nhit) atomicAdd(&P.counters[6], nhit);
}

// Any-hit for every shadow slot of the queued paths; a lane keeps a path
// until all its slots are traced so unoccluded contributions are added to the
// path radiance in slot order (lights, then environment).
template <int RING, bool MOTION = false>
__global__ __launch_bounds__(kThreads) void k_shadow(Paths P, const JobDev* __restrict__ jobs, int2* gstack,
                                                   bool one_job) {
  __shared__ int2 ring[RING * kThreads];
  const uint32_t tid = blockIdx.x * kThreads + threadIdx.x;
  HybridStack<RING> st{ring + threadIdx.x, kThreads, gstack + tid, gridDim.x * kThreads, 0, 0};
  WaveFeed<kPathChunk, kPathChunk> feed(P.counters[2], P.counters + 16);
  Trav4<true, RING, MOTION> tv;
  tv.cur = -1;  // -1: no traversal in progress for the current slot
  uint32_t p = 0, k = 0, cnt = 0, item = 0;
  const Node4* nodes = one_job ? jobs[0].S.nodes4 : nullptr;
  const float4* tris = one_job ? jobs[0].S.tris : nullptr;
  const float4* motion = MOTION && one_job ? jobs[0].S.motion : nullptr;
  float tm = 0.0f;  // the path's shutter time (scene.rs:102-140 take ray.time)
  bool active = false;
  for (;;) {
    const uint64_t idle = __ballot(!active);
    if (__popcll(idle) >= kRefillLanes) {
      if (feed.take(idle, !active, item)) {
        p = P.q_shadow[item];
        if (!one_job) {
          const JobDev& J = jobs[P.pjob[p]];
          nodes = J.S.nodes4;
          tris = J.S.tris;
          if (MOTION) motion = J.S.motion;
        }
        if (MOTION) tm = motion ? P.time[p] : 0.0f;
        cnt = P.scount[p];
        k = 0;
        tv.cur = -1;
        // the first slot's ray with the path's data (one memory round trip)
        const float4 a = P.sray[2 * (size_t)p * P.slots], b = P.sray[2 * (size_t)p * P.slots + 1];
        if (cnt > 0 && b.w > a.w) tv.start(make_tray(xyz(a), xyz(b), a.w, b.w, tm), st);
        active = true;
      }
      if (!__any(active)) break;
    }
    if (!active) continue;
    bool occ = false, traced = false;
    if (k < cnt) {
      // a slot whose ray has an empty range (distance <= 0) cannot be
      // occluded (scene.rs:123-125); others start a traversal on first visit
      const size_t slot = (size_t)p * P.slots + k;
      if (tv.cur < 0) {
        const float4 a = P.sray[2 * slot], b = P.sray[2 * slot + 1];
        if (b.w > a.w) tv.start(make_tray(xyz(a), xyz(b), a.w, b.w, tm), st);
      }
      if (tv.cur >= 0) {
        if (tv.step(nodes, tris, st, motion)) {
          THit h;
          tv.result(&h);
          occ = h.gid != PYRE_MISS;
          traced = true;
          tv.cur = -1;
        }
      } else {
        traced = true;
      }
      if (traced) {
        if (!occ) {
          // the lane owns path p: its radiance is updated in memory, slot by
          // slot in order, rather than held in registers for the whole path
          const float4 cc = P.scontrib[slot];
          float4 L = P.lum[p];
          L.x = L.x + cc.x;
          L.y = L.y + cc.y;
          L.z = L.z + cc.z;
          P.lum[p] = L;
        }
        ++k;
      }
    }
    if (k >= cnt) active = false;
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

// The mesh of triangle gid: the largest m with tri_offset[m] <= gid. The
// lookup table brackets it within [mesh_blk[b], mesh_blk[b + 1]] for gid's
// block b (one cache line; the two entries are equal unless a mesh starts
// inside the block), which the search then narrows.
__device__ __forceinline__ uint32_t mesh_of(const SceneDev& S, uint32_t gid) {
  const uint32_t b = gid >> kMeshBlkShift;
  uint32_t lo = S.mesh_blk[b], hi = S.mesh_blk[b + 1] + 1;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) / 2;
    if (S.tri_offset[mid] <= gid)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

struct ShadeOut {
  bool cont, shadow, query;
  uint32_t n_shadow;
};

// One bounce of integrator.rs:76-222 for path p (PART = kShadeAll), or its
// first part (kShadeSplit, k_shade_split): the ray's nearest hit among the
// BVH's and the quad lights'; an escaped ray or a light hit ends the path
// here (and the call returns false); a surface hit returns true, writing
// nothing, to be shaded whole by k_shade from the surface-hit queue (its
// light loop finds no nearer light again).
constexpr int kShadeAll = 0, kShadeSplit = 1;
template <int PART>
__device__ bool shade_path(const Paths& P, const Sched& sc, const JobDev& J, const PassDev& pp, uint32_t depth,
                           uint32_t p, ShadeOut& o) {
  const SceneDev& S = J.S;
  float4 ra = P.ray[2 * (size_t)p], rb = P.ray[2 * (size_t)p + 1];
  V3 ro = xyz(ra), rd = xyz(rb);
  float r_tmin = ra.w;
  float4 hv = P.hit[p];
  uint32_t gid = __builtin_bit_cast(uint32_t, hv.x);
  // a camera path starts with beta = 1 and no BSDF pdf (integrator.rs:63-74):
  // the first bounce needs no stored state for them
  const float4 bp = depth == 0 ? make_float4(1.0f, 1.0f, 1.0f, 0.0f) : P.beta_pdf[p];
  V3 beta = xyz(bp);
  float last_pdf = bp.w;
  float4 lf = P.lum[p];
  V3 l = xyz(lf);
  uint32_t flags = __builtin_bit_cast(uint32_t, lf.w);
  bool specular = (flags & FL_SPECULAR) != 0;

  // Scene::intersect light loop (scene.rs:87-95): strict t < t_max.
  int kind = gid != PYRE_MISS ? 1 : 0;
  float t_max = kind ? hv.y : rb.w;
  uint32_t light_id = 0, light_tri = 0;
  V3 lpos = splat(0.0f);
  {
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
    // emissive meshes after the quads, with the same strict t < t_max
    for (uint32_t mi = 0; mi < S.n_mlights; ++mi) {
      float t;
      uint32_t tri;
      if (mlight_intersect(S.mlights[mi], S.etris, ro, rd, r_tmin, t_max, &t, &tri) && t < t_max) {
        t_max = t;
        kind = 2;
        light_id = S.n_lights + mi;
        light_tri = tri;
        lpos = ro + rd * t;
      }
    }
    if (PART == kShadeSplit && kind == 1) return true;
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
    return false;
  }

  uint32_t mesh = 0;
  V3 pos, ns;
  if (PART == kShadeAll && kind == 1) {
    mesh = mesh_of(S, gid);
    pos = ro + rd * hv.y;  // Ray::at (math.rs:26-28)
    if (S.motion) {
      // a moving mesh's hit: Ray::at on the shifted ray, shifted back
      // (mesh.rs:118-146)
      const float4 m0 = S.motion[2 * mesh], m1 = S.motion[2 * mesh + 1];
      const V3 off = lerp(xyz(m0), xyz(m1), rclamp(P.time[p], 0.0f, 1.0f));
      if (!(off.x == 0.0f && off.y == 0.0f && off.z == 0.0f)) pos = ((ro - off) + rd * hv.y) + off;
    }
    ns = surface_normal(S, gid, mesh, hv.z, hv.w);
  } else {
    pos = lpos;
    ns = light_id < S.n_lights ? S.lights[light_id].normal
                               : mlight_normal(S.mlights[light_id - S.n_lights], S.etris, light_tri);
  }
  if (!(flags & FL_AOV)) {  // integrator.rs:96-110
    flags |= FL_AOV;
    V3 alb = PART == kShadeAll && kind == 1 ? S.mats[S.mesh_info[mesh] & 0x7FFFFFFFu].base : splat(1.0f);
    P.aov0[p] = f4(alb, t_max);
    P.aov1[p] = f4(ns, 0.0f);
  }
  if (kind == 2) {  // a light hit: integrator.rs:113-129
    const bool quad = light_id < S.n_lights;
    const V3 em = quad ? S.lights[light_id].emission : S.mlights[light_id - S.n_lights].emission;
    V3 le = dot(ns, -rd) > 0.0f ? em : splat(0.0f);  // light.rs:146-152
    if (specular) {
      l = l + beta * le;
    } else {
      float pdf_light = quad ? quad_pdf(S.lights[light_id], ro, rd)
                             : mlight_pdf(S.mlights[light_id - S.n_lights], S.etris, ro, rd);
      float w = power_heuristic(last_pdf, pdf_light);
      l = l + beta * le * w;
    }
    P.lum[p] = f4(l, __builtin_bit_cast(float, flags));
    return false;
  }
  if constexpr (PART == kShadeSplit) return false;  // not reached: every kind returned above

  const MatDev mat = S.mats[S.mesh_info[mesh] & 0x7FFFFFFFu];
  Frame frame = frame_from_normal(ns);
  V3 wo = to_local(frame, -rd);
  if (wo.z <= 0.0f) {  // back face: integrator.rs:136-139
    P.lum[p] = f4(l, __builtin_bit_cast(float, flags));
    return false;
  }
  bool reseed = false;
  if (pp.mode == MODE_SPLIT) {
    // the vertex where a cache render would end the path and query the cache
    if (depth >= 1 && (flags & FL_DIFFUSE) && !(flags & FL_SPLIT)) {
      reseed = true;
      flags |= FL_SPLIT;
    }
  } else if (pp.mode != MODE_PLAIN) {
    const float rough = mat.kind == PYRE_MAT_LAMBERTIAN ? 1.0f : mat.roughness;
    const float qf[13] = {pos.x, pos.y, pos.z, -rd.x, -rd.y, -rd.z, ns.x, ns.y, ns.z, rough,
                          mat.base.x, mat.base.y, mat.base.z};
    if (!(flags & FL_TRAIN) && depth >= 1 && (flags & FL_DIFFUSE) && !pp.cache_off) {
      for (int k = 0; k < 13; ++k) P.qfeat[13 * (size_t)p + k] = qf[k];
      o.query = true;
      P.lum[p] = f4(l, __builtin_bit_cast(float, flags));
      return false;
    }
    if (flags & FL_TRAIN) {
      const uint32_t tp = p - pp.train0;
      uint32_t k = P.vcount[tp];
      if (k < pp.vmax) {
        float4* v = P.vrec + ((size_t)tp * pp.vmax + k) * kVrec;
        v[0] = make_float4(qf[0], qf[1], qf[2], qf[3]);
        v[1] = make_float4(qf[4], qf[5], qf[6], qf[7]);
        v[2] = make_float4(qf[8], qf[9], qf[10], qf[11]);
        v[3] = make_float4(qf[12], l.x, l.y, l.z);
        v[4] = make_float4(beta.x, beta.y, beta.z, 0.0f);
        P.vcount[tp] = k + 1;
      }
    }
  }
  Rng rng = depth == 0 ? camera_stream(sc, pp, p) : (reseed ? split_stream(sc, pp, p) : load_rng(P, p));
  uint32_t ns_count = 0;
  const size_t slot0 = (size_t)p * P.slots;
  // NEE toward a point sampled on a light (integrator.rs:142-167)
  auto nee_point = [&](V3 lp, V3 lwi, float lpdf, V3 li) {
    if (lpdf <= 0.0f || is_zero(li)) return;
    V3 wi = to_local(frame, lwi);
    if (wi.z <= 0.0f) return;
    V3 f = bsdf_eval(mat, wo, wi);
    if (is_zero(f)) return;
    V3 p_off = pos + kRayEps * ns;
    V3 target = lp - kRayEps * lwi;
    // Scene::occluded ray (scene.rs:120-135)
    V3 to_from = target - p_off;
    float dist = length(to_from);
    V3 dir = to_from / dist;
    float pdf_bsdf = bsdf_pdf(mat, wo, wi);
    float w = power_heuristic(lpdf, pdf_bsdf);
    V3 c = beta * f * li * wi.z * w / lpdf;
    size_t sl = slot0 + ns_count++;
    bool empty = !(dist > 0.0f);
    P.sray[2 * sl] = f4(p_off, 1e-3f);
    P.sray[2 * sl + 1] = f4(empty ? v3(0.0f, 0.0f, 1.0f) : dir, empty ? -1.0f : dist - 1e-3f);
    P.scontrib[sl] = f4(c, 0.0f);
  };
  for (uint32_t li = 0; li < S.n_lights; ++li) {  // each quad light
    V2 u = rng_vec2(rng);
    const QuadLightDev& L = S.lights[li];
    V3 lp, lwi;
    float ldist, lpdf;
    if (quad_sample(L, pos, u, &lp, &lwi, &ldist, &lpdf)) nee_point(lp, lwi, lpdf, L.emission);
  }
  for (uint32_t mi = 0; mi < S.n_mlights; ++mi) {  // then each emissive mesh
    V2 u = rng_vec2(rng);
    const MeshLightDev& L = S.mlights[mi];
    V3 lp, lwi;
    float ldist, lpdf;
    if (mlight_sample(L, S.etris, S.ecdf, pos, u, &lp, &lwi, &ldist, &lpdf)) nee_point(lp, lwi, lpdf, L.emission);
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
  o.shadow = ns_count > 0;
  o.n_shadow = ns_count;
  // BSDF sample (integrator.rs:190-201)
  bool alive = true;
  V3 bwi, bf;
  float bpdf;
  if (!bsdf_sample(mat, wo, rng_vec2(rng), S.phi, &bwi, &bf, &bpdf)) alive = false;
  if (alive && bpdf <= 0.0f) alive = false;
  float cos_wi = alive ? fabsf(bwi.z) : 0.0f;
  if (alive && cos_wi == 0.0f) alive = false;
  if (alive) {
    beta = beta * (bf * cos_wi / bpdf);
    if (depth >= J.min_rr) {  // integrator.rs:203-209
      float q = rclamp(1.0f - max_element(beta), 0.05f, 0.95f);
      if (rng_f32(rng) < q)
        alive = false;
      else
        beta = beta / (1.0f - q);
    }
  }
  if (alive) {
    flags &= ~(FL_SPECULAR | FL_DIFFUSE);
    if (mat.kind == PYRE_MAT_LAMBERTIAN || mat.roughness >= pp.diffuse_rough) flags |= FL_DIFFUSE;
    V3 wi_world = to_world(frame, bwi);
    P.ray[2 * (size_t)p] = f4(pos + kRayEps * ns, 1e-4f);
    P.ray[2 * (size_t)p + 1] = f4(wi_world, kInf);
    P.beta_pdf[p] = f4(beta, bpdf);
    o.cont = depth + 1 < J.max_depth;
  }
  store_rng(P, p, rng);
  P.lum[p] = f4(l, __builtin_bit_cast(float, flags));
  return false;
}

// The shade stage's form for a bounce, from the queue's size and its BVH hits
// (counted by k_extend; both unchanged until k_rotate): when at most a fifth
// of the paths hit a surface, the bounce is shaded in two kernels --
// k_shade_split ends escaped rays and light hits and queues the surface hits,
// k_shade shades the queued hits on full waves -- so that the rare hits'
// chains of dependent loads do not stretch every wave. With more hits one
// kernel over every path is faster (it overlaps the hits' chains with the
// escaped rays' work): a bounce with a quarter of its paths hitting took 7.7
// ms split and 7.1 ms whole on the 1080p production frame, and the closed
// Cornell box, where almost every ray hits, lost 2.9 ms per frame split.
__device__ __forceinline__ bool shade_split(const Paths& P) { return 5u * P.counters[6] <= P.counters[0]; }

// The first part of a split bounce for every path of the current queue (the
// kernel exits at once when the bounce is shaded whole).
__global__ __launch_bounds__(kThreads) void k_shade_split(Paths P, Sched sc, uint32_t depth, bool one_job) {
  __shared__ BlockQueues<1> qs;
  if (!shade_split(P)) return;
  const PassDev pp = cur_pass(sc);
  const uint32_t n = P.counters[0];
  const uint32_t stride = gridDim.x * kThreads;
  uint32_t* const ctr[1] = {&P.counters[5]};
  // block-uniform trip count: the appends below synchronise the block
  for (uint32_t b0 = blockIdx.x * kThreads; b0 < n; b0 += stride) {
    const uint32_t i = b0 + threadIdx.x;
    bool hit = false;
    uint32_t p = 0;
    if (i < n) {
      p = depth == 0 ? i : P.q_cur[i];
      ShadeOut o;
      hit = shade_path<kShadeSplit>(P, sc, sc.jobs[one_job ? 0u : P.pjob[p]], pp, depth, p, o);
    }
    const bool flag[1] = {hit};
    uint32_t slot[1];
    block_append<1>(qs, ctr, flag, slot);
    if (hit) P.q_hit[slot[0]] = p;
  }
}

// A whole bounce for every path of the current queue, or, when the bounce is
// split (shade_split), for the surface hits that k_shade_split queued.
// MINW: minimum waves per SIMD asked of the register allocator (0 = none).
template <int MINW>
__global__ __launch_bounds__(kThreads, MINW) void k_shade(Paths P, Sched sc, uint32_t depth, bool one_job) {
  __shared__ BlockQueues<3> qs;
  __shared__ unsigned long long s_nsh[kWaves];
  const PassDev pp = cur_pass(sc);
  const bool split = shade_split(P);
  const uint32_t n = split ? P.counters[5] : P.counters[0];
  const uint32_t stride = gridDim.x * kThreads;
  uint32_t* const ctr[3] = {&P.counters[1], &P.counters[2], pp.mode == MODE_CACHE ? &P.counters[3] : nullptr};
  unsigned long long nsh = 0;
  // block-uniform trip count: the appends below synchronise the block
  for (uint32_t b0 = blockIdx.x * kThreads; b0 < n; b0 += stride) {
    const uint32_t i = b0 + threadIdx.x;
    ShadeOut o;
    o.cont = o.shadow = o.query = false;
    o.n_shadow = 0;
    uint32_t p = 0;
    if (i < n) {
      p = split ? P.q_hit[i] : (depth == 0 ? i : P.q_cur[i]);
      shade_path<kShadeAll>(P, sc, sc.jobs[one_job ? 0u : P.pjob[p]], pp, depth, p, o);
    }
    nsh += o.n_shadow;
    const bool flag[3] = {o.cont, o.shadow, o.query};
    uint32_t slot[3];
    block_append<3>(qs, ctr, flag, slot);
    if (o.cont) P.q_next[slot[0]] = p;
    if (o.shadow) P.q_shadow[slot[1]] = p;
    if (o.query) P.qpath[slot[2]] = p;
  }
  // shadow-ray statistics: one atomic per block
  for (int k = 32; k > 0; k >>= 1) nsh += __shfl_xor(nsh, k);
  if ((threadIdx.x & 63u) == 0) s_nsh[threadIdx.x >> 6] = nsh;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kWaves; ++w) t += s_nsh[w];
    if (t) atomicAdd(&P.stats[1], t);
  }
}

// main.rs:275-285 in sample order; FP64 radiance sums beside the f32 ones.
// One item per (slice, pixel): the slice's samples of that pixel in order.
__global__ __launch_bounds__(kThreads) void k_accumulate(Paths P, Sched sc) {
  const PassDev pp = cur_pass(sc);
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t it = blockIdx.x * kThreads + threadIdx.x; it < pp.pixels; it += stride) {
    const SliceDev slc = sc.slices[find_slice(sc, pp, it, [](const SliceDev& q) { return q.pix0; })];
    const JobDev& J = sc.jobs[slc.job];
    const uint32_t px = it - slc.pix0;
    float* acc = J.acc + 10 * (size_t)px;
    double* acc64 = J.acc64 + 3 * (size_t)px;
    float a[10];
    for (int k = 0; k < 10; ++k) a[k] = acc[k];
    double d0 = acc64[0], d1 = acc64[1], d2 = acc64[2];
    for (uint32_t sl = 0; sl < slc.n_samples; ++sl) {
      uint32_t p = slc.path0 + sl * J.npix + px;
      // a path's AOVs exist once its first vertex was shaded (FL_AOV);
      // before that they are zero (integrator.rs:63-74), and are not stored
      const float4 L = P.lum[p];
      const bool has_aov = (__builtin_bit_cast(uint32_t, L.w) & FL_AOV) != 0u;
      const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      const float4 A = has_aov ? P.aov0[p] : zero, N = has_aov ? P.aov1[p] : zero;
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
    for (int k = 0; k < 10; ++k) acc[k] = a[k];
    acc64[0] = d0;
    acc64[1] = d1;
    acc64[2] = d2;
    if (J.welford) {
      // viewer.rs:165-177: after pass n (1-based) the mean becomes
      // d + (s - d) / n, component by component
      float* w = J.welford + 3 * (size_t)px;
      float w0 = w[0], w1 = w[1], w2 = w[2];
      for (uint32_t sl = 0; sl < slc.n_samples; ++sl) {
        const float4 L = P.lum[slc.path0 + sl * J.npix + px];
        const float n = (float)(slc.n_before + sl + 1u);
        w0 = w0 + (L.x - w0) / n;
        w1 = w1 + (L.y - w1) / n;
        w2 = w2 + (L.z - w2) / n;
      }
      w[0] = w0;
      w[1] = w1;
      w[2] = w2;
    }
  }
}

// Block-wide append of n items per thread to one queue: one global atomic
// per block; returns the thread's first slot. Every thread of the block must
// call it.
__device__ __forceinline__ uint32_t block_append_n(BlockQueues<1>& sh, uint32_t* ctr, uint32_t n) {
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  uint32_t x = n;  // inclusive scan over the wave
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) sh.cnt[0][wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (int w = 0; w < kWaves; ++w) tot += sh.cnt[0][w];
    uint32_t b = tot ? atomicAdd(ctr, tot) : 0u;
    for (int w = 0; w < kWaves; ++w) {
      sh.base[0][w] = b;
      b += sh.cnt[0][w];
    }
  }
  __syncthreads();
  return sh.base[0][wave] + x - n;
}

// Training paths: per recorded vertex, target = (L_end - L_before) / beta,
// compacted into (feats, targets) up to cap samples. One thread per training
// path appends its usable vertices (throughput above 1e-6 and a finite target
// in every channel) in vertex order.
__global__ __launch_bounds__(kThreads) void k_train_targets(Paths P, Sched sc, float* feats, float* targets,
                                                            uint32_t cap) {
  __shared__ BlockQueues<1> qs;
  const PassDev pp = cur_pass(sc);
  const uint32_t ntp = pp.paths - pp.train0;
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t b0 = blockIdx.x * kThreads; b0 < ntp; b0 += stride) {  // block-uniform
    const uint32_t tp = b0 + threadIdx.x;
    uint32_t nv = 0;
    uint64_t ok = 0;
    float4 L = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (tp < ntp) {
      nv = min(P.vcount[tp], pp.vmax);
      L = P.lum[pp.train0 + tp];
      for (uint32_t k = 0; k < nv; ++k) {
        const float4* v = P.vrec + ((size_t)tp * pp.vmax + k) * kVrec;
        const float4 lb = v[3], bb = v[4];
        const float t0 = (L.x - lb.y) / bb.x, t1 = (L.y - lb.z) / bb.y, t2 = (L.z - lb.w) / bb.z;
        if (bb.x > 1e-6f && bb.y > 1e-6f && bb.z > 1e-6f && isfinite(t0) && isfinite(t1) && isfinite(t2))
          ok |= 1ull << k;
      }
    }
    uint32_t slot = block_append_n(qs, &P.counters[4], (uint32_t)__popcll(ok));
    for (uint32_t k = 0; k < nv; ++k) {
      if (!((ok >> k) & 1ull)) continue;
      if (slot < cap) {
        const float4* v = P.vrec + ((size_t)tp * pp.vmax + k) * kVrec;
        const float4 v0 = v[0], v1 = v[1], v2 = v[2], lb = v[3], bb = v[4];
        float* f = feats + 13 * (size_t)slot;
        f[0] = v0.x;
        f[1] = v0.y;
        f[2] = v0.z;
        f[3] = v0.w;
        f[4] = v1.x;
        f[5] = v1.y;
        f[6] = v1.z;
        f[7] = v1.w;
        f[8] = v2.x;
        f[9] = v2.y;
        f[10] = v2.z;
        f[11] = v2.w;
        f[12] = lb.x;
        float* t = targets + 3 * (size_t)slot;
        t[0] = fmaxf((L.x - lb.y) / bb.x, 0.0f);
        t[1] = fmaxf((L.y - lb.z) / bb.y, 0.0f);
        t[2] = fmaxf((L.z - lb.w) / bb.z, 0.0f);
      }
      ++slot;
    }
  }
}

// stats[3]: training samples produced (counters[4] of the pass just traced)
__global__ void k_count_train(Paths P) { P.stats[3] += P.counters[4]; }

// CACHE: L += beta x max(cache, 0) for every path that ended in the cache.
__global__ __launch_bounds__(kThreads) void k_cache_apply(Paths P) {
  const uint32_t n = P.counters[3];
  if (blockIdx.x == 0 && threadIdx.x == 0) P.stats[2] += n;
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    uint32_t p = P.qpath[i];
    float4 L = P.lum[p], b = P.beta_pdf[p];
    L.x = L.x + b.x * fmaxf(P.qrgb[3 * (size_t)i], 0.0f);
    L.y = L.y + b.y * fmaxf(P.qrgb[3 * (size_t)i + 1], 0.0f);
    L.z = L.z + b.z * fmaxf(P.qrgb[3 * (size_t)i + 2], 0.0f);
    P.lum[p] = L;
  }
}

__global__ void k_scale(const float* acc, size_t n, float inv_spp, float* out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = acc[i] * inv_spp;
}

// Film sinks on the device (film.rs:117-264) from a film's sums: the mean is
// sum * inv_spp as for pyre_session_read, then save_png's / save_aov_pngs' /
// save_exr's conversions.
__device__ __forceinline__ uint8_t u8_sat(float x) {  // Rust's `f32 as u8`
  if (!(x > 0.0f)) return 0;
  if (x >= 255.0f) return 255;
  return (uint8_t)x;
}

// save_png's byte `(clamp(x, 0, 1).powf(1 / 2.2) * 255) as u8` (film.rs:181-198)
// as the host's libm computes it: t[k - 1] is the smallest f32 in [0, 1] whose
// byte is at least k (the byte is monotone in x; thresholds found on the host
// by bisection over the f32 bit patterns), so the byte of x is the number of
// thresholds <= clamp(x, 0, 1) and device packing equals host packing.
struct Gamma8 {
  float t[255];
};

__device__ __forceinline__ uint8_t gamma8(float x, const Gamma8& G) {
  const float c = rmin(rmax(x, 0.0f), 1.0f);  // NaN -> 0 (f32::max ignores NaN)
  uint32_t lo = 0, hi = 255;                   // count of thresholds <= c lies in [lo, hi]
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (G.t[mid - 1] <= c)
      lo = mid;
    else
      hi = mid - 1;
  }
  return (uint8_t)lo;
}

static uint8_t host_gamma8(float x) {
  const float v = ::powf(rmin(rmax(x, 0.0f), 1.0f), 1.0f / 2.2f) * 255.0f;
  if (!(v > 0.0f)) return 0;
  if (v >= 255.0f) return 255;
  return (uint8_t)v;
}

static const Gamma8& gamma8_thresholds() {
  static const Gamma8 G = [] {
    Gamma8 g{};
    for (uint32_t k = 1; k <= 255; ++k) {
      uint32_t lo = 0, hi = 0x3F800000u;  // bit patterns of 0 and 1.0f; host_gamma8(1) = 255
      while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (host_gamma8(__builtin_bit_cast(float, mid)) >= k)
          hi = mid;
        else
          lo = mid + 1;
      }
      g.t[k - 1] = __builtin_bit_cast(float, lo);
    }
    return g;
  }();
  return G;
}

// depth range over the hits (depth > 0): positive floats order like their bits
__global__ void k_depth_range(const float* acc, uint32_t npix, float inv_spp, uint32_t* range) {
  uint32_t lo = 0x7F800000u, hi = 0u;  // +inf, 0
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += gridDim.x * blockDim.x) {
    const float d = acc[10 * (size_t)i + 9] * inv_spp;
    if (d > 0.0f) {
      const uint32_t b = __builtin_bit_cast(uint32_t, d);
      lo = min(lo, b);
      hi = max(hi, b);
    }
  }
  for (int k = 32; k > 0; k >>= 1) {
    lo = min(lo, (uint32_t)__shfl_xor((int)lo, k));
    hi = max(hi, (uint32_t)__shfl_xor((int)hi, k));
  }
  if (__lane_id() == 0) {
    atomicMin(&range[0], lo);
    atomicMax(&range[1], hi);
  }
}

__global__ void k_pack(const float* acc, uint32_t npix, float inv_spp, uint32_t what, const uint32_t* range,
                       Gamma8 G, uint8_t* o8, float* of) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  float m[10];
  for (int k = 0; k < 10; ++k) m[k] = acc[10 * (size_t)i + k] * inv_spp;
  if (what == PYRE_PACK_BEAUTY_PNG || what == PYRE_PACK_ALBEDO_PNG) {
    const int b = what == PYRE_PACK_BEAUTY_PNG ? 0 : 3;
    for (int c = 0; c < 3; ++c) o8[3 * (size_t)i + c] = gamma8(m[b + c], G);
  } else if (what == PYRE_PACK_NORMAL_PNG) {
    for (int c = 0; c < 3; ++c) o8[3 * (size_t)i + c] = u8_sat(rmin(rmax(m[6 + c] * 0.5f + 0.5f, 0.0f), 1.0f) * 255.0f);
  } else if (what == PYRE_PACK_DEPTH_PNG) {
    const float lo = range[0] == 0x7F800000u ? kInf : __builtin_bit_cast(float, range[0]);
    const float hi = __builtin_bit_cast(float, range[1]);
    const float r = rmax(hi - lo, 1e-6f);
    const float d = m[9];
    o8[i] = u8_sat((d > 0.0f ? rmin(rmax(1.0f - (d - lo) / r, 0.0f), 1.0f) : 0.0f) * 255.0f);
  } else {
    const float v[10] = {m[2], m[1], m[6], m[7], m[8], m[0], m[9], m[5], m[4], m[3]};
    for (int k = 0; k < 10; ++k) of[(size_t)k * npix + i] = v[k];
  }
}

__global__ void k_fill_f32(float* a, size_t n, float v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
__global__ void k_fill_f64(double* a, size_t n, double v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
__global__ void k_fill_u64(unsigned long long* a, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0ull;
}

inline unsigned nblocks(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// Launch configuration of the path stages: paths per pass (queue capacity),
// persistent traversal blocks per CU and the LDS part of the traversal stack
// (ring entries per thread). Results do not depend on it.
struct Tuning {
  // 2^27 paths (about 35 GB of path state with two shadow slots) hold a
  // 64-spp 1080p frame in one pass: fewer passes leave fewer nearly empty
  // deep bounces, and the frame runs about 1.8x faster than at 2^24.
  uint32_t pool_paths = 1u << 27;
  uint32_t blocks_per_cu = 8;
  uint32_t ring = 8;
  uint32_t shade_waves = 4;  // 4: shade kernel compiled for 4 waves per SIMD (1: no occupancy target)
};

// The context's pyre_tuning resolved to values.
Tuning ctx_tuning(const pyre_ctx* ctx) {
  Tuning t;
  if (ctx->tune.pool_paths) t.pool_paths = ctx->tune.pool_paths;
  t.blocks_per_cu = tune_blocks_per_cu(ctx);
  t.ring = tune_ring(ctx);
  t.shade_waves = ctx->tune.shade_waves == 1 ? 1u : 4u;
  return t;
}

// Device bytes per pool path (Work::alloc): ray 32, stream 32, throughput and
// pdf 16, radiance and flags 16, AOVs 32, hit 16, shutter time 4, the four
// queues, the job id and the shadow-slot count 24, and per shadow slot a ray
// (32) and its contribution (16); a radiance-cache pool adds a query (52),
// its path (4), the cache output (12) and a vertex count (4).
size_t path_bytes(uint32_t slots, bool cache) { return 172 + 48 * (size_t)slots + (cache ? 72 : 0); }

// Paths a pool may hold: the tuning's pool size, bounded by the tuning's
// share of the device's free memory. `reclaimable` = bytes of a pool that is
// released before the new one is allocated; `fixed` = bytes the new pool
// needs whatever its size.
uint64_t pool_limit(const pyre_ctx* ctx, const Tuning& t, size_t per_path, size_t fixed, size_t reclaimable) {
  size_t free_b = 0, total_b = 0;
  PYRE_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  const double budget = (double)tune_memory_fraction(ctx) * (double)(free_b + reclaimable) - (double)fixed;
  const double by_mem = budget > 0.0 ? budget / (double)per_path : 0.0;
  return std::min<uint64_t>(t.pool_paths, (uint64_t)by_mem);
}

// Device path state for up to `cap` paths, plus the job/slice/pass tables.
// Movable, so a context can keep one for the next one-shot render.
struct Work {
  size_t cap = 0;
  uint32_t slots = 1, vmax = 0, tcap = 0, tpaths = 0;  // tpaths: training paths with vertex records
  bool cache = false;
  size_t n_jobs = 0, n_slices = 0, n_passes = 0;
  Paths P{};
  JobDev* jobs = nullptr;
  SliceDev* slices = nullptr;
  PassDev* passes = nullptr;
  uint32_t* pass_idx = nullptr;
  int2* gstack = nullptr;  // traversal stack spill columns
  size_t gstack_cap = 0;
  std::vector<void*> bufs;
  size_t bytes = 0;  // device memory held
  // the pool's own ray / cache counters; a session points P.stats at its
  // own while it uses the pool, and unbind() points it back before the pool
  // is handed to another render (the session's buffer dies with it)
  unsigned long long* own_stats = nullptr;
  float* tfeat = nullptr;
  float* ttarget = nullptr;
  template <class T>
  T* get(size_t n) {
    T* p = dev_alloc<T>(n);
    bufs.push_back(p);
    bytes += std::max<size_t>(n, 1) * sizeof(T);
    return p;
  }
  Work() = default;
  Work(const Work&) = delete;
  Work& operator=(const Work&) = delete;
  Work& operator=(Work&& o) noexcept {
    if (this == &o) return *this;
    release();
    cap = o.cap;
    slots = o.slots;
    vmax = o.vmax;
    tcap = o.tcap;
    tpaths = o.tpaths;
    cache = o.cache;
    n_jobs = o.n_jobs;
    n_slices = o.n_slices;
    n_passes = o.n_passes;
    P = o.P;
    jobs = o.jobs;
    slices = o.slices;
    passes = o.passes;
    pass_idx = o.pass_idx;
    gstack = o.gstack;
    gstack_cap = o.gstack_cap;
    tfeat = o.tfeat;
    ttarget = o.ttarget;
    bufs = std::move(o.bufs);
    bytes = o.bytes;
    own_stats = o.own_stats;
    o.bufs.clear();
    o.bytes = 0;
    o.own_stats = nullptr;
    o.cap = 0;
    o.gstack = nullptr;
    o.gstack_cap = 0;
    return *this;
  }
  void release() {
    for (void* p : bufs) (void)hipFree(p);
    bufs.clear();
    bytes = 0;
    own_stats = nullptr;
    P.stats = nullptr;
    cap = 0;
    gstack = nullptr;
    gstack_cap = 0;
  }
  // Buffers large enough for this request (radiance-cache sizes must match).
  bool fits(size_t cap_, uint32_t slots_, uint32_t vmax_, uint32_t tpaths_, uint32_t tcap_, bool cache_,
            size_t nj, size_t ns, size_t np) const {
    return !bufs.empty() && cap >= cap_ && slots >= slots_ && vmax == vmax_ && tpaths == tpaths_ && tcap == tcap_ &&
           cache == cache_ && n_jobs >= nj && n_slices >= ns && n_passes >= np;
  }
  void alloc(size_t cap_, uint32_t slots_, uint32_t vmax_, uint32_t tpaths_, uint32_t tcap_, bool cache_,
             size_t n_jobs_, size_t n_slices_, size_t n_passes_) {
    release();
    cap = cap_;
    slots = slots_;
    vmax = vmax_;
    tpaths = tpaths_;
    tcap = tcap_;
    cache = cache_;
    n_jobs = n_jobs_;
    n_slices = n_slices_;
    n_passes = n_passes_;
    P.ray = get<float4>(2 * cap);
    P.rng = get<ulonglong2>(2 * cap);
    P.beta_pdf = get<float4>(cap);
    P.lum = get<float4>(cap);
    P.aov0 = get<float4>(cap);
    P.aov1 = get<float4>(cap);
    P.hit = get<float4>(cap);
    P.time = get<float>(cap);
    P.sray = get<float4>(2 * cap * slots);
    P.scontrib = get<float4>(cap * slots);
    P.scount = get<uint32_t>(cap);
    P.q_cur = get<uint32_t>(cap);
    P.q_next = get<uint32_t>(cap);
    P.q_shadow = get<uint32_t>(cap);
    P.q_hit = get<uint32_t>(cap);
    P.pjob = get<uint32_t>(cap);
    P.counters = get<uint32_t>(kCounters);
    P.stats = own_stats = get<unsigned long long>(4);
    P.slots = slots;
    if (vmax) {
      P.vrec = get<float4>((size_t)tpaths * vmax * kVrec);
      P.vcount = get<uint32_t>(cap);
      tfeat = get<float>(13 * (size_t)tcap);
      ttarget = get<float>(3 * (size_t)tcap);
    }
    if (cache_) {
      P.qfeat = get<float>(13 * cap);
      P.qpath = get<uint32_t>(cap);
      P.qrgb = get<float>(3 * cap);
    }
    jobs = get<JobDev>(n_jobs_);
    slices = get<SliceDev>(n_slices_);
    passes = get<PassDev>(n_passes_);
    pass_idx = get<uint32_t>(1);
  }
  // Spill columns for `threads` traversal threads on trees of binary depth
  // `depth` (3 pending children per level); allocated before any capture.
  void ensure_gstack(size_t threads, uint32_t depth) {
    const size_t need = threads * (3 * (size_t)depth + 8);
    if (need <= gstack_cap) return;
    PYRE_HIP_CHECK(hipDeviceSynchronize());
    if (gstack) {
      bufs.erase(std::find(bufs.begin(), bufs.end(), (void*)gstack));
      (void)hipFree(gstack);
      bytes -= gstack_cap * sizeof(int2);
    }
    gstack = get<int2>(need);
    gstack_cap = need;
  }
  Sched sched() const { return Sched{jobs, slices, passes, pass_idx}; }
  void unbind() { P.stats = own_stats; }
  ~Work() { release(); }
};

SceneDev scene_dev(pyre_scene* s) {
  PhiTable phi;
  phi.rank = s->ctx->d_phi_rank;
  phi.val = s->ctx->d_phi_val;
  SceneDev S{s->d_nodes4, s->d_tris, s->d_verts, s->d_normals, s->d_tri_offset, s->d_mesh_blk, s->d_mesh_info, s->d_mats,
             s->d_lights, s->n_meshes, s->n_lights, s->has_env ? 1u : 0u, s->env, s->d_motion, phi};
  S.mlights = s->d_mlights;
  S.etris = s->d_etris;
  S.ecdf = s->d_ecdf;
  S.n_mlights = s->n_mlights;
  return S;
}

JobDev job_dev(pyre_scene* s, const pyre_camera_desc& cam, const pyre_integrator_params& ip, uint32_t w, uint32_t h,
               float* acc, double* acc64, float* welford = nullptr) {
  JobDev J{};
  J.S = scene_dev(s);
  J.cam = make_camera(cam);
  J.width = w;
  J.height = h;
  J.npix = w * h;
  J.max_depth = ip.max_depth;
  J.min_rr = ip.min_rr_depth;
  J.acc = acc;
  J.acc64 = acc64;
  J.welford = welford;
  return J;
}

__global__ void k_set_job(JobDev* dst, JobDev v) { *dst = v; }

unsigned stage_grid(const pyre_ctx* ctx) { return (unsigned)std::max(1, ctx->cus) * 8; }
unsigned trace_grid_of(const pyre_ctx* ctx, const Tuning& t) {
  return (unsigned)std::max(1, ctx->cus) * t.blocks_per_cu;
}

// What a pass needs to know about its jobs: deepest path, whether any job
// has lights or an environment (shadow stage), deepest tree.
struct PassShape {
  uint32_t max_depth = 0, tree_depth = 0;
  bool nee = false;
  bool one_job = false;  // every path belongs to job 0 (sessions): its tree is read once per kernel
  bool motion = false;   // some job's scene has moving meshes
};

// Blocks of a persistent stage kernel that fit on one CU at once (its
// registers and LDS), at most `want`: a block that is not resident from the
// start would run its static share of the queue after everyone else.
template <class K>
unsigned resident_blocks(K kernel, unsigned want) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, kThreads, 0) != hipSuccess || n <= 0) return want;
  return std::min<unsigned>(want, (unsigned)n);
}

// Resident blocks per CU of every stage kernel, queried once per process
// before the first launch (stage_occ() is called outside any stream capture,
// so no runtime query is made while a pass is being captured).
struct StageOcc {
  unsigned extend[3][2], shadow[3][2];  // [ring 4 / 8 / 16][static / moving]
  unsigned generate, accumulate, shade_split, shade1, shade4, train_targets, cache_apply;
};

const StageOcc& stage_occ() {
  static const StageOcc o = [] {
    StageOcc r{};
    r.extend[0][0] = resident_blocks(k_extend<4, false>, 64);
    r.extend[1][0] = resident_blocks(k_extend<8, false>, 64);
    r.extend[2][0] = resident_blocks(k_extend<16, false>, 64);
    r.extend[1][1] = resident_blocks(k_extend<8, true>, 64);
    r.shadow[0][0] = resident_blocks(k_shadow<4, false>, 64);
    r.shadow[1][0] = resident_blocks(k_shadow<8, false>, 64);
    r.shadow[2][0] = resident_blocks(k_shadow<16, false>, 64);
    r.shadow[1][1] = resident_blocks(k_shadow<8, true>, 64);
    r.generate = resident_blocks(k_generate, 8);
    r.accumulate = resident_blocks(k_accumulate, 8);
    r.shade_split = resident_blocks(k_shade_split, 8);
    r.shade1 = resident_blocks(k_shade<1>, 8);
    r.shade4 = resident_blocks(k_shade<4>, 8);
    r.train_targets = resident_blocks(k_train_targets, 8);
    r.cache_apply = resident_blocks(k_cache_apply, 8);
    return r;
  }();
  return o;
}

constexpr int ring_index(int ring) { return ring == 4 ? 0 : ring == 16 ? 2 : 1; }

template <int RING, bool MOTION>
void launch_trace_stages(const Work& W, Paths P, unsigned cus, unsigned bpc, bool shadow, hipStream_t st,
                         bool extend, bool one_job, bool first) {
  const StageOcc& occ = stage_occ();
  if (extend)
    hipLaunchKernelGGL((k_extend<RING, MOTION>), dim3(cus * std::min(bpc, occ.extend[ring_index(RING)][MOTION])),
                       dim3(kThreads), 0, st, P, (const JobDev*)W.jobs, W.gstack, one_job, first);
  if (shadow)
    hipLaunchKernelGGL((k_shadow<RING, MOTION>), dim3(cus * std::min(bpc, occ.shadow[ring_index(RING)][MOTION])),
                       dim3(kThreads), 0, st, P, (const JobDev*)W.jobs, W.gstack, one_job);
}

// tgrid = CUs x t.blocks_per_cu (the spill columns are sized for it); each
// stage kernel runs on as many of those blocks per CU as are resident at once.
void launch_trace(const Work& W, const Paths& P, const Tuning& t, unsigned tgrid, bool extend, bool shadow,
                  hipStream_t st, bool one_job, bool motion, bool first) {
  const unsigned bpc = std::max(1u, t.blocks_per_cu), cus = std::max(1u, tgrid / bpc);
  if (motion)  // scenes with moving meshes: one ring size
    launch_trace_stages<8, true>(W, P, cus, bpc, shadow, st, extend, one_job, first);
  else if (t.ring == 4)
    launch_trace_stages<4, false>(W, P, cus, bpc, shadow, st, extend, one_job, first);
  else if (t.ring == 16)
    launch_trace_stages<16, false>(W, P, cus, bpc, shadow, st, extend, one_job, first);
  else
    launch_trace_stages<8, false>(W, P, cus, bpc, shadow, st, extend, one_job, first);
}

// All stages of the pass passes[*pass_idx] on the stream; accumulate = false
// skips the film update.
void launch_pass(pyre_ctx* ctx, Work& W, const PassShape& shape, const Tuning& t, bool accumulate, hipStream_t st) {
  // grid-stride stages: as many blocks per CU as are resident at once (a
  // second, partial round of blocks would leave most of the GPU idle)
  const unsigned cus = (unsigned)std::max(1, ctx->cus);
  const StageOcc& occ = stage_occ();
  const unsigned gen_bpc = occ.generate, acc_bpc = occ.accumulate;
  const unsigned tgrid = trace_grid_of(ctx, t);
  const Sched sc = W.sched();
  Paths P = W.P;
  hipLaunchKernelGGL(k_counters, dim3(1), dim3(kCounters), 0, st, P.counters, sc);
  hipLaunchKernelGGL(k_generate, dim3(cus * gen_bpc), dim3(kThreads), 0, st, P, sc);
  for (uint32_t d = 0; d < shape.max_depth; ++d) {
    launch_trace(W, P, t, tgrid, true, false, st, shape.one_job, shape.motion, d == 0);
    // the shade stage's form is chosen on the device (shade_split)
    hipLaunchKernelGGL(k_shade_split, dim3(cus * occ.shade_split), dim3(kThreads), 0, st, P, sc, d, shape.one_job);
    if (t.shade_waves >= 4)
      hipLaunchKernelGGL(k_shade<4>, dim3(cus * occ.shade4), dim3(kThreads), 0, st, P, sc, d, shape.one_job);
    else
      hipLaunchKernelGGL(k_shade<1>, dim3(cus * occ.shade1), dim3(kThreads), 0, st, P, sc, d, shape.one_job);
    if (shape.nee) launch_trace(W, P, t, tgrid, false, true, st, shape.one_job, shape.motion, false);
    hipLaunchKernelGGL(k_rotate, dim3(1), dim3(kCounters), 0, st, P.counters, P.stats);
    std::swap(P.q_cur, P.q_next);
  }
  if (accumulate) hipLaunchKernelGGL(k_accumulate, dim3(cus * acc_bpc), dim3(kThreads), 0, st, P, sc);
}

bool graphs_enabled(const pyre_ctx* ctx) { return ctx->tune.eager == 0; }

hipGraphExec_t capture(hipStream_t st, const std::function<void()>& body) {
  hipGraph_t g = nullptr;
  PYRE_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  body();
  PYRE_HIP_CHECK(hipStreamEndCapture(st, &g));
  hipGraphExec_t exec = nullptr;
  hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  PYRE_HIP_CHECK(e);
  return exec;
}

template <class F>
int sguard(pyre_ctx* ctx, F&& f) {
  try {
    PYRE_HIP_CHECK(hipSetDevice(ctx->device));
    f();
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

}  // namespace

struct pyre_session {
  pyre_ctx* ctx = nullptr;
  pyre_scene* s = nullptr;
  pyre_camera_desc cam{};
  pyre_integrator_params ip{};
  uint32_t width = 0, height = 0, npix = 0;
  uint32_t per_pass = 0;  // samples per pixel a pass of the current pool holds (0: no pool yet)
  uint32_t train_paths = 0, slots = 1, vmax = 0, tcap = 0;
  Tuning tune;
  PassShape shape;
  Work W;
  float* acc = nullptr;
  double* acc64 = nullptr;
  float* welford = nullptr;
  unsigned long long* stats = nullptr;  // ray / cache counters over the session's lifetime
  bool bound = false;                   // W's job table and counters are this session's
  uint32_t split = 0;                   // MODE_SPLIT passes: continuation number
  bool cache_off = false;               // CACHE passes: image paths run as plain paths
  uint32_t spp_done = 0, next_sample = 0, train_frames = 0;
  uint64_t stats_seen[2] = {0, 0};  // ray counters already reported
  bool use_graph = true;
  hipGraphExec_t exec = nullptr;        // PLAIN pass
  hipGraphExec_t exec_cache = nullptr;   // CACHE pass trace + training targets
  hipGraphExec_t exec_finish = nullptr;  // cache inference + radiance + accumulate, for cache_of
  const pyre_nrc* cache_of = nullptr;
  ~pyre_session() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (exec_cache) (void)hipGraphExecDestroy(exec_cache);
    if (exec_finish) (void)hipGraphExecDestroy(exec_finish);
    dev_free(acc);
    dev_free(acc64);
    dev_free(welford);
    dev_free(stats);
  }
};

namespace {

// Training paths of a CACHE pass: how many, their sample index and pixel map.
struct TrainPaths {
  uint32_t paths = 0, s_base = 0, pix_mul = 1, pix_add = 0;
};

void set_pass(pyre_session* S, uint32_t s_base, uint32_t n_samples, uint32_t mode, float diffuse_rough,
              const TrainPaths& tr = TrainPaths{}) {
  PassDev pp{};
  pp.slice0 = 0;
  pp.n_slices = 1;
  pp.train0 = n_samples * S->npix;
  pp.paths = pp.train0 + tr.paths;
  pp.pixels = n_samples ? S->npix : 0u;
  pp.mode = mode;
  pp.pix_mul = tr.pix_mul;
  pp.pix_add = tr.pix_add;
  pp.t_sbase = tr.s_base;
  pp.vmax = S->W.vmax;
  pp.diffuse_rough = diffuse_rough;
  pp.split = S->split;
  pp.cache_off = S->cache_off ? 1u : 0u;
  SliceDev sl{0u, s_base, n_samples, 0u, 0u, S->spp_done};
  hipLaunchKernelGGL(k_set_pass, dim3(1), dim3(1), 0, S->ctx->stream, S->W.passes, S->W.slices, pp, sl);
}

// Samples [s_base, s_base + n_samples) of every pixel (a CACHE pass may
// also carry training paths, whose samples land in the training buffers).
// Every stage reads its pass from device memory, so one captured graph per
// kind of pass (plain, cache) serves any sample count; a context tuned
// `eager` launches the stages directly.
void session_pass(pyre_session* S, uint32_t s_base, uint32_t n_samples, uint32_t mode, pyre_nrc* nrc,
                  float diffuse_rough, bool accumulate, const TrainPaths& tr = TrainPaths{},
                  const std::function<void()>& between = nullptr) {
  hipStream_t st = S->ctx->stream;
  require(mode == MODE_CACHE || tr.paths == 0, "training paths ride in cache passes only");
  require(S->bound && (size_t)n_samples * S->npix + tr.paths <= S->W.cap, "pass exceeds the path pool");
  (void)stage_occ();  // occupancy queries before any capture
  set_pass(S, s_base, n_samples, mode, diffuse_rough, tr);
  if (mode == MODE_CACHE) {
    // trace (+ training targets), then `between` (the frame's training
    // steps), then cache inference for the paths that ended in the cache,
    // their radiance, and the film update
    require(nrc != nullptr, "cache pass without a radiance cache");
    require(accumulate, "cache passes always accumulate");
    auto trace = [&] {
      launch_pass(S->ctx, S->W, S->shape, S->tune, false, st);
      if (S->W.vmax) {
        const unsigned tt_bpc = stage_occ().train_targets;
        hipLaunchKernelGGL(k_train_targets, dim3((unsigned)std::max(1, S->ctx->cus) * tt_bpc), dim3(kThreads), 0, st, S->W.P, S->W.sched(),
                           S->W.tfeat, S->W.ttarget, S->W.tcap);
        hipLaunchKernelGGL(k_count_train, dim3(1), dim3(1), 0, st, S->W.P);
      }
    };
    auto finish = [&] {
      nrc_infer_dev_count(nrc, S->W.P.qfeat, S->W.P.qpath, S->W.P.counters + 3, (uint32_t)S->W.cap, S->W.P.qrgb,
                          st);
      const unsigned ap_bpc = stage_occ().cache_apply, acc_bpc = stage_occ().accumulate;
      hipLaunchKernelGGL(k_cache_apply, dim3((unsigned)std::max(1, S->ctx->cus) * ap_bpc), dim3(kThreads), 0, st, S->W.P);
      hipLaunchKernelGGL(k_accumulate, dim3((unsigned)std::max(1, S->ctx->cus) * acc_bpc), dim3(kThreads), 0, st, S->W.P, S->W.sched());
    };
    if (!S->use_graph) {
      trace();
      if (between) between();
      finish();
      return;
    }
    if (S->exec_finish && S->cache_of != nrc) {
      (void)hipGraphExecDestroy(S->exec_finish);
      S->exec_finish = nullptr;
    }
    if (!S->exec_cache) S->exec_cache = capture(st, trace);
    if (!S->exec_finish) {
      S->exec_finish = capture(st, finish);
      S->cache_of = nrc;
    }
    PYRE_HIP_CHECK(hipGraphLaunch(S->exec_cache, st));
    if (between) between();
    PYRE_HIP_CHECK(hipGraphLaunch(S->exec_finish, st));
    return;
  }
  if (mode == MODE_PLAIN && accumulate && S->use_graph) {
    if (!S->exec) S->exec = capture(st, [&] { launch_pass(S->ctx, S->W, S->shape, S->tune, true, st); });
    PYRE_HIP_CHECK(hipGraphLaunch(S->exec, st));
    return;
  }
  launch_pass(S->ctx, S->W, S->shape, S->tune, accumulate, st);
}

void clear_sums(pyre_ctx* ctx, float* acc, double* acc64, size_t npix) {
  hipStream_t st = ctx->stream;
  hipLaunchKernelGGL(k_fill_f32, dim3(nblocks(10 * npix, 256)), dim3(256), 0, st, acc, 10 * npix, 0.0f);
  hipLaunchKernelGGL(k_fill_f64, dim3(nblocks(3 * npix, 256)), dim3(256), 0, st, acc64, 3 * npix, 0.0);
}

void session_init(pyre_session* S, pyre_ctx* ctx, pyre_scene* s, const pyre_camera_desc& cam,
                  const pyre_integrator_params& ip, uint32_t w, uint32_t h, uint32_t train_paths) {
  require(w > 0 && h > 0, "width and height must be positive");
  require(s->built, "render before pyre_bvh_build");
  require((uint64_t)w * h < (1ull << 28), "image too large");
  require(train_paths < (1u << 24), "too many training paths");
  S->ctx = ctx;
  S->s = s;
  S->cam = cam;
  S->ip = ip;
  S->width = w;
  S->height = h;
  S->npix = w * h;
  S->tune = ctx_tuning(ctx);
  S->train_paths = train_paths;
  S->slots = std::max(1u, s->n_lights + s->n_mlights + (s->has_env ? 1u : 0u));
  // vertices recorded per training path (at most 64: one bit each in the target pass)
  S->vmax = train_paths ? std::min(64u, std::max(1u, ip.max_depth)) : 0;
  S->tcap = train_paths ? std::min<uint32_t>(train_paths * S->vmax, 1u << 20) : 0;
  S->shape.max_depth = ip.max_depth;
  S->shape.tree_depth = s->depth;
  S->shape.nee = s->n_lights + s->n_mlights + (s->has_env ? 1u : 0u) > 0;
  S->shape.one_job = true;
  S->shape.motion = s->has_motion;
  S->acc = dev_alloc<float>(10 * (size_t)S->npix);
  S->acc64 = dev_alloc<double>(3 * (size_t)S->npix);
  S->welford = dev_alloc<float>(3 * (size_t)S->npix);
  S->stats = dev_alloc<unsigned long long>(4);
  clear_sums(ctx, S->acc, S->acc64, S->npix);
  hipStream_t st = ctx->stream;
  hipLaunchKernelGGL(k_fill_f32, dim3(nblocks(3 * (size_t)S->npix, 256)), dim3(256), 0, st, S->welford,
                     3 * (size_t)S->npix, 0.0f);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, S->stats, (size_t)4);
  S->spp_done = 0;
  S->use_graph = graphs_enabled(ctx);
}

// A path pool for passes of up to `want` samples per pixel (plus the
// session's training paths): at most the tuning's pool and its share of free
// device memory, at least one sample per pixel. A pool that already holds
// that many is kept; growing it drops the captured pass graphs, which hold
// the old buffers' addresses.
void session_reserve(pyre_session* S, uint32_t want) {
  want = std::max(1u, want);
  const size_t tp = S->train_paths;
  const size_t fixed = tp * S->vmax * kVrec * sizeof(float4) + (size_t)S->tcap * 16 * sizeof(float);
  const bool cache = tp > 0;
  const size_t held = (S->W.cap > tp ? S->W.cap - tp : 0) / S->npix;
  uint64_t per = want;
  if (!(S->W.fits(S->npix + tp, S->slots, S->vmax, S->train_paths, S->tcap, cache, 1, 1, 1) && held >= want)) {
    const uint64_t limit = pool_limit(S->ctx, S->tune, path_bytes(S->slots, cache), fixed, S->W.bytes);
    per = limit > tp ? (limit - tp) / S->npix : 0;
    per = std::min<uint64_t>(per, ((1ull << 31) - 1 - tp) / S->npix);
    per = std::max<uint64_t>(1, std::min<uint64_t>(per, want));
    if (!S->W.fits(per * S->npix + tp, S->slots, S->vmax, S->train_paths, S->tcap, cache, 1, 1, 1)) {
      for (hipGraphExec_t* e : {&S->exec, &S->exec_cache, &S->exec_finish})
        if (*e) {
          (void)hipGraphExecDestroy(*e);
          *e = nullptr;
        }
      S->cache_of = nullptr;
      S->W.release();
      S->W.alloc(per * S->npix + tp, S->slots, S->vmax, S->train_paths, S->tcap, cache, 1, 1, 1);
      S->bound = false;
    }
  }
  S->per_pass = (uint32_t)per;
  if (S->bound) return;
  // a new pool, or one another render left in the context: this session's
  // job table, counters and spill columns
  S->W.ensure_gstack((size_t)trace_grid_of(S->ctx, S->tune) * kThreads, S->s->depth);
  S->W.P.stats = S->stats;  // the session's counters outlive its pools
  S->bound = true;
  hipStream_t st = S->ctx->stream;
  hipLaunchKernelGGL(k_set_job, dim3(1), dim3(1), 0, st, S->W.jobs,
                     job_dev(S->s, S->cam, S->ip, S->width, S->height, S->acc, S->acc64, S->welford));
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, st, S->W.pass_idx, 0u);
  PYRE_HIP_CHECK(hipGetLastError());
}

void session_render(pyre_session* S, uint32_t spp, uint32_t mode = MODE_PLAIN, float diffuse_rough = 0.25f) {
  if (!spp) return;
  session_reserve(S, spp);
  for (uint32_t s0 = 0; s0 < spp; s0 += S->per_pass) {
    uint32_t ns = std::min(S->per_pass, spp - s0);
    session_pass(S, S->next_sample, ns, mode, nullptr, diffuse_rough, true);
    S->next_sample += ns;
    S->spp_done += ns;
  }
  PYRE_HIP_CHECK(hipGetLastError());
}

void read_stats(pyre_ctx* ctx, const Work& W, uint64_t* ext, uint64_t* shadow) {
  unsigned long long hs[2] = {0, 0};
  PYRE_HIP_CHECK(hipMemcpyAsync(hs, W.P.stats, 16, hipMemcpyDeviceToHost, ctx->stream));
  PYRE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  *ext = hs[0];
  *shadow = hs[1];
}

// Rays traced since the previous report (the device counters run for the
// session's lifetime).
void finish_stats(pyre_session* S) {
  uint64_t ext = 0, sh = 0;
  unsigned long long hs[2] = {0, 0};
  PYRE_HIP_CHECK(hipMemcpyAsync(hs, S->stats, 16, hipMemcpyDeviceToHost, S->ctx->stream));
  PYRE_HIP_CHECK(hipStreamSynchronize(S->ctx->stream));
  ext = hs[0];
  sh = hs[1];
  S->s->last_extension_rays = ext - S->stats_seen[0];
  S->s->last_shadow_rays = sh - S->stats_seen[1];
  S->stats_seen[0] = ext;
  S->stats_seen[1] = sh;
}

void read_mean(pyre_ctx* ctx, const float* acc, size_t npix, pyre_pixel_sample* out, float inv_spp) {
  hipStream_t st = ctx->stream;
  const size_t n10 = 10 * npix;
  float* o = dev_alloc<float>(n10);
  hipLaunchKernelGGL(k_scale, dim3(nblocks(n10, 256)), dim3(256), 0, st, acc, n10, inv_spp, o);
  hipError_t e = hipMemcpyAsync(out, o, n10 * sizeof(float), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  dev_free(o);
  PYRE_HIP_CHECK(e);
}

// ------------------------------------------------------------ batch
// Passes over a shared path pool: jobs sorted by work (largest first), each
// pass filled greedily with whole samples of the jobs that still have
// samples left (at most one slice per job and pass).
struct Schedule {
  std::vector<SliceDev> slices;
  std::vector<PassDev> passes;
};

Schedule make_schedule(const pyre_job* jobs, size_t n, uint32_t pool) {
  std::vector<size_t> order(n);
  for (size_t i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    return (uint64_t)jobs[a].width * jobs[a].height * jobs[a].spp > (uint64_t)jobs[b].width * jobs[b].height * jobs[b].spp;
  });
  std::vector<uint32_t> left(n), next(n);
  for (size_t i = 0; i < n; ++i) {
    left[i] = jobs[i].spp;
    next[i] = jobs[i].spp_offset;
  }
  Schedule sc;
  for (;;) {
    PassDev pp{};
    pp.slice0 = (uint32_t)sc.slices.size();
    pp.pix_mul = 1;
    uint32_t free = pool;
    for (size_t oi = 0; oi < n; ++oi) {
      const size_t j = order[oi];
      if (!left[j]) continue;
      const uint32_t npix = jobs[j].width * jobs[j].height;
      const uint32_t ns = std::min(left[j], free / npix);
      if (!ns) continue;
      sc.slices.push_back(SliceDev{(uint32_t)j, next[j], ns, pp.paths, pp.pixels, next[j] - jobs[j].spp_offset});
      pp.paths += ns * npix;
      pp.pixels += npix;
      free -= ns * npix;
      left[j] -= ns;
      next[j] += ns;
    }
    pp.n_slices = (uint32_t)sc.slices.size() - pp.slice0;
    pp.train0 = pp.paths;
    if (!pp.n_slices) break;
    sc.passes.push_back(pp);
  }
  return sc;
}

struct Batch {
  pyre_ctx* ctx;
  const pyre_job* jobs;
  size_t n;
  PassShape shape;
  uint32_t slots = 1, max_npix = 0;
  uint64_t total_paths = 0;  // sum of width * height * spp over the jobs
  std::vector<float*> acc;   // job j's film sums (in the context's batch sums)
  std::vector<double*> acc64;
};

// Paths per pass of a batch: the tuning's pool within its share of free
// device memory (`reclaimable` = the cached pool's bytes), no more than the
// whole batch needs (the path buffers stay cached in the context), and at
// least one image of the largest job.
uint32_t batch_pool(const pyre_ctx* ctx, const Tuning& t, const Batch& B, size_t reclaimable) {
  const uint64_t lim = pool_limit(ctx, t, path_bytes(B.slots, false), 0, reclaimable);
  return (uint32_t)std::max<uint64_t>(std::min<uint64_t>(lim, B.total_paths), B.max_npix);
}

// Runs the whole schedule with tuning t into the jobs' sums (accumulate) or,
// for a trial, its first `max_passes` passes without touching them; returns
// the GPU time in ms.
float run_batch(Batch& B, const Tuning& t, bool accumulate, uint32_t max_passes, uint64_t* ext, uint64_t* shadow,
                uint32_t* n_passes, uint32_t* pool_used = nullptr) {
  pyre_ctx* ctx = B.ctx;
  hipStream_t st = ctx->stream;
  (void)stage_occ();  // occupancy queries before any capture
  // the context's cached path buffers, when they are large enough (one-shot
  // renders and batches leave theirs there)
  Work W;
  if (ctx->render_cache) {
    W = std::move(*static_cast<Work*>(ctx->render_cache.get()));
    ctx->render_cache.reset();
  }
  W.unbind();
  const uint32_t pool = batch_pool(ctx, t, B, W.bytes);
  if (pool_used) *pool_used = pool;
  Schedule sc = make_schedule(B.jobs, B.n, pool);
  const uint32_t np = std::min<uint32_t>((uint32_t)sc.passes.size(), max_passes);
  const size_t n_sl = std::max<size_t>(1, sc.slices.size()), n_pa = std::max<size_t>(1, sc.passes.size());
  if (!W.fits(pool, B.slots, 0, 0, 0, false, B.n, n_sl, n_pa)) {
    W.release();
    W.alloc(pool, B.slots, 0, 0, 0, false, B.n, n_sl, n_pa);
  }
  W.ensure_gstack((size_t)trace_grid_of(ctx, t) * kThreads, B.shape.tree_depth);
  std::vector<JobDev> jd(B.n);
  for (size_t j = 0; j < B.n; ++j)
    jd[j] = job_dev(B.jobs[j].scene, B.jobs[j].camera, B.jobs[j].integrator, B.jobs[j].width, B.jobs[j].height,
                    B.acc[j], B.acc64[j]);
  PYRE_HIP_CHECK(hipMemcpyAsync(W.jobs, jd.data(), jd.size() * sizeof(JobDev), hipMemcpyHostToDevice, st));
  PYRE_HIP_CHECK(hipMemcpyAsync(W.slices, sc.slices.data(), sc.slices.size() * sizeof(SliceDev),
                                hipMemcpyHostToDevice, st));
  PYRE_HIP_CHECK(hipMemcpyAsync(W.passes, sc.passes.data(), sc.passes.size() * sizeof(PassDev),
                                hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, st, W.pass_idx, 0u);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, W.P.stats, (size_t)4);
  PYRE_HIP_CHECK(hipStreamSynchronize(st));  // host tables may go out of scope after this
  auto one_pass = [&] {
    launch_pass(ctx, W, B.shape, t, accumulate, st);
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, st, W.pass_idx);
  };
  hipEvent_t e0, e1;
  PYRE_HIP_CHECK(hipEventCreate(&e0));
  PYRE_HIP_CHECK(hipEventCreate(&e1));
  hipGraphExec_t exec = nullptr;
  float ms = 0.0f;
  try {
    if (graphs_enabled(ctx) && np > 1) exec = capture(st, one_pass);
    PYRE_HIP_CHECK(hipEventRecord(e0, st));
    for (uint32_t k = 0; k < np; ++k) {
      if (exec)
        PYRE_HIP_CHECK(hipGraphLaunch(exec, st));
      else
        one_pass();
    }
    PYRE_HIP_CHECK(hipEventRecord(e1, st));
    PYRE_HIP_CHECK(hipEventSynchronize(e1));
    PYRE_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    PYRE_HIP_CHECK(hipGetLastError());
  } catch (...) {
    if (exec) (void)hipGraphExecDestroy(exec);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ext) read_stats(ctx, W, ext, shadow);
  if (n_passes) *n_passes = np;
  auto keep = std::make_shared<Work>();
  *keep = std::move(W);
  ctx->render_cache = std::static_pointer_cast<void>(keep);
  return ms;
}

// Trial runs of the first passes: pool size and blocks per CU at ring 8,
// then the ring size; the fastest paths-per-second wins.
Tuning autotune(Batch& B, Tuning base) {
  auto rate = [&](const Tuning& t) {
    uint64_t ext = 0, sh = 0;
    uint32_t np = 0;
    float ms = run_batch(B, t, false, 2, &ext, &sh, &np);
    return ms > 0.0f ? (double)(ext + sh) / ms : 0.0;
  };
  Tuning best = base;
  best.ring = 8;
  double best_rate = -1.0;
  for (uint32_t pool : {1u << 22, 1u << 24, 1u << 26})
    for (uint32_t bpc : {4u, 8u}) {
      Tuning t = best;
      t.pool_paths = pool;
      t.blocks_per_cu = bpc;
      double r = rate(t);
      if (r > best_rate) {
        best_rate = r;
        best.pool_paths = pool;
        best.blocks_per_cu = bpc;
      }
    }
  for (uint32_t ring : {4u, 16u}) {
    Tuning t = best;
    t.ring = ring;
    double r = rate(t);
    if (r > best_rate) {
      best_rate = r;
      best.ring = ring;
    }
  }
  return best;
}

}  // namespace

extern "C" {

int pyre_render_stats(pyre_ctx* ctx, pyre_scene* s, uint64_t* extension_rays, uint64_t* shadow_rays) {
  if (!ctx || !s) return PYRE_EINVAL;
  if (extension_rays) *extension_rays = s->last_extension_rays;
  if (shadow_rays) *shadow_rays = s->last_shadow_rays;
  return PYRE_OK;
}

static int render_one_shot(pyre_ctx* ctx, pyre_scene* s, const pyre_camera_desc* camera,
                           const pyre_integrator_params* integrator, uint32_t width, uint32_t height, uint32_t spp,
                           uint32_t spp_offset, pyre_pixel_sample* out, double* accum64, uint32_t mode,
                           float diffuse_rough, uint32_t split) {
  if (!ctx || !s || !camera || !integrator) return PYRE_EINVAL;
  return sguard(ctx, [&] {
    pyre_session S;
    // path buffers of the previous one-shot render on this context, if any
    if (ctx->render_cache) {
      S.W = std::move(*static_cast<Work*>(ctx->render_cache.get()));
      ctx->render_cache.reset();
    }
    session_init(&S, ctx, s, *camera, *integrator, width, height, 0);
    hipStream_t st = ctx->stream;
    const size_t n3 = 3 * (size_t)S.npix;
    if (accum64) PYRE_HIP_CHECK(hipMemcpyAsync(S.acc64, accum64, n3 * sizeof(double), hipMemcpyHostToDevice, st));
    S.next_sample = spp_offset;
    S.split = split;
    session_render(&S, spp, mode, diffuse_rough);
    if (out) read_mean(ctx, S.acc, S.npix, out, 1.0f / (float)spp);  // main.rs:255
    if (accum64) PYRE_HIP_CHECK(hipMemcpyAsync(accum64, S.acc64, n3 * sizeof(double), hipMemcpyDeviceToHost, st));
    finish_stats(&S);
    // keep the path buffers for the next one-shot render on this context,
    // with their own counters again (this session's die with it)
    S.W.unbind();
    auto keep = std::make_shared<Work>();
    *keep = std::move(S.W);
    ctx->render_cache = std::static_pointer_cast<void>(keep);
  });
}

int pyre_render(pyre_ctx* ctx, pyre_scene* s, const pyre_camera_desc* camera, const pyre_integrator_params* integrator,
                uint32_t width, uint32_t height, uint32_t spp, uint32_t spp_offset, pyre_pixel_sample* out,
                double* accum64) {
  return render_one_shot(ctx, s, camera, integrator, width, height, spp, spp_offset, out, accum64, MODE_PLAIN, 0.25f,
                         0u);
}

int pyre_render_split(pyre_ctx* ctx, pyre_scene* s, const pyre_camera_desc* camera,
                      const pyre_integrator_params* integrator, uint32_t width, uint32_t height, uint32_t spp,
                      uint32_t spp_offset, float diffuse_roughness, uint32_t continuation, double* accum64) {
  if (!accum64) return PYRE_EINVAL;
  return render_one_shot(ctx, s, camera, integrator, width, height, spp, spp_offset, nullptr, accum64, MODE_SPLIT,
                         diffuse_roughness, continuation);
}

int pyre_render_batch(pyre_ctx* ctx, const pyre_job* jobs, size_t n, const pyre_batch_params* params,
                      pyre_batch_stats* stats) {
  if (!ctx || (!jobs && n)) return PYRE_EINVAL;
  return sguard(ctx, [&] {
    if (stats) *stats = pyre_batch_stats{};
    if (n == 0) return;
    require(n < (1u << 20), "too many jobs");
    Batch B{ctx, jobs, n, {}, 1, 0, 0, {}, {}};
    for (size_t j = 0; j < n; ++j) {
      const pyre_job& J = jobs[j];
      require(J.scene && J.scene->built, "job " + std::to_string(j) + ": scene missing or BVH not built");
      require(J.scene->ctx == ctx, "job " + std::to_string(j) + ": scene belongs to another context");
      require(J.width > 0 && J.height > 0 && (uint64_t)J.width * J.height <= (1u << 24),
              "job " + std::to_string(j) + ": bad image size");
      require(J.spp > 0, "job " + std::to_string(j) + ": spp must be positive");
      B.max_npix = std::max(B.max_npix, J.width * J.height);
      B.total_paths += (uint64_t)J.width * J.height * J.spp;
      B.shape.max_depth = std::max(B.shape.max_depth, J.integrator.max_depth);
      B.shape.tree_depth = std::max(B.shape.tree_depth, J.scene->depth);
      const uint32_t sl = J.scene->n_lights + J.scene->n_mlights + (J.scene->has_env ? 1u : 0u);
      B.shape.nee = B.shape.nee || sl > 0;
      B.shape.one_job = n == 1;
      B.shape.motion = B.shape.motion || J.scene->has_motion;
      B.slots = std::max(B.slots, sl);
    }
    // every job's sums in the context's batch buffers, job after job (grown
    // here when the batch has more pixels than any before it; the previous
    // batch has finished with them: a batch returns after its copies)
    std::vector<size_t> off(n + 1, 0);
    for (size_t j = 0; j < n; ++j) off[j + 1] = off[j] + (size_t)jobs[j].width * jobs[j].height;
    if (ctx->batch_npix < off[n]) {
      ctx->batch_npix = 0;
      dev_free(ctx->batch_acc);
      dev_free(ctx->batch_acc64);
      ctx->batch_acc = dev_alloc<float>(10 * off[n]);
      ctx->batch_acc64 = dev_alloc<double>(3 * off[n]);
      ctx->batch_npix = off[n];
    }
    clear_sums(ctx, ctx->batch_acc, ctx->batch_acc64, off[n]);
    B.acc.assign(n, nullptr);
    B.acc64.assign(n, nullptr);
    hipStream_t st = ctx->stream;
    // job buffers may be host or device memory (hipMemcpyDefault finds out)
    for (size_t j = 0; j < n; ++j) {
      const size_t npix = off[j + 1] - off[j];
      B.acc[j] = ctx->batch_acc + 10 * off[j];
      B.acc64[j] = ctx->batch_acc64 + 3 * off[j];
      if (jobs[j].accum64)
        PYRE_HIP_CHECK(hipMemcpyAsync(B.acc64[j], jobs[j].accum64, 3 * npix * sizeof(double), hipMemcpyDefault, st));
    }
    Tuning t = ctx_tuning(ctx);
    if (params) {
      if (params->pool_paths) t.pool_paths = params->pool_paths;
      if (params->trace_blocks_per_cu) t.blocks_per_cu = params->trace_blocks_per_cu;
      if (params->stack_ring == 4 || params->stack_ring == 8 || params->stack_ring == 16) t.ring = params->stack_ring;
      if (params->autotune) t = autotune(B, t);
    }
    uint64_t ext = 0, sh = 0;
    uint32_t np = 0;
    uint32_t pool_used = 0;
    float ms = run_batch(B, t, true, 0xFFFFFFFFu, &ext, &sh, &np, &pool_used);
    // means (the sums scaled in place: the batch's sums are not kept) and
    // FP64 sums, copied out with one synchronisation at the end
    for (size_t j = 0; j < n; ++j) {
      const size_t npix = off[j + 1] - off[j];
      if (jobs[j].out) {
        hipLaunchKernelGGL(k_scale, dim3(nblocks(10 * npix, 256)), dim3(256), 0, st, B.acc[j], 10 * npix,
                           1.0f / (float)jobs[j].spp, B.acc[j]);
        PYRE_HIP_CHECK(hipMemcpyAsync(jobs[j].out, B.acc[j], 10 * npix * sizeof(float), hipMemcpyDefault, st));
      }
      if (jobs[j].accum64)
        PYRE_HIP_CHECK(hipMemcpyAsync(jobs[j].accum64, B.acc64[j], 3 * npix * sizeof(double), hipMemcpyDefault, st));
    }
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
    if (stats) {
      stats->extension_rays = ext;
      stats->shadow_rays = sh;
      stats->passes = np;
      stats->gpu_ms = ms;
      stats->pool_paths = pool_used;
      stats->trace_blocks_per_cu = t.blocks_per_cu;
      stats->stack_ring = t.ring;
    }
  });
}

int pyre_session_create(pyre_ctx* ctx, pyre_scene* s, const pyre_camera_desc* camera,
                        const pyre_integrator_params* integrator, uint32_t width, uint32_t height,
                        uint32_t nrc_train_paths, pyre_session** out) {
  if (!ctx || !s || !camera || !integrator || !out) return PYRE_EINVAL;
  *out = nullptr;
  pyre_session* S = new (std::nothrow) pyre_session();
  if (!S) return PYRE_ENOMEM;
  int rc = sguard(ctx, [&] {
    session_init(S, ctx, s, *camera, *integrator, width, height, nrc_train_paths);
    PYRE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
  if (rc != PYRE_OK) {
    delete S;
    return rc;
  }
  *out = S;
  return PYRE_OK;
}

void pyre_session_destroy(pyre_session* S) {
  if (!S) return;
  (void)hipSetDevice(S->ctx->device);
  (void)hipStreamSynchronize(S->ctx->stream);
  delete S;
}

int pyre_session_render(pyre_ctx* ctx, pyre_session* S, uint32_t spp) {
  if (!ctx || !S) return PYRE_EINVAL;
  return sguard(ctx, [&] {
    session_render(S, spp);
    finish_stats(S);
  });
}

int pyre_session_reserve(pyre_ctx* ctx, pyre_session* S, uint32_t spp_per_pass) {
  if (!ctx || !S) return PYRE_EINVAL;
  return sguard(ctx, [&] {
    session_reserve(S, spp_per_pass);
    PYRE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

int pyre_session_render_nrc(pyre_ctx* ctx, pyre_session* S, pyre_nrc* nrc, const pyre_nrc_render_params* p,
                            float* losses) {
  if (!ctx || !S || !nrc || !p) return PYRE_EINVAL;
  return sguard(ctx, [&] {
    require(S->vmax > 0, "session created without radiance-cache buffers (nrc_train_paths = 0)");
    require(p->train_paths > 0 && p->train_paths <= S->train_paths,
            "train_paths exceeds the session's nrc_train_paths");
    require(p->train_batch >= 16, "train_batch below 16");
    hipStream_t st = ctx->stream;
    session_reserve(S, p->spp_per_frame);
    // Per-frame loss sums and step counts stay on the device until the end:
    // the training steps are queued without knowing how many samples the
    // frame produced (each step reads the count and skips itself when its
    // batch would hold fewer than 16), so nothing waits on the host.
    const size_t nf = std::max(1u, p->frames);
    float* dloss = dev_alloc<float>(2 * nf);
    uint32_t* dsteps = reinterpret_cast<uint32_t*>(dloss + nf);
    hipLaunchKernelGGL(k_fill_f32, dim3(nblocks(2 * nf, 256)), dim3(256), 0, st, dloss, 2 * nf, 0.0f);
    struct Free {
      float* p;
      ~Free() { dev_free(p); }
    } free_loss{dloss};
    for (uint32_t f = 0; f < p->frames; ++f) {
      // One pass per frame (or several, past the pool size) over the frame's
      // spp_per_frame samples (none in a warmup frame) and, in its last pass,
      // the frame's training paths (full unbiased paths from scattered
      // pixels, their own sample streams). After tracing, the frame trains on
      // its own samples, then the paths that ended in the cache are given
      // the updated cache's radiance. The plain frames after the warmup
      // trace their image samples to the end, as the plain renderer does.
      const uint32_t spp = f < p->warmup_frames ? 0u : p->spp_per_frame;
      S->cache_off = f >= p->warmup_frames && f - p->warmup_frames < p->plain_frames;
      TrainPaths tr;
      tr.paths = p->train_paths;
      tr.s_base = kTrainStreamBase + S->train_frames;
      tr.pix_mul = (uint32_t)(2654435761ull % S->npix) | 1u;
      tr.pix_add = (uint32_t)(pixel_seed(S->train_frames, 7, 11) % S->npix);
      auto train = [&] {
        for (uint32_t k = 0; k < p->train_steps; ++k) {
          const uint64_t off = (uint64_t)k * p->train_batch;
          if (off + 16 > S->W.tcap) break;  // past the sample buffer: never filled
          nrc_train_step_counted(nrc, S->W.tfeat, S->W.ttarget, S->W.P.counters + 4, S->W.tcap, (uint32_t)off,
                                 p->train_batch, st, dloss + f, dsteps + f);
        }
      };
      const uint32_t n_pass = spp ? (spp + S->per_pass - 1) / S->per_pass : 1u;
      for (uint32_t i = 0, s0 = 0; i < n_pass; ++i) {
        const uint32_t ns = std::min(S->per_pass, spp - s0);
        const bool last = i + 1 == n_pass;
        session_pass(S, S->next_sample, ns, MODE_CACHE, nrc, p->diffuse_roughness, true,
                     last ? tr : TrainPaths{}, last ? std::function<void()>(train) : nullptr);
        S->next_sample += ns;
        S->spp_done += ns;
        s0 += ns;
      }
      S->train_frames++;
    }
    S->cache_off = false;
    if (p->frames) {
      std::vector<float> h(2 * nf);
      PYRE_HIP_CHECK(hipMemcpyAsync(h.data(), dloss, 8 * nf, hipMemcpyDeviceToHost, st));
      PYRE_HIP_CHECK(hipStreamSynchronize(st));
      float last = 0.0f;
      for (uint32_t f = 0; f < p->frames; ++f) {
        uint32_t n;
        std::memcpy(&n, &h[nf + f], 4);
        const float l = n ? h[f] / (float)n : 0.0f;
        if (losses) losses[f] = l;
        if (n) last = l;
      }
      nrc_set_last_loss(nrc, last);
    }
    finish_stats(S);
  });
}

int pyre_session_read(pyre_ctx* ctx, pyre_session* S, pyre_pixel_sample* mean, double* radiance_sum64,
                      uint32_t* spp_done) {
  if (!ctx || !S) return PYRE_EINVAL;
  return sguard(ctx, [&] {
    hipStream_t st = ctx->stream;
    if (mean) read_mean(ctx, S->acc, S->npix, mean, S->spp_done ? 1.0f / (float)S->spp_done : 0.0f);
    if (radiance_sum64)
      PYRE_HIP_CHECK(hipMemcpyAsync(radiance_sum64, S->acc64, 3 * sizeof(double) * (size_t)S->npix,
                                    hipMemcpyDeviceToHost, st));
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
    if (spp_done) *spp_done = S->spp_done;
  });
}

int pyre_session_pack(pyre_ctx* ctx, pyre_session* S, uint32_t what, void* out) {
  if (!ctx || !S || !out || what > PYRE_PACK_EXR) return PYRE_EINVAL;
  return sguard(ctx, [&] {
    hipStream_t st = ctx->stream;
    const float inv = S->spp_done ? 1.0f / (float)S->spp_done : 0.0f;
    const size_t bytes = what == PYRE_PACK_EXR ? 40 * (size_t)S->npix
                         : what == PYRE_PACK_DEPTH_PNG ? (size_t)S->npix : 3 * (size_t)S->npix;
    uint8_t* buf = dev_alloc<uint8_t>(bytes);
    uint32_t* range = dev_alloc<uint32_t>(2);
    const uint32_t init[2] = {0x7F800000u, 0u};
    hipError_t e = hipMemcpyAsync(range, init, sizeof(init), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && what == PYRE_PACK_DEPTH_PNG)
      hipLaunchKernelGGL(k_depth_range, dim3(std::min<size_t>(nblocks(S->npix, 256), 1024)), dim3(256), 0, st, S->acc,
                         S->npix, inv, range);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_pack, dim3(nblocks(S->npix, 256)), dim3(256), 0, st, S->acc, S->npix, inv, what, range,
                         gamma8_thresholds(), buf, (float*)buf);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, buf, bytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    dev_free(buf);
    dev_free(range);
    PYRE_HIP_CHECK(e);
  });
}

int pyre_session_stats(pyre_ctx* ctx, pyre_session* S, uint64_t out[4]) {
  if (!ctx || !S || !out) return PYRE_EINVAL;
  return sguard(ctx, [&] {
    unsigned long long hs[4] = {0, 0, 0, 0};
    PYRE_HIP_CHECK(hipMemcpyAsync(hs, S->stats, sizeof(hs), hipMemcpyDeviceToHost, ctx->stream));
    PYRE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    out[0] = hs[0];
    out[1] = hs[1];
    out[2] = hs[2];
    out[3] = hs[3];
  });
}

int pyre_session_read_viewer(pyre_ctx* ctx, pyre_session* S, float* rgb) {
  if (!ctx || !S || !rgb) return PYRE_EINVAL;
  return sguard(ctx, [&] {
    PYRE_HIP_CHECK(hipMemcpyAsync(rgb, S->welford, 3 * sizeof(float) * (size_t)S->npix, hipMemcpyDeviceToHost,
                                  ctx->stream));
    PYRE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

int pyre_session_reset(pyre_ctx* ctx, pyre_session* S) {
  if (!ctx || !S) return PYRE_EINVAL;
  return sguard(ctx, [&] {
    clear_sums(ctx, S->acc, S->acc64, S->npix);
    hipLaunchKernelGGL(k_fill_f32, dim3(nblocks(3 * (size_t)S->npix, 256)), dim3(256), 0, ctx->stream, S->welford,
                       3 * (size_t)S->npix, 0.0f);
    S->spp_done = 0;
    PYRE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

}  // extern "C"
