// Batched closest-hit / any-hit kernels (Scene::intersect / occluded over a
// ray array) built on the traversal cores in traverse.hpp.
//
// Default: the 4-wide tree, persistent threads with dynamic ray fetch, an
// 8-entry LDS stack ring per thread spilling to a per-thread global column. The binary
// tree kernels (LDS stack as deep as the tree) remain selectable with
// PYRE_TRAVERSAL=binary for comparison.
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "traverse.hpp"

namespace pyre {
namespace {

constexpr int kTraceThreads = 256;
constexpr int kRing = 8;

template <bool ANY, int STACK>
__global__ __launch_bounds__(kTraceThreads) void k_trace(const Node2* __restrict__ nodes,
                                                         const float4* __restrict__ tris,
                                                         const float4* __restrict__ rays, uint32_t n,
                                                         pyre_hit8* __restrict__ out_hit,
                                                         float2* __restrict__ out_uv,
                                                         uint32_t* __restrict__ out_occ, int32_t* gstack,
                                                         int gdepth, uint32_t* overflow) {
  __shared__ int32_t lds_stack[STACK > 0 ? STACK * kTraceThreads : 1];
  const uint32_t i = blockIdx.x * kTraceThreads + threadIdx.x;
  if (i >= n) return;
  int32_t* stack;
  size_t stride;
  int cap;
  if constexpr (STACK > 0) {
    stack = lds_stack + threadIdx.x;
    stride = kTraceThreads;
    cap = STACK;
  } else {
    stack = gstack + i;
    stride = n;
    cap = gdepth;
  }
  const float4 ra = rays[2 * (size_t)i], rb = rays[2 * (size_t)i + 1];
  const TRay r = make_tray(v3(ra.x, ra.y, ra.z), v3(rb.x, rb.y, rb.z), ra.w, rb.w);
  THit h;
  if (!traverse<ANY>(nodes, tris, r, stack, stride, cap, &h)) atomicOr(overflow, 1u);
  if (ANY) {
    out_occ[i] = h.gid != PYRE_MISS ? 1u : 0u;
  } else {
    pyre_hit8 o;
    o.prim = h.gid;
    o.t = h.t;
    out_hit[i] = o;
    if (out_uv) out_uv[i] = make_float2(h.u, h.v);
  }
}

// Persistent 4-wide traversal with dynamic ray fetch: a lane that finishes
// its ray takes the next one of its wave's chunk as soon as a quarter of the
// wave is idle, instead of waiting for the wave's slowest ray.
// MOTION: scenes with moving meshes; times = per-ray shutter times (null: 0).
template <bool ANY, int RING, bool MOTION = false>
__global__ __launch_bounds__(kTraceThreads, 8) void k_trace4(const Node4* __restrict__ nodes,
                                                          const float4* __restrict__ tris,
                                                          const float4* __restrict__ rays, uint32_t n,
                                                          pyre_hit8* __restrict__ out_hit,
                                                          float2* __restrict__ out_uv,
                                                          uint32_t* __restrict__ out_occ, int2* gstack,
                                                          uint32_t* fetch, const float* __restrict__ times,
                                                          const float4* __restrict__ motion) {
  __shared__ int2 ring[RING * kTraceThreads];
  const uint32_t tid = blockIdx.x * kTraceThreads + threadIdx.x;
  const uint32_t nthreads = gridDim.x * kTraceThreads;
  HybridStack<RING> st{ring + threadIdx.x, kTraceThreads, gstack + tid, nthreads, 0, 0};
  WaveFeed<256, 64> feed(n, fetch);
  Trav4<ANY, RING, MOTION> tv;
  uint32_t ray = 0;
  bool active = false;
  for (;;) {
    const uint64_t idle = __ballot(!active);
    if (__popcll(idle) >= kRefillLanes) {
      if (feed.take(idle, !active, ray)) {
        const float4 ra = rays[2 * (size_t)ray], rb = rays[2 * (size_t)ray + 1];
        const float tm = MOTION && times ? times[ray] : 0.0f;
        tv.start(make_tray(v3(ra.x, ra.y, ra.z), v3(rb.x, rb.y, rb.z), ra.w, rb.w, tm), st);
        active = true;
      }
      if (!__any(active)) break;  // slice exhausted and every lane idle
    }
    if (active && tv.step(nodes, tris, st, motion)) {
      THit h;
      tv.result(&h);
      if (ANY) {
        out_occ[ray] = h.gid != PYRE_MISS ? 1u : 0u;
      } else {
        pyre_hit8 o;
        o.prim = h.gid;
        o.t = h.t;
        out_hit[ray] = o;
        if (out_uv) out_uv[ray] = make_float2(h.u, h.v);
      }
      active = false;
    }
  }
}

// Grid-stride variant (each lane keeps its ray to the end), selectable with
// PYRE_TRACE_FETCH=static for comparison.
template <bool ANY>
__global__ __launch_bounds__(kTraceThreads) void k_trace4s(const Node4* __restrict__ nodes,
                                                           const float4* __restrict__ tris,
                                                           const float4* __restrict__ rays, uint32_t n,
                                                           pyre_hit8* __restrict__ out_hit,
                                                           float2* __restrict__ out_uv,
                                                           uint32_t* __restrict__ out_occ, int2* gstack) {
  __shared__ int2 ring[kRing * kTraceThreads];
  const uint32_t tid = blockIdx.x * kTraceThreads + threadIdx.x;
  const uint32_t nthreads = gridDim.x * kTraceThreads;
  HybridStack<kRing> st{ring + threadIdx.x, kTraceThreads, gstack + tid, nthreads, 0, 0};
  for (uint32_t i = tid; i < n; i += nthreads) {
    const float4 ra = rays[2 * (size_t)i], rb = rays[2 * (size_t)i + 1];
    THit h;
    traverse4<ANY, kRing>(nodes, tris, make_tray(v3(ra.x, ra.y, ra.z), v3(rb.x, rb.y, rb.z), ra.w, rb.w), st, &h);
    if (ANY) {
      out_occ[i] = h.gid != PYRE_MISS ? 1u : 0u;
    } else {
      pyre_hit8 o;
      o.prim = h.gid;
      o.t = h.t;
      out_hit[i] = o;
      if (out_uv) out_uv[i] = make_float2(h.u, h.v);
    }
  }
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// LDS stack ring entries per thread (PYRE_STACK_RING = 4, 8 or 16; default 8).
int trace_ring() {
  const char* e = std::getenv("PYRE_STACK_RING");
  const int r = e ? std::atoi(e) : 8;
  return (r == 4 || r == 16) ? r : 8;
}

bool use_binary() {
  const char* e = std::getenv("PYRE_TRAVERSAL");
  return e && std::strcmp(e, "binary") == 0;
}

template <bool ANY>
void launch(pyre_scene* s, const float4* rays, size_t n, pyre_hit8* hit, float2* uv, uint32_t* occ, hipStream_t st,
            bool check, const float* times) {
  require(s->built, "trace before pyre_bvh_build");
  require(n <= 0xFFFFFFFFull, "too many rays in one call");
  if (n == 0) return;
  uint32_t* ovf = s->d_flag;
  if (s->has_motion) {
    // moving meshes: the persistent 4-wide kernel with shifted leaf tests
    int2* gst = trace_gstack(s);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, s->d_fetch, 1u);
    hipLaunchKernelGGL((k_trace4<ANY, 8, true>), dim3(trace_grid(s)), dim3(kTraceThreads), 0, st, s->d_nodes4,
                       s->d_tris, rays, (uint32_t)n, hit, uv, occ, gst, s->d_fetch, times, s->d_motion);
    PYRE_HIP_CHECK(hipGetLastError());
    return;
  }
  if (!use_binary()) {
    int2* gst = trace_gstack(s);
    const char* fe = std::getenv("PYRE_TRACE_FETCH");
    if (fe && std::strcmp(fe, "static") == 0) {
      hipLaunchKernelGGL((k_trace4s<ANY>), dim3(trace_grid(s)), dim3(kTraceThreads), 0, st, s->d_nodes4, s->d_tris,
                         rays, (uint32_t)n, hit, uv, occ, gst);
      PYRE_HIP_CHECK(hipGetLastError());
      return;
    }
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, s->d_fetch, 1u);
    const int ring = trace_ring();
    if (ring == 4)
      hipLaunchKernelGGL((k_trace4<ANY, 4>), dim3(trace_grid(s)), dim3(kTraceThreads), 0, st, s->d_nodes4, s->d_tris,
                         rays, (uint32_t)n, hit, uv, occ, gst, s->d_fetch, nullptr, nullptr);
    else if (ring == 16)
      hipLaunchKernelGGL((k_trace4<ANY, 16>), dim3(trace_grid(s)), dim3(kTraceThreads), 0, st, s->d_nodes4,
                         s->d_tris, rays, (uint32_t)n, hit, uv, occ, gst, s->d_fetch, nullptr, nullptr);
    else
      hipLaunchKernelGGL((k_trace4<ANY, 8>), dim3(trace_grid(s)), dim3(kTraceThreads), 0, st, s->d_nodes4, s->d_tris,
                         rays, (uint32_t)n, hit, uv, occ, gst, s->d_fetch, nullptr, nullptr);
    PYRE_HIP_CHECK(hipGetLastError());
    return;
  }
  dim3 grid((unsigned)((n + kTraceThreads - 1) / kTraceThreads)), block(kTraceThreads);
  int32_t* gst = nullptr;
  const uint32_t need = s->depth + 1;
  if (need <= 16)
    hipLaunchKernelGGL((k_trace<ANY, 16>), grid, block, 0, st, s->d_nodes, s->d_tris, rays, (uint32_t)n, hit, uv, occ,
                       gst, 0, ovf);
  else if (need <= 32)
    hipLaunchKernelGGL((k_trace<ANY, 32>), grid, block, 0, st, s->d_nodes, s->d_tris, rays, (uint32_t)n, hit, uv, occ,
                       gst, 0, ovf);
  else if (need <= 48)
    hipLaunchKernelGGL((k_trace<ANY, 48>), grid, block, 0, st, s->d_nodes, s->d_tris, rays, (uint32_t)n, hit, uv, occ,
                       gst, 0, ovf);
  else {
    gst = dev_alloc<int32_t>((size_t)need * n);
    hipLaunchKernelGGL((k_trace<ANY, 0>), grid, block, 0, st, s->d_nodes, s->d_tris, rays, (uint32_t)n, hit, uv, occ,
                       gst, (int)need, ovf);
  }
  PYRE_HIP_CHECK(hipGetLastError());
  if (gst) {
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
    dev_free(gst);
  }
  if (check) {
    uint32_t h_ovf = 0;
    PYRE_HIP_CHECK(hipMemcpyAsync(&h_ovf, ovf, 4, hipMemcpyDeviceToHost, st));
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
    if (h_ovf) throw Error(PYRE_EOVERFLOW, "traversal stack overflow (tree deeper than recorded depth)");
  }
}

}  // namespace

// Persistent grid of the 4-wide kernels: blocks per CU (PYRE_TRACE_BLOCKS_PER_CU,
// default 8) x CUs.
unsigned trace_grid(pyre_scene* s) {
  int per_cu = 8;
  if (const char* e = std::getenv("PYRE_TRACE_BLOCKS_PER_CU")) per_cu = std::max(1, std::atoi(e));
  return (unsigned)std::max(1, s->ctx->cus) * (unsigned)per_cu;
}

// Spill columns for the persistent kernels: a 4-wide node visit leaves at
// most 3 pending siblings per level of the current path, and the 4-wide
// tree is no deeper than the binary one: 3 * depth entries (+ 8 spare).
int2* trace_gstack(pyre_scene* s) {
  const size_t threads = (size_t)trace_grid(s) * kTraceThreads;
  const size_t need = threads * (3 * (size_t)s->depth + 8);
  if (!s->d_fetch) s->d_fetch = dev_alloc<uint32_t>(64);
  if (need > s->gstack_cap) {
    PYRE_HIP_CHECK(hipDeviceSynchronize());
    dev_free(s->d_gstack);
    s->d_gstack = dev_alloc<int2>(need);
    s->gstack_cap = need;
  }
  return s->d_gstack;
}

void trace_closest(pyre_scene* s, const float4* rays, size_t n, pyre_hit8* hits, float2* uv, hipStream_t st,
                   bool check, const float* times) {
  launch<false>(s, rays, n, hits, uv, nullptr, st, check, times);
}

void trace_any(pyre_scene* s, const float4* rays, size_t n, uint32_t* occluded, hipStream_t st, bool check,
               const float* times) {
  launch<true>(s, rays, n, nullptr, nullptr, occluded, st, check, times);
}

}  // namespace pyre
