// Batched closest-hit / any-hit kernels (Scene::intersect / occluded over a
// ray array) built on the traversal cores in traverse.hpp.
//
// Default: the 4-wide tree with child boxes quantised to 16 bits (Node4Q),
// persistent threads with dynamic ray fetch, an 8-entry LDS stack ring per
// thread spilling to a per-thread global column. The same kernel on the
// exact float boxes, the binary-tree kernels (LDS stack as deep as the
// tree), a grid-stride form of the 4-wide kernel and a form that tests leaf
// triangles one per step remain selectable through the context's tuning
// (pyre_tuning.traversal) for comparison.
#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>

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
// its ray takes the next one of its wave's chunk as soon as half of the
// wave is idle, instead of waiting for the wave's slowest ray.
// MOTION: scenes with moving meshes; times = per-ray shutter times (null: 0).
// LEAFQ: leaf triangles tested one per step beside the node visits
// (Trav4Q, pyre_tuning.traversal = 3); false: all of a node's hit leaves
// tested at its visit (Trav4, the default).
// QNODE: the tree of quantised nodes (nodes is a Node4Q array, grid its
// quantisation grid); otherwise nodes is a Node4 array.
template <bool ANY, int RING, bool MOTION = false, bool LEAFQ = false, bool QNODE = false>
__global__ __launch_bounds__(kTraceThreads, 8) void k_trace4(const void* __restrict__ nodes_v,
                                                          const float4* __restrict__ tris,
                                                          const float4* __restrict__ rays, uint32_t n,
                                                          pyre_hit8* __restrict__ out_hit,
                                                          float2* __restrict__ out_uv,
                                                          uint32_t* __restrict__ out_occ, int2* gstack,
                                                          uint32_t* fetch, const float* __restrict__ times,
                                                          const float4* __restrict__ motion, QGrid grid) {
  using TV = std::conditional_t<LEAFQ, Trav4Q<ANY, RING, MOTION>, Trav4<ANY, RING, MOTION, QNODE>>;
  const auto* __restrict__ nodes = static_cast<const typename TV::NodeT*>(nodes_v);
  __shared__ int2 ring[RING * kTraceThreads];
  const uint32_t tid = blockIdx.x * kTraceThreads + threadIdx.x;
  const uint32_t nthreads = gridDim.x * kTraceThreads;
  HybridStack<RING> st{ring + threadIdx.x, kTraceThreads, gstack + tid, nthreads, 0, 0};
  WaveFeed<256, 64> feed(n, fetch);
  TV tv;
  uint32_t ray = 0;
  bool active = false;
  for (;;) {
    const uint64_t idle = __ballot(!active);
    if (__popcll(idle) >= kRefillLanes) {
      if (feed.take(idle, !active, ray)) {
        const float4 ra = rays[2 * (size_t)ray], rb = rays[2 * (size_t)ray + 1];
        const float tm = MOTION && times ? times[ray] : 0.0f;
        tv.start(make_tray(v3(ra.x, ra.y, ra.z), v3(rb.x, rb.y, rb.z), ra.w, rb.w, tm), st, &grid);
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
// pyre_tuning.traversal = 2 for comparison.
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

// Persistent grid of the 4-wide kernels: the tuning's blocks per CU x CUs.
unsigned trace_grid(const pyre_ctx* ctx) {
  return (unsigned)std::max(1, ctx->cus) * tune_blocks_per_cu(ctx);
}

// Spill-column entries per thread: a 4-wide node visit leaves at most 3
// pending siblings per level of the current path, and the 4-wide tree is no
// deeper than the binary one: 3 * depth entries (+ 8 spare).
size_t spill_entries(uint32_t depth) { return 3 * (size_t)depth + 8; }

template <bool ANY>
void launch(pyre_scene* s, const float4* rays, size_t n, pyre_hit8* hit, float2* uv, uint32_t* occ, hipStream_t st,
            bool check, const float* times) {
  require(s->built, "trace before pyre_bvh_build");
  require(n <= 0xFFFFFFFFull, "too many rays in one call");
  if (n == 0) return;
  pyre_ctx* ctx = s->ctx;
  require(ctx->trace_gstack && ctx->trace_depth >= s->depth, "trace scratch smaller than the scene's tree");
  const uint32_t mode = ctx->tune.traversal;
  require(mode <= 4, "pyre_tuning.traversal must be 0 to 4");
  require(mode != 1 || s->depth + 1 <= 48, "binary traversal supports trees of depth <= 47");
  // calls sharing the context's scratch run in the order they were made,
  // whatever streams they are on
  PYRE_HIP_CHECK(hipStreamWaitEvent(st, ctx->trace_ev, 0));
  const unsigned grid = trace_grid(ctx);
  // spill columns are laid out for the grid in use (gstride = its threads)
  int2* gst = ctx->trace_gstack;
  if (mode == 0 || mode == 4 || s->has_motion) {
    // the persistent 4-wide kernel: mode 0 on the quantised tree, mode 4 on
    // the exact float boxes; scenes with moving meshes shift the leaf tests
    // and take the exact boxes (the quantised form's extra per-ray terms
    // beside the motion state would spill registers)
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, ctx->trace_fetch, 1u);
    const bool q = mode == 0 && !s->has_motion;
    const void* nodes = q ? (const void*)s->d_nodes4q : (const void*)s->d_nodes4;
    const QGrid g = q ? s->qgrid : QGrid{};
    const uint32_t ring = tune_ring(ctx);
#define PYRE_TRACE4(RING, MOT, QN)                                                                                  \
  hipLaunchKernelGGL((k_trace4<ANY, RING, MOT, false, QN>), dim3(grid), dim3(kTraceThreads), 0, st, nodes, s->d_tris, \
                     rays, (uint32_t)n, hit, uv, occ, gst, ctx->trace_fetch, times, s->d_motion, g)
    if (s->has_motion) {
      PYRE_TRACE4(8, true, false);
    } else if (ring == 4) {
      if (q)
        PYRE_TRACE4(4, false, true);
      else
        PYRE_TRACE4(4, false, false);
    } else if (ring == 16) {
      if (q)
        PYRE_TRACE4(16, false, true);
      else
        PYRE_TRACE4(16, false, false);
    } else {
      if (q)
        PYRE_TRACE4(8, false, true);
      else
        PYRE_TRACE4(8, false, false);
    }
#undef PYRE_TRACE4
  } else if (mode == 3) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, ctx->trace_fetch, 1u);
    hipLaunchKernelGGL((k_trace4<ANY, 8, false, true>), dim3(grid), dim3(kTraceThreads), 0, st, s->d_nodes4,
                       s->d_tris, rays, (uint32_t)n, hit, uv, occ, gst, ctx->trace_fetch, nullptr, nullptr, QGrid{});
  } else if (mode == 2) {
    hipLaunchKernelGGL((k_trace4s<ANY>), dim3(grid), dim3(kTraceThreads), 0, st, s->d_nodes4, s->d_tris, rays,
                       (uint32_t)n, hit, uv, occ, gst);
  } else {
    dim3 g((unsigned)((n + kTraceThreads - 1) / kTraceThreads)), block(kTraceThreads);
    const uint32_t need = s->depth + 1;
    if (need <= 16)
      hipLaunchKernelGGL((k_trace<ANY, 16>), g, block, 0, st, s->d_nodes, s->d_tris, rays, (uint32_t)n, hit, uv, occ,
                         nullptr, 0, ctx->trace_flag);
    else if (need <= 32)
      hipLaunchKernelGGL((k_trace<ANY, 32>), g, block, 0, st, s->d_nodes, s->d_tris, rays, (uint32_t)n, hit, uv, occ,
                         nullptr, 0, ctx->trace_flag);
    else
      hipLaunchKernelGGL((k_trace<ANY, 48>), g, block, 0, st, s->d_nodes, s->d_tris, rays, (uint32_t)n, hit, uv, occ,
                         nullptr, 0, ctx->trace_flag);
  }
  PYRE_HIP_CHECK(hipGetLastError());
  PYRE_HIP_CHECK(hipEventRecord(ctx->trace_ev, st));
  if (check && mode == 1) {
    uint32_t h_ovf = 0;
    PYRE_HIP_CHECK(hipMemcpyAsync(&h_ovf, ctx->trace_flag, 4, hipMemcpyDeviceToHost, st));
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
    if (h_ovf) throw Error(PYRE_EOVERFLOW, "traversal stack overflow (tree deeper than recorded depth)");
  }
}

}  // namespace

void trace_scratch_reserve(pyre_ctx* ctx, uint32_t depth) {
  if (!ctx->trace_ev) PYRE_HIP_CHECK(hipEventCreateWithFlags(&ctx->trace_ev, hipEventDisableTiming));
  if (!ctx->trace_fetch) ctx->trace_fetch = dev_alloc<uint32_t>(64);
  if (!ctx->trace_flag) {
    ctx->trace_flag = dev_alloc<uint32_t>(1);
    PYRE_HIP_CHECK(hipMemset(ctx->trace_flag, 0, 4));
  }
  if (ctx->trace_gstack && depth <= ctx->trace_depth) return;
  // every grid the tuning allows (at most 8 blocks per CU) fits the columns
  const size_t threads = (size_t)std::max(1, ctx->cus) * 8 * kTraceThreads;
  PYRE_HIP_CHECK(hipDeviceSynchronize());  // earlier trace calls may still use the old columns
  dev_free(ctx->trace_gstack);
  ctx->trace_gstack = dev_alloc<int2>(threads * spill_entries(depth));
  ctx->trace_depth = depth;
}

void trace_scratch_release(pyre_ctx* ctx) {
  dev_free(ctx->trace_gstack);
  dev_free(ctx->trace_fetch);
  dev_free(ctx->trace_flag);
  ctx->trace_depth = 0;
  if (ctx->trace_ev) (void)hipEventDestroy(ctx->trace_ev);
  ctx->trace_ev = nullptr;
}

// ------------------------------------------------ no-allocation variant
// Traversal scratch of pyre_trace_closest_dev_noalloc: a fixed array in the
// device code, one copy per GPU, held until the process ends. Its first 256
// bytes hold the persistent kernel's work counter (cleared by a fill kernel
// before every launch), the rest the stack spill columns of as many
// 256-thread blocks as fit. On the host `scratch` names no device address:
// only the pointer hipGetSymbolAddress returns for the current device is
// used (noalloc_scratch_init, at context creation).
constexpr size_t kNoallocScratchBytes = 64ull << 20;
constexpr size_t kNoallocCounterBytes = 256;
static __device__ __align__(256) int2 scratch[kNoallocScratchBytes / sizeof(int2)];

void noalloc_scratch_init(pyre_ctx* ctx) {
  // one lookup per device (the context's device is current here)
  static std::mutex mu;
  static std::map<int, void*> per_device;
  std::lock_guard<std::mutex> lock(mu);
  auto it = per_device.find(ctx->device);
  if (it == per_device.end()) {
    void* ptr = nullptr;
    PYRE_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    it = per_device.emplace(ctx->device, ptr).first;
  }
  ctx->noalloc_scratch = it->second;
}

// Persistent blocks the fixed scratch has spill columns for (at most the
// tuning's blocks per CU); 0 when not even one block fits.
unsigned noalloc_blocks(const pyre_ctx* ctx, uint32_t depth) {
  const size_t per_block = (size_t)kTraceThreads * spill_entries(depth) * sizeof(int2);
  const size_t fit = (kNoallocScratchBytes - kNoallocCounterBytes) / per_block;
  return (unsigned)std::min<size_t>(fit, trace_grid(ctx));
}

void trace_closest_noalloc(pyre_scene* s, const float4* rays, size_t n, pyre_hit8* hits, hipStream_t st) {
  require(s->built, "trace before pyre_bvh_build");
  require(n <= 0xFFFFFFFFull, "too many rays in one call");
  pyre_ctx* ctx = s->ctx;
  require(ctx->noalloc_scratch != nullptr, "no-allocation scratch not initialised");
  const unsigned grid = noalloc_blocks(ctx, s->depth);
  require(grid >= 1, "tree depth " + std::to_string(s->depth) + " needs more stack than the " +
                         std::to_string(kNoallocScratchBytes >> 20) + " MiB no-allocation scratch holds");
  if (n == 0) return;
  uint32_t* fetch = static_cast<uint32_t*>(ctx->noalloc_scratch);
  int2* gst = reinterpret_cast<int2*>(static_cast<char*>(ctx->noalloc_scratch) + kNoallocCounterBytes);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, fetch, 1u);
  if (s->has_motion)
    hipLaunchKernelGGL((k_trace4<false, 8, true>), dim3(grid), dim3(kTraceThreads), 0, st, s->d_nodes4, s->d_tris,
                       rays, (uint32_t)n, hits, nullptr, nullptr, gst, fetch, nullptr, s->d_motion, QGrid{});
  else
    hipLaunchKernelGGL((k_trace4<false, 8, false, false, true>), dim3(grid), dim3(kTraceThreads), 0, st,
                       s->d_nodes4q, s->d_tris, rays, (uint32_t)n, hits, nullptr, nullptr, gst, fetch, nullptr,
                       nullptr, s->qgrid);
  PYRE_HIP_CHECK(hipGetLastError());
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
