// GPU binned-SAH BVH build for gfx950.
//
// Split decisions restate the reference's top-down build
// (crates/pyre/src/geometry/bvh.rs:180-346): bin centroids along the longest
// centroid axis, prefix/suffix SAH sweep with cost tc + ic*(nL*AL + nR*AR)/A,
// leaf when n == 1, when the centroid extent is < 1e-9, or when n <= max_leaf
// and the best cost is >= ic*n; partition with `c < split_pos`; fall back to a
// leaf or an equal-count split when the partition is one-sided. Bounds are
// unions (exact min/max), so computing them in parallel gives the same bits
// as the reference's sequential loop, and the tree has the reference's
// topology. Two differences, both documented: partitions are stable (the
// reference swaps to the end, bvh.rs:389-409), which only changes the order
// of primitives inside a leaf and the equal-count fallback split; nodes are
// numbered breadth-first at the top and per subtree below, with explicit
// child links instead of the reference's DFS `left = self + 1`.
//
// Structure (MI355X-first):
//   wide phase   nodes with > kSubtreeMax primitives, one level at a time;
//                each node is cut into kChunk-primitive chunks, one 256-thread
//                workgroup per chunk bins into an LDS histogram and flushes it
//                with a handful of global atomics; partitions are stable
//                block-scan scatters.
//   subtree phase every node with <= kSubtreeMax primitives roots a subtree
//                built by one wave64 workgroup entirely in LDS (ballot-ranked
//                partitions, wave reductions), with no global atomics.
//   flatten      build nodes -> 64-byte two-box traversal nodes; triangles are
//                reordered into leaf order as (v0, e1, e2) float4 triples.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>

#include "internal.hpp"

namespace pyre {
namespace {

constexpr int kMaxBins = 64;
// primitives per wave-built subtree: 256 keeps 16-32 subtree waves per CU
// at 2^20 primitives (1024: 4-8, and a 33 % longer build)
constexpr uint32_t kSubtreeMax = 256;
constexpr uint32_t kChunk = 4096;       // primitives per wide-phase workgroup
constexpr int kWideThreads = 256;
constexpr int kPerThread = kChunk / kWideThreads;
constexpr int kStackMax = 40;

constexpr uint32_t FLAG_LEAF = 1u, FLAG_USED = 2u;
constexpr uint32_t MODE_LEAF = 0u, MODE_SPLIT = 1u, MODE_EQUAL = 2u;

struct BuildParams {
  uint32_t bins, max_leaf;
  float tc, ic;
};

struct Counters {
  uint32_t next_big;      // big nodes in the next level
  uint32_t n_subtrees;    // subtree roots so far
  uint32_t node_cursor;   // next free build-node id (wide phase)
  uint32_t region_cursor; // next free build-node id in the subtree region
  uint32_t total_chunks;
  uint32_t depth;
  uint32_t pad[2];
};

__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }

__device__ __forceinline__ int32_t wave_min_i(int32_t v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int32_t wave_max_i(int32_t v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ uint32_t wave_sum_u(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// 12 keys: bounds min(3) max(3), centroid min(3) max(3).
struct KeyAcc {
  int32_t k[12];
  __device__ void init() {
    for (int i = 0; i < 3; ++i) {
      k[i] = f2key(kInf);
      k[3 + i] = f2key(-kInf);
      k[6 + i] = f2key(kInf);
      k[9 + i] = f2key(-kInf);
    }
  }
  __device__ void add(float4 lo, float4 hi, float4 c) {
    k[0] = min(k[0], f2key(lo.x));
    k[1] = min(k[1], f2key(lo.y));
    k[2] = min(k[2], f2key(lo.z));
    k[3] = max(k[3], f2key(hi.x));
    k[4] = max(k[4], f2key(hi.y));
    k[5] = max(k[5], f2key(hi.z));
    k[6] = min(k[6], f2key(c.x));
    k[7] = min(k[7], f2key(c.y));
    k[8] = min(k[8], f2key(c.z));
    k[9] = max(k[9], f2key(c.x));
    k[10] = max(k[10], f2key(c.y));
    k[11] = max(k[11], f2key(c.z));
  }
  __device__ void wave_reduce() {
    for (int i = 0; i < 3; ++i) {
      k[i] = wave_min_i(k[i]);
      k[3 + i] = wave_max_i(k[3 + i]);
      k[6 + i] = wave_min_i(k[6 + i]);
      k[9 + i] = wave_max_i(k[9 + i]);
    }
  }
};

__device__ __forceinline__ void keys_atomic(int32_t* bkey, int32_t* ckey, const KeyAcc& a) {
  for (int i = 0; i < 3; ++i) {
    atomicMin(&bkey[i], a.k[i]);
    atomicMax(&bkey[3 + i], a.k[3 + i]);
    atomicMin(&ckey[i], a.k[6 + i]);
    atomicMax(&ckey[3 + i], a.k[9 + i]);
  }
}

__device__ __forceinline__ void keys_empty(int32_t* bkey, int32_t* ckey) {
  for (int i = 0; i < 3; ++i) {
    bkey[i] = f2key(kInf);
    bkey[3 + i] = f2key(-kInf);
    ckey[i] = f2key(kInf);
    ckey[3 + i] = f2key(-kInf);
  }
}

// bvh.rs:164-172
__device__ __forceinline__ int max_axis(V3 v) {
  if (v.x >= v.y && v.x >= v.z) return 0;
  if (v.y >= v.z) return 1;
  return 2;
}

// math.rs:69-72
__device__ __forceinline__ float surface_area(V3 lo, V3 hi) {
  V3 d = hi - lo;
  return 2.0f * (d.x * d.y + d.x * d.z + d.y * d.z);
}

// Bin of a centroid coordinate (bvh.rs:218-229).
__device__ __forceinline__ int bin_of(float c, float cmin, float scale, int bins) {
  int32_t b = f2i32((c - cmin) * scale);
  if (b < 0) b = 0;
  if (b >= bins) b = bins - 1;
  return b;
}

struct NodeGeom {
  V3 blo, bhi, clo, chi;
};
__device__ __forceinline__ NodeGeom geom_from_keys(const int32_t* bkey, const int32_t* ckey) {
  NodeGeom g;
  g.blo = v3(key2f(bkey[0]), key2f(bkey[1]), key2f(bkey[2]));
  g.bhi = v3(key2f(bkey[3]), key2f(bkey[4]), key2f(bkey[5]));
  g.clo = v3(key2f(ckey[0]), key2f(ckey[1]), key2f(ckey[2]));
  g.chi = v3(key2f(ckey[3]), key2f(ckey[4]), key2f(ckey[5]));
  return g;
}

// LDS histogram: per bin count + 6 bound keys.
struct BinsLds {
  uint32_t cnt[kMaxBins];
  int32_t key[kMaxBins][6];
};

__device__ void bins_init(BinsLds& b, int bins, int tid, int nthreads) {
  for (int i = tid; i < bins; i += nthreads) {
    b.cnt[i] = 0;
    b.key[i][0] = b.key[i][1] = b.key[i][2] = f2key(kInf);
    b.key[i][3] = b.key[i][4] = b.key[i][5] = f2key(-kInf);
  }
}

__device__ __forceinline__ void bins_add(BinsLds& b, int bin, float4 lo, float4 hi) {
  atomicAdd(&b.cnt[bin], 1u);
  atomicMin(&b.key[bin][0], f2key(lo.x));
  atomicMin(&b.key[bin][1], f2key(lo.y));
  atomicMin(&b.key[bin][2], f2key(lo.z));
  atomicMax(&b.key[bin][3], f2key(hi.x));
  atomicMax(&b.key[bin][4], f2key(hi.y));
  atomicMax(&b.key[bin][5], f2key(hi.z));
}

// SAH sweep over LDS bins by one wave (bvh.rs:231-270): lane i < bins-1
// evaluates split i with its prefix/suffix unions (exact, so identical to the
// reference's sequential sweeps); the wave argmin keeps the first index of the
// minimum like the reference's strict `cost < best_cost`. Returns best_split
// and best_cost (inf when no split has both sides non-empty).
__device__ void wave_sah(const BinsLds& b, int bins, float parent_area, const BuildParams& p, float* best_cost_out,
                         int* best_split_out) {
  int lane = (int)__lane_id();
  float cost = kInf;
  if (lane < bins - 1) {
    uint32_t lc = 0, rc = 0;
    V3 llo = splat(kInf), lhi = splat(-kInf), rlo = splat(kInf), rhi = splat(-kInf);
    for (int i = 0; i <= lane; ++i) {
      lc += b.cnt[i];
      llo = vmin(llo, v3(key2f(b.key[i][0]), key2f(b.key[i][1]), key2f(b.key[i][2])));
      lhi = vmax(lhi, v3(key2f(b.key[i][3]), key2f(b.key[i][4]), key2f(b.key[i][5])));
    }
    for (int i = lane + 1; i < bins; ++i) {
      rc += b.cnt[i];
      rlo = vmin(rlo, v3(key2f(b.key[i][0]), key2f(b.key[i][1]), key2f(b.key[i][2])));
      rhi = vmax(rhi, v3(key2f(b.key[i][3]), key2f(b.key[i][4]), key2f(b.key[i][5])));
    }
    if (lc != 0 && rc != 0)
      cost = p.tc + p.ic * ((float)lc * surface_area(llo, lhi) + (float)rc * surface_area(rlo, rhi)) / parent_area;
  }
  // argmin with lowest index on ties (costs are never NaN: both sides non-empty).
  float bc = cost;
  int bi = lane;
  for (int o = 32; o > 0; o >>= 1) {
    float oc = __shfl_xor(bc, o);
    int oi = __shfl_xor(bi, o);
    if (oc < bc || (oc == bc && oi < bi)) {
      bc = oc;
      bi = oi;
    }
  }
  *best_cost_out = bc;
  *best_split_out = bi;
}

// Leaf / split decision of bvh.rs:272-315 given the sweep result.
// Returns MODE_LEAF or MODE_SPLIT (with split_pos).
__device__ __forceinline__ uint32_t decide(float best_cost, int best_split, uint32_t n, float cmin, float extent,
                                           const BuildParams& p, float* split_pos) {
  float leaf_cost = p.ic * (float)n;
  if (!isfinite(best_cost)) {
    *split_pos = cmin + 0.5f * extent;
    return MODE_SPLIT;
  }
  if (n <= p.max_leaf && best_cost >= leaf_cost) return MODE_LEAF;
  *split_pos = cmin + (float)(best_split + 1) * extent / (float)p.bins;
  return MODE_SPLIT;
}

// ------------------------------------------------------------------ scans
// Exclusive scan of u32 (3 kernels: per-block scan, scan of block sums,
// add-back). Blocks of 256 threads x 4 items.
constexpr int kScanThreads = 256, kScanItems = 4, kScanTile = kScanThreads * kScanItems;

__device__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds_waves, uint32_t* total) {
  // lds_waves: at least (blockDim/64 + 1) entries
  int lane = (int)__lane_id();
  int wid = (int)(threadIdx.x >> 6);
  int nw = (int)(blockDim.x >> 6);
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds_waves[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int w = 0; w < nw; ++w) {
      uint32_t t = lds_waves[w];
      lds_waves[w] = acc;
      acc += t;
    }
    lds_waves[nw] = acc;
  }
  __syncthreads();
  uint32_t r = x - v + lds_waves[wid];
  if (total) *total = lds_waves[nw];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_blocks(const uint32_t* in, uint32_t* out, uint32_t* sums,
                                                              size_t n) {
  __shared__ uint32_t w[kScanThreads / 64 + 1];
  size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = (base + i < n) ? in[base + i] : 0u;
    s += v[i];
  }
  uint32_t tot;
  uint32_t off = block_exclusive_scan(s, w, &tot);
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = off;
    off += v[i];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_sums(uint32_t* sums, uint32_t nb, uint32_t* total) {
  __shared__ uint32_t w[1024 / 64 + 1];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t b0 = 0; b0 < nb; b0 += 1024) {
    uint32_t i = b0 + threadIdx.x;
    uint32_t v = i < nb ? sums[i] : 0u;
    uint32_t tot;
    uint32_t off = block_exclusive_scan(v, w, &tot);
    uint32_t c = carry;
    if (i < nb) sums[i] = off + c;
    __syncthreads();
    if (threadIdx.x == 0) carry = c + tot;
    __syncthreads();
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(uint32_t* out, const uint32_t* sums, size_t n) {
  size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
  uint32_t a = sums[blockIdx.x];
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) out[base + i] += a;
}

struct Scratch {
  std::vector<void*> bufs;
  template <class T>
  T* get(size_t n) {
    T* p = dev_alloc<T>(n);
    bufs.push_back(p);
    return p;
  }
  ~Scratch() {
    for (void* p : bufs) (void)hipFree(p);
  }
};

// out = exclusive scan of in; *total_dev (device) = sum.
void exclusive_scan(Scratch& S, const uint32_t* in, uint32_t* out, size_t n, uint32_t* total_dev, hipStream_t st) {
  size_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb == 0) nb = 1;
  uint32_t* sums = S.get<uint32_t>(nb);
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, out, sums, n);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, st, sums, (uint32_t)nb, total_dev);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(kScanThreads), 0, st, out, sums, n);
}

// ------------------------------------------------------------------ prep
// Per triangle bounds and centroid (mesh.rs:32-40); root bounds reduced per
// thread over a grid-stride range, then per wave and per block (LDS), so
// the root's keys see one set of atomics per block.
// Mesh of scene-global triangle gid: the largest m with tri_offset[m] <= gid.
__device__ __forceinline__ uint32_t mesh_of_tri(const uint32_t* tri_offset, uint32_t n_meshes, uint32_t gid) {
  uint32_t lo = 0, hi = n_meshes;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (tri_offset[mid] <= gid)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// motion (scenes with moving meshes, else null): 2 float4 per mesh; a moving
// mesh's triangle gets the box swept over the shutter (its box at offset_t0
// and at offset_t1) and the centroid at the mid-shutter offset.
__global__ __launch_bounds__(256) void k_prep(const float* __restrict__ verts, uint32_t n, float4* plo, float4* phi,
                                              float4* pc, uint32_t* ref, BNode* root, const float4* motion,
                                              const uint32_t* tri_offset, uint32_t n_meshes) {
  __shared__ int32_t part[4][12];
  KeyAcc acc;
  acc.init();
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float* v = verts + 9 * (size_t)i;
    V3 a = v3(v[0], v[1], v[2]), b = v3(v[3], v[4], v[5]), c = v3(v[6], v[7], v[8]);
    V3 lo = vmin(vmin(a, b), c);
    V3 hi = vmax(vmax(a, b), c);
    V3 ce = (a + b + c) / 3.0f;
    if (motion) {
      const uint32_t m = mesh_of_tri(tri_offset, n_meshes, i);
      const float4 m0 = motion[2 * m], m1 = motion[2 * m + 1];
      const V3 o0 = v3(m0.x, m0.y, m0.z), o1 = v3(m1.x, m1.y, m1.z);
      if (o0.x != 0.0f || o0.y != 0.0f || o0.z != 0.0f || o1.x != 0.0f || o1.y != 0.0f || o1.z != 0.0f) {
        const V3 l0 = lo + o0, l1 = lo + o1, h0 = hi + o0, h1 = hi + o1;
        lo = vmin(l0, l1);
        hi = vmax(h0, h1);
        ce = ce + 0.5f * (o0 + o1);
      }
    }
    float4 l4 = make_float4(lo.x, lo.y, lo.z, 0.0f), h4 = make_float4(hi.x, hi.y, hi.z, 0.0f),
           c4 = make_float4(ce.x, ce.y, ce.z, 0.0f);
    plo[i] = l4;
    phi[i] = h4;
    pc[i] = c4;
    ref[i] = i;
    acc.add(l4, h4, c4);
  }
  acc.wave_reduce();
  const int wave = threadIdx.x >> 6;
  if (__lane_id() == 0)
    for (int k = 0; k < 12; ++k) part[wave][k] = acc.k[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
      for (int k = 0; k < 12; ++k)
        acc.k[k] = (k % 6) < 3 ? min(acc.k[k], part[w][k]) : max(acc.k[k], part[w][k]);
    keys_atomic(root->bkey, root->ckey, acc);
  }
}

// Root record, counters and the first work list (level list or, for small
// scenes, the single subtree root).
__global__ void k_init_root(BNode* nodes, uint32_t n, Counters* ctr, uint32_t* lvl, uint32_t* sub_roots,
                            uint32_t* sub_base, uint32_t small) {
  if (small) {
    sub_roots[0] = 0;
    sub_base[0] = 1;
    ctr->n_subtrees = 1;
    ctr->region_cursor = 1 + 2 * n;
  } else {
    lvl[0] = 0;
    ctr->n_subtrees = 0;
    ctr->region_cursor = 0;
  }
  ctr->node_cursor = 1;
  ctr->next_big = 0;
  ctr->total_chunks = 0;
  ctr->depth = 0;
  BNode& r = nodes[0];
  keys_empty(r.bkey, r.ckey);
  r.start = 0;
  r.count = n;
  r.left = r.right = -1;
  r.axis = 0;
  r.flags = FLAG_USED;
  r.split_pos = 0.0f;
  r.mode = 0;
}

// ---------------------------------------------------------- wide phase
// Chunks of the level's nodes: chunk c covers primitives
// [chunk_first[c], +kChunk) of node lvl[chunk_node[c]].
__global__ void k_chunk_counts(const BNode* nodes, const uint32_t* lvl, uint32_t L, uint32_t* nch) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < L)
    nch[j] = (nodes[lvl[j]].count + kChunk - 1) / kChunk;
  else if (j == L)
    nch[j] = 0;
}

__global__ void k_chunk_fill(const BNode* nodes, const uint32_t* lvl, uint32_t L, const uint32_t* nch_off,
                             uint32_t* chunk_node, uint32_t* chunk_first) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= L) return;
  uint32_t cnt = nodes[lvl[j]].count;
  uint32_t nc = (cnt + kChunk - 1) / kChunk;
  for (uint32_t k = 0; k < nc; ++k) {
    chunk_node[nch_off[j] + k] = j;
    chunk_first[nch_off[j] + k] = k * kChunk;
  }
}

__device__ __forceinline__ bool node_axis(const NodeGeom& g, int* axis, float* extent, float* cmin) {
  V3 cd = g.chi - g.clo;
  int a = max_axis(cd);
  *axis = a;
  *extent = axis_value(cd, a);
  *cmin = axis_value(g.clo, a);
  return !(*extent < 1e-9f);
}

__global__ __launch_bounds__(kWideThreads) void k_wide_bin(const BNode* nodes, const uint32_t* lvl,
                                                           const uint32_t* chunk_node, const uint32_t* chunk_first,
                                                           const uint32_t* total_chunks, const uint32_t* ref,
                                                           const float4* plo, const float4* phi, const float4* pc,
                                                           int32_t* gbins, BuildParams p) {
  if (blockIdx.x >= *total_chunks) return;
  __shared__ BinsLds B;
  uint32_t j = chunk_node[blockIdx.x];
  const BNode& nd = nodes[lvl[j]];
  NodeGeom g = geom_from_keys(nd.bkey, nd.ckey);
  int axis;
  float extent, cmin;
  if (!node_axis(g, &axis, &extent, &cmin)) return;  // leaf: coincident centroids
  int bins = (int)p.bins;
  float scale = (float)bins / extent;
  bins_init(B, bins, threadIdx.x, blockDim.x);
  __syncthreads();
  uint32_t first = chunk_first[blockIdx.x];
  uint32_t end = min(first + kChunk, nd.count);
  for (uint32_t k = first + threadIdx.x; k < end; k += blockDim.x) {
    uint32_t pid = ref[nd.start + k];
    float4 c = pc[pid];
    float cv = axis == 0 ? c.x : (axis == 1 ? c.y : c.z);
    bins_add(B, bin_of(cv, cmin, scale, bins), plo[pid], phi[pid]);
  }
  __syncthreads();
  int32_t* gb = gbins + (size_t)j * kMaxBins * 7;
  for (int i = threadIdx.x; i < bins; i += blockDim.x) {
    if (B.cnt[i] == 0) continue;
    atomicAdd((uint32_t*)&gb[i * 7], B.cnt[i]);
    for (int q = 0; q < 3; ++q) atomicMin(&gb[i * 7 + 1 + q], B.key[i][q]);
    for (int q = 3; q < 6; ++q) atomicMax(&gb[i * 7 + 1 + q], B.key[i][q]);
  }
}

__global__ void k_gbins_init(int32_t* gbins, uint32_t L) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L * kMaxBins) return;
  int32_t* b = gbins + (size_t)i * 7;
  b[0] = 0;
  b[1] = b[2] = b[3] = f2key(kInf);
  b[4] = b[5] = b[6] = f2key(-kInf);
}

// One wave per level node: SAH decision from the global histogram.
__global__ __launch_bounds__(64) void k_wide_decide(BNode* nodes, const uint32_t* lvl, const int32_t* gbins,
                                                    BuildParams p) {
  __shared__ BinsLds B;
  uint32_t j = blockIdx.x;
  BNode& nd = nodes[lvl[j]];
  NodeGeom g = geom_from_keys(nd.bkey, nd.ckey);
  int axis;
  float extent, cmin;
  bool splittable = node_axis(g, &axis, &extent, &cmin);
  if (!splittable) {  // bvh.rs:205-208
    if (threadIdx.x == 0) nd.mode = MODE_LEAF;
    return;
  }
  const int32_t* gb = gbins + (size_t)j * kMaxBins * 7;
  for (int i = threadIdx.x; i < (int)p.bins; i += 64) {
    B.cnt[i] = (uint32_t)gb[i * 7];
    for (int q = 0; q < 6; ++q) B.key[i][q] = gb[i * 7 + 1 + q];
  }
  __syncthreads();
  float parent_area = rmax(surface_area(g.blo, g.bhi), 1e-9f);
  float bc;
  int bs;
  wave_sah(B, (int)p.bins, parent_area, p, &bc, &bs);
  if (threadIdx.x == 0) {
    float sp = 0.0f;
    nd.mode = decide(bc, bs, nd.count, cmin, extent, p, &sp);
    nd.split_pos = sp;
    nd.axis = (uint32_t)axis;
  }
}

__global__ __launch_bounds__(kWideThreads) void k_wide_count(const BNode* nodes, const uint32_t* lvl,
                                                             const uint32_t* chunk_node, const uint32_t* chunk_first,
                                                             const uint32_t* total_chunks, const uint32_t* ref,
                                                             const float4* pc, uint32_t* chunk_left) {
  if (blockIdx.x >= *total_chunks) return;
  __shared__ uint32_t w[kWideThreads / 64 + 1];
  uint32_t j = chunk_node[blockIdx.x];
  const BNode& nd = nodes[lvl[j]];
  if (nd.mode != MODE_SPLIT) {
    if (threadIdx.x == 0) chunk_left[blockIdx.x] = 0;
    return;
  }
  uint32_t first = chunk_first[blockIdx.x];
  uint32_t end = min(first + kChunk, nd.count);
  uint32_t cnt = 0;
  for (uint32_t k = first + threadIdx.x; k < end; k += blockDim.x) {
    float4 c = pc[ref[nd.start + k]];
    float cv = nd.axis == 0 ? c.x : (nd.axis == 1 ? c.y : c.z);
    cnt += cv < nd.split_pos ? 1u : 0u;
  }
  uint32_t tot;
  block_exclusive_scan(cnt, w, &tot);
  if (threadIdx.x == 0) chunk_left[blockIdx.x] = tot;
}

// Single workgroup: per node left counts and chunk offsets, fallback rules,
// child allocation (deterministic: by scan over the level order), and the
// next level's big-node list / subtree-root list.
__global__ __launch_bounds__(1024) void k_wide_finalize(BNode* nodes, const uint32_t* lvl, uint32_t L,
                                                        const uint32_t* nch_off, const uint32_t* chunk_left,
                                                        uint32_t* chunk_loff, uint32_t* next_lvl, uint32_t* sub_roots,
                                                        uint32_t* sub_base, Counters* ctr, BuildParams p,
                                                        uint32_t level) {
  __shared__ uint32_t w[1024 / 64 + 1];
  __shared__ uint32_t carry_split, carry_big, carry_small, carry_region;
  if (threadIdx.x == 0) {
    carry_split = 0;
    carry_big = 0;
    carry_small = 0;
    carry_region = 0;
  }
  __syncthreads();
  const uint32_t node_base = ctr->node_cursor;
  const uint32_t sub_base0 = ctr->n_subtrees;
  const uint32_t region0 = ctr->region_cursor;
  for (uint32_t j0 = 0; j0 < L; j0 += 1024) {
    uint32_t j = j0 + threadIdx.x;
    bool valid = j < L;
    uint32_t is_split = 0, nbig = 0, nsmall = 0, region = 0;
    uint32_t nleft = 0, mode = MODE_LEAF, id = 0, count = 0;
    if (valid) {
      id = lvl[j];
      BNode& nd = nodes[id];
      count = nd.count;
      mode = nd.mode;
      if (mode == MODE_SPLIT) {
        uint32_t c0 = nch_off[j], c1 = nch_off[j + 1];
        uint32_t acc = 0;
        for (uint32_t c = c0; c < c1; ++c) {
          chunk_loff[c] = acc;
          acc += chunk_left[c];
        }
        nleft = acc;
        if (nleft == 0 || nleft == count) {  // bvh.rs:280-307, 317-333
          if (count <= p.max_leaf)
            mode = MODE_LEAF;
          else {
            mode = MODE_EQUAL;
            nleft = count / 2;
            for (uint32_t c = c0; c < c1; ++c) chunk_loff[c] = min((c - c0) * kChunk, nleft);
          }
        }
      }
      if (mode == MODE_LEAF) {
        nd.flags = FLAG_USED | FLAG_LEAF;
        nd.count = min(count, 65535u);  // bvh.rs:207 (u16 leaf count)
        nd.left = nd.right = -1;
      } else {
        is_split = 1;
        nd.mode = mode;
        uint32_t nl = nleft, nr = count - nleft;
        nbig = (nl > kSubtreeMax ? 1u : 0u) + (nr > kSubtreeMax ? 1u : 0u);
        nsmall = 2u - nbig;
        region = (nl <= kSubtreeMax ? 2u * nl : 0u) + (nr <= kSubtreeMax ? 2u * nr : 0u);
      }
    }
    uint32_t t_split, t_big, t_small, t_region;
    uint32_t o_split = block_exclusive_scan(is_split, w, &t_split);
    uint32_t o_big = block_exclusive_scan(nbig, w, &t_big);
    uint32_t o_small = block_exclusive_scan(nsmall, w, &t_small);
    uint32_t o_region = block_exclusive_scan(region, w, &t_region);
    if (valid && is_split) {
      BNode& nd = nodes[id];
      uint32_t lid = node_base + 2u * (carry_split + o_split);
      uint32_t rid = lid + 1u;
      nd.left = (int32_t)lid;
      nd.right = (int32_t)rid;
      nd.flags = FLAG_USED;
      nd.split_pos = nd.split_pos;
      uint32_t nl = nleft, nr = count - nleft;
      BNode* ch[2] = {&nodes[lid], &nodes[rid]};
      uint32_t cs[2] = {nd.start, nd.start + nl};
      uint32_t cn[2] = {nl, nr};
      uint32_t bi = carry_big + o_big, si = carry_small + o_small, ri = carry_region + o_region;
      for (int k = 0; k < 2; ++k) {
        BNode& c = *ch[k];
        keys_empty(c.bkey, c.ckey);
        c.start = cs[k];
        c.count = cn[k];
        c.left = c.right = -1;
        c.axis = 0;
        c.flags = FLAG_USED;
        c.split_pos = 0.0f;
        c.mode = 0;
        if (cn[k] > kSubtreeMax) {
          next_lvl[bi++] = (k == 0) ? lid : rid;
        } else {
          sub_roots[sub_base0 + si] = (k == 0) ? lid : rid;
          sub_base[sub_base0 + si] = region0 + ri;
          // subtree depth base: stored in the node's mode field until built
          c.mode = level + 1;
          si++;
          ri += 2u * cn[k];
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      carry_split += t_split;
      carry_big += t_big;
      carry_small += t_small;
      carry_region += t_region;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    ctr->node_cursor = node_base + 2u * carry_split;
    ctr->next_big = carry_big;
    ctr->n_subtrees = sub_base0 + carry_small;
    ctr->region_cursor = region0 + carry_region;
    if (carry_split > 0) ctr->depth = max(ctr->depth, level + 1);
  }
}

// Stable partition of each split node's chunk into tmp, plus children's
// bounds and centroid bounds.
__global__ __launch_bounds__(kWideThreads) void k_wide_scatter(BNode* nodes, const uint32_t* lvl,
                                                               const uint32_t* chunk_node,
                                                               const uint32_t* chunk_first,
                                                               const uint32_t* total_chunks,
                                                               const uint32_t* chunk_loff, const uint32_t* ref,
                                                               uint32_t* tmp, const float4* plo, const float4* phi,
                                                               const float4* pc) {
  if (blockIdx.x >= *total_chunks) return;
  __shared__ uint32_t w[kWideThreads / 64 + 1];
  __shared__ int32_t ck[2][12];
  uint32_t j = chunk_node[blockIdx.x];
  const BNode& nd = nodes[lvl[j]];
  if (nd.flags & FLAG_LEAF) return;
  if (threadIdx.x < 24) {
    int k = threadIdx.x % 12;
    bool is_min = (k < 3) || (k >= 6 && k < 9);
    ck[threadIdx.x / 12][k] = is_min ? f2key(kInf) : f2key(-kInf);
  }
  uint32_t first = chunk_first[blockIdx.x];
  uint32_t end = min(first + kChunk, nd.count);
  uint32_t nleft = (uint32_t)(nodes[nd.right].start - nd.start);
  // each thread owns kPerThread consecutive primitives of the chunk
  uint32_t k0 = first + threadIdx.x * kPerThread;
  uint32_t flags = 0, nl = 0, nvalid = 0;
  for (int q = 0; q < kPerThread; ++q) {
    uint32_t k = k0 + q;
    if (k >= end) break;
    nvalid++;
    bool f;
    if (nd.mode == MODE_EQUAL) {
      f = k < nleft;
    } else {
      float4 c = pc[ref[nd.start + k]];
      float cv = nd.axis == 0 ? c.x : (nd.axis == 1 ? c.y : c.z);
      f = cv < nd.split_pos;
    }
    if (f) {
      flags |= 1u << q;
      nl++;
    }
  }
  uint32_t loff = block_exclusive_scan(nl, w, nullptr);
  uint32_t chunk_l = chunk_loff[blockIdx.x];
  uint32_t chunk_r = first - chunk_l;  // right elements of this node before the chunk
  uint32_t roff = (threadIdx.x * kPerThread) - loff;
  KeyAcc acc[2];
  acc[0].init();
  acc[1].init();
  uint32_t li = 0, ri = 0;
  for (uint32_t q = 0; q < nvalid; ++q) {
    uint32_t pid = ref[nd.start + k0 + q];
    float4 lo = plo[pid], hi = phi[pid], c = pc[pid];
    if (flags & (1u << q)) {
      tmp[nd.start + chunk_l + loff + li] = pid;
      li++;
      acc[0].add(lo, hi, c);
    } else {
      tmp[nd.start + nleft + chunk_r + roff + ri] = pid;
      ri++;
      acc[1].add(lo, hi, c);
    }
  }
  acc[0].wave_reduce();
  acc[1].wave_reduce();
  __syncthreads();
  if (__lane_id() == 0) {
    for (int s = 0; s < 2; ++s)
      for (int i = 0; i < 12; ++i) {
        bool is_min = (i < 3) || (i >= 6 && i < 9);
        if (is_min)
          atomicMin(&ck[s][i], acc[s].k[i]);
        else
          atomicMax(&ck[s][i], acc[s].k[i]);
      }
  }
  __syncthreads();
  if (threadIdx.x < 24) {
    int s = threadIdx.x / 12, i = threadIdx.x % 12;
    BNode& c = nodes[s == 0 ? nd.left : nd.right];
    int32_t* dst = i < 6 ? &c.bkey[i] : &c.ckey[i - 6];
    bool is_min = (i < 3) || (i >= 6 && i < 9);
    if (is_min)
      atomicMin(dst, ck[s][i]);
    else
      atomicMax(dst, ck[s][i]);
  }
}

__global__ __launch_bounds__(kWideThreads) void k_wide_copyback(const BNode* nodes, const uint32_t* lvl,
                                                                const uint32_t* chunk_node,
                                                                const uint32_t* chunk_first,
                                                                const uint32_t* total_chunks, uint32_t* ref,
                                                                const uint32_t* tmp) {
  if (blockIdx.x >= *total_chunks) return;
  const BNode& nd = nodes[lvl[chunk_node[blockIdx.x]]];
  if (nd.flags & FLAG_LEAF) return;
  uint32_t first = chunk_first[blockIdx.x];
  uint32_t end = min(first + kChunk, nd.count);
  for (uint32_t k = first + threadIdx.x; k < end; k += blockDim.x) ref[nd.start + k] = tmp[nd.start + k];
}

// -------------------------------------------------------- subtree phase
struct StackEntry {
  uint32_t a, m, id, depth;
  int32_t bkey[6];
  int32_t ckey[6];
};

// One wave64 workgroup builds the whole subtree under a node with at most
// kSubtreeMax primitives, keeping its primitive references in LDS.
__global__ __launch_bounds__(64) void k_subtree(BNode* nodes, const uint32_t* sub_roots, const uint32_t* sub_base,
                                                uint32_t n_sub, uint32_t* ref, const float4* plo, const float4* phi,
                                                const float4* pc, BuildParams p, Counters* ctr) {
  uint32_t r = blockIdx.x;
  if (r >= n_sub) return;
  __shared__ uint32_t sref[2][kSubtreeMax];
  __shared__ BinsLds B;
  __shared__ StackEntry stack[kStackMax];
  const int lane = (int)threadIdx.x;
  const uint32_t root = sub_roots[r];
  const uint32_t gstart = nodes[root].start;
  const uint32_t total = nodes[root].count;
  const uint32_t root_depth = nodes[root].mode;  // set by the wide phase (0 for the tree root)
  uint32_t cursor = sub_base[r];
  for (uint32_t k = lane; k < total; k += 64) sref[0][k] = ref[gstart + k];
  int sp = 0;
  if (lane == 0) {
    StackEntry& e = stack[0];
    e.a = 0;
    e.m = total;
    e.id = root;
    e.depth = root_depth;
    for (int i = 0; i < 6; ++i) {
      e.bkey[i] = nodes[root].bkey[i];
      e.ckey[i] = nodes[root].ckey[i];
    }
  }
  sp = 1;
  uint32_t max_depth = root_depth;
  __syncthreads();
  while (sp > 0) {
    sp--;
    StackEntry e = stack[sp];
    __syncthreads();
    max_depth = max(max_depth, e.depth);
    NodeGeom g = geom_from_keys(e.bkey, e.ckey);
    BNode& nd = nodes[e.id];
    bool leaf = false;
    int axis = 0;
    float extent = 0.0f, cmin = 0.0f, split_pos = 0.0f;
    uint32_t mode = MODE_SPLIT, nleft = 0;
    if (e.m == 1) {
      leaf = true;  // bvh.rs:198-200
    } else if (!node_axis(g, &axis, &extent, &cmin)) {
      leaf = true;  // bvh.rs:205-208
    } else {
      int bins = (int)p.bins;
      float scale = (float)bins / extent;
      bins_init(B, bins, lane, 64);
      __syncthreads();
      for (uint32_t k = lane; k < e.m; k += 64) {
        uint32_t pid = sref[0][e.a + k];
        float4 c = pc[pid];
        float cv = axis == 0 ? c.x : (axis == 1 ? c.y : c.z);
        bins_add(B, bin_of(cv, cmin, scale, bins), plo[pid], phi[pid]);
      }
      __syncthreads();
      float parent_area = rmax(surface_area(g.blo, g.bhi), 1e-9f);
      float bc;
      int bs;
      wave_sah(B, bins, parent_area, p, &bc, &bs);
      mode = decide(bc, bs, e.m, cmin, extent, p, &split_pos);
      if (mode == MODE_LEAF) {
        leaf = true;
      } else {
        uint32_t cnt = 0;
        for (uint32_t k = lane; k < e.m; k += 64) {
          float4 c = pc[sref[0][e.a + k]];
          float cv = axis == 0 ? c.x : (axis == 1 ? c.y : c.z);
          cnt += cv < split_pos ? 1u : 0u;
        }
        nleft = wave_sum_u(cnt);
        if (nleft == 0 || nleft == e.m) {
          if (e.m <= p.max_leaf)
            leaf = true;
          else {
            mode = MODE_EQUAL;
            nleft = e.m / 2;
          }
        }
      }
    }
    if (leaf) {
      if (lane == 0) {
        for (int i = 0; i < 6; ++i) {
          nd.bkey[i] = e.bkey[i];
          nd.ckey[i] = e.ckey[i];
        }
        nd.start = gstart + e.a;
        nd.count = min(e.m, 65535u);
        nd.left = nd.right = -1;
        nd.axis = 0;
        nd.flags = FLAG_USED | FLAG_LEAF;
        nd.split_pos = 0.0f;
        nd.mode = MODE_LEAF;
      }
      __syncthreads();
      continue;
    }
    // stable partition sref[0] -> sref[1]
    KeyAcc acc[2];
    acc[0].init();
    acc[1].init();
    uint32_t lbase = 0, rbase = 0;
    for (uint32_t b0 = 0; b0 < e.m; b0 += 64) {
      uint32_t k = b0 + lane;
      bool valid = k < e.m;
      uint32_t pid = valid ? sref[0][e.a + k] : 0u;
      bool f = false;
      if (valid) {
        if (mode == MODE_EQUAL) {
          f = k < nleft;
        } else {
          float4 c = pc[pid];
          float cv = axis == 0 ? c.x : (axis == 1 ? c.y : c.z);
          f = cv < split_pos;
        }
      }
      uint64_t fm = __ballot(f);
      uint64_t vm = __ballot(valid);
      uint64_t rm = vm & ~fm;
      if (valid) {
        float4 lo = plo[pid], hi = phi[pid], c = pc[pid];
        if (f) {
          sref[1][e.a + lbase + (uint32_t)__popcll(fm & lanemask_lt())] = pid;
          acc[0].add(lo, hi, c);
        } else {
          sref[1][e.a + nleft + rbase + (uint32_t)__popcll(rm & lanemask_lt())] = pid;
          acc[1].add(lo, hi, c);
        }
      }
      lbase += (uint32_t)__popcll(fm);
      rbase += (uint32_t)__popcll(rm);
    }
    acc[0].wave_reduce();
    acc[1].wave_reduce();
    __syncthreads();
    for (uint32_t k = lane; k < e.m; k += 64) sref[0][e.a + k] = sref[1][e.a + k];
    uint32_t lid = cursor, rid = cursor + 1;
    cursor += 2;
    if (lane == 0) {
      for (int i = 0; i < 6; ++i) {
        nd.bkey[i] = e.bkey[i];
        nd.ckey[i] = e.ckey[i];
      }
      nd.start = gstart + e.a;
      nd.count = e.m;
      nd.left = (int32_t)lid;
      nd.right = (int32_t)rid;
      nd.axis = (uint32_t)axis;
      nd.flags = FLAG_USED;
      nd.split_pos = split_pos;
      nd.mode = mode;
      // children: larger first on the stack so the smaller is built first
      // (bounds the stack at log2(kSubtreeMax) + 1 entries)
      StackEntry c[2];
      c[0].a = e.a;
      c[0].m = nleft;
      c[0].id = lid;
      c[1].a = e.a + nleft;
      c[1].m = e.m - nleft;
      c[1].id = rid;
      for (int s = 0; s < 2; ++s) {
        c[s].depth = e.depth + 1;
        for (int i = 0; i < 3; ++i) {
          c[s].bkey[i] = acc[s].k[i];
          c[s].bkey[3 + i] = acc[s].k[3 + i];
          c[s].ckey[i] = acc[s].k[6 + i];
          c[s].ckey[3 + i] = acc[s].k[9 + i];
        }
      }
      int small = c[0].m <= c[1].m ? 0 : 1;
      stack[sp] = c[1 - small];
      stack[sp + 1] = c[small];
    }
    sp += 2;
    __syncthreads();
  }
  for (uint32_t k = lane; k < total; k += 64) ref[gstart + k] = sref[0][k];
  if (lane == 0) atomicMax(&ctr->depth, max_depth);
}

// ------------------------------------------------------------- flatten
__global__ void k_shift(uint32_t* a, uint32_t n, uint32_t by) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] += by;
}

__global__ void k_interior_flags(const BNode* nodes, uint32_t n, uint32_t* flag) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = ((nodes[i].flags & FLAG_USED) && !(nodes[i].flags & FLAG_LEAF)) ? 1u : 0u;
}

__device__ __forceinline__ void child_box(const BNode& c, float* lo, float* hi) {
  for (int i = 0; i < 3; ++i) {
    lo[i] = key2f(c.bkey[i]);
    hi[i] = key2f(c.bkey[3 + i]);
  }
}

__global__ void k_flatten(const BNode* nodes, uint32_t n, const uint32_t* flag, const uint32_t* map, Node2* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const BNode& nd = nodes[i];
  const BNode& c0 = nodes[nd.left];
  const BNode& c1 = nodes[nd.right];
  float l0[3], h0[3], l1[3], h1[3];
  child_box(c0, l0, h0);
  child_box(c1, l1, h1);
  Node2 o;
  o.x = make_float4(l0[0], h0[0], l1[0], h1[0]);
  o.y = make_float4(l0[1], h0[1], l1[1], h1[1]);
  o.z = make_float4(l0[2], h0[2], l1[2], h1[2]);
  bool leaf0 = (c0.flags & FLAG_LEAF) != 0, leaf1 = (c1.flags & FLAG_LEAF) != 0;
  o.meta = make_int4(leaf0 ? (int)c0.start : (int)map[nd.left], leaf1 ? (int)c1.start : (int)map[nd.right],
                     leaf0 ? (int)c0.count : 0, leaf1 ? (int)c1.count : 0);
  out[map[i]] = o;
}

// Root that is a leaf: one node with the leaf in child 0 and no child 1.
__global__ void k_flatten_leaf_root(const BNode* nodes, Node2* out) {
  const BNode& r = nodes[0];
  float l[3], h[3];
  child_box(r, l, h);
  Node2 o;
  o.x = make_float4(l[0], h[0], 0.0f, 0.0f);
  o.y = make_float4(l[1], h[1], 0.0f, 0.0f);
  o.z = make_float4(l[2], h[2], 0.0f, 0.0f);
  o.meta = make_int4((int)r.start, -1, (int)r.count, 0);
  out[0] = o;
}

// moving (per mesh, or null for static scenes): e1.w of a moving mesh's
// triangle holds the mesh index + 1 (its motion slot for the leaf tests).
__global__ void k_reorder_tris(const float* verts, const uint32_t* ref, uint32_t n, float4* tris,
                               const uint32_t* moving, const uint32_t* tri_offset, uint32_t n_meshes) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  uint32_t gid = ref[k];
  const float* v = verts + 9 * (size_t)gid;
  V3 v0 = v3(v[0], v[1], v[2]), v1 = v3(v[3], v[4], v[5]), v2 = v3(v[6], v[7], v[8]);
  V3 e1 = v1 - v0, e2 = v2 - v0;  // mesh.rs:154-155
  uint32_t slot = 0;
  if (moving) {
    const uint32_t m = mesh_of_tri(tri_offset, n_meshes, gid);
    slot = moving[m] ? m + 1 : 0u;
  }
  tris[3 * (size_t)k] = make_float4(v0.x, v0.y, v0.z, __builtin_bit_cast(float, gid));
  tris[3 * (size_t)k + 1] = make_float4(e1.x, e1.y, e1.z, __builtin_bit_cast(float, slot));
  tris[3 * (size_t)k + 2] = make_float4(e2.x, e2.y, e2.z, 0.0f);
}

// e2.w of the last slot of every leaf holds kLeafEnd (the traversal walks a
// leaf slot by slot until that mark).
__global__ void k_mark_leaf_ends(const BNode* nodes, uint32_t nb, float4* tris) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const BNode& nd = nodes[i];
  if ((nd.flags & FLAG_USED) && (nd.flags & FLAG_LEAF) && nd.count > 0)
    tris[3 * (size_t)(nd.start + nd.count - 1) + 2].w = __builtin_bit_cast(float, kLeafEnd);
}

// Leaf count: grid-stride per thread, then per wave and block, one atomic
// per block.
__global__ __launch_bounds__(256) void k_count_leaves(const BNode* nodes, uint32_t n, uint32_t* cnt) {
  __shared__ uint32_t part[4];
  uint32_t v = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if ((nodes[i].flags & FLAG_USED) && (nodes[i].flags & FLAG_LEAF)) ++v;
  v = wave_sum_u(v);
  if (__lane_id() == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) v += part[w];
    if (v) atomicAdd(cnt, v);
  }
}

inline unsigned blocks(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// ------------------------------------------------- 4-wide traversal tree
// The binary tree (the reference's topology) is collapsed into 4-wide nodes
// for traversal: each 4-wide node takes a binary interior node's two
// children and repeatedly opens the interior child with the largest surface
// area until it has four children (the standard greedy collapse). The set of
// children depends only on the subtree, so every binary node's candidate is
// planned in parallel; reachability from the root marks the ones used.
__device__ __forceinline__ float node_area(const BNode& b) {
  V3 lo = v3(key2f(b.bkey[0]), key2f(b.bkey[1]), key2f(b.bkey[2]));
  V3 hi = v3(key2f(b.bkey[3]), key2f(b.bkey[4]), key2f(b.bkey[5]));
  return surface_area(lo, hi);
}

__device__ __forceinline__ bool is_interior(const BNode& b) { return (b.flags & FLAG_USED) && !(b.flags & FLAG_LEAF); }

__global__ void k_collapse_plan(const BNode* nodes, uint32_t n, int4* cand) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !is_interior(nodes[i])) return;
  int kids[4] = {nodes[i].left, nodes[i].right, -1, -1};
  int nk = 2;
  while (nk < 4) {
    int best = -1;
    float best_a = -1.0f;
    for (int k = 0; k < nk; ++k) {
      const BNode& c = nodes[kids[k]];
      if (is_interior(c)) {
        float a = node_area(c);
        if (a > best_a) {
          best_a = a;
          best = k;
        }
      }
    }
    if (best < 0) break;
    int open = kids[best];
    kids[best] = nodes[open].left;
    kids[nk++] = nodes[open].right;
  }
  cand[i] = make_int4(kids[0], kids[1], kids[2], kids[3]);
}

__global__ void k_collapse_mark(const BNode* nodes, uint32_t n, const int4* cand, uint32_t* used) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !used[i]) return;
  int4 c = cand[i];
  int kk[4] = {c.x, c.y, c.z, c.w};
  for (int k = 0; k < 4; ++k)
    if (kk[k] >= 0 && is_interior(nodes[kk[k]])) used[kk[k]] = 1u;
}

__global__ void k_collapse_root(uint32_t* used) { used[0] = 1u; }

__global__ void k_collapse_write(const BNode* nodes, uint32_t n, const int4* cand, const uint32_t* used,
                                 const uint32_t* map, Node4* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !used[i]) return;
  int4 c = cand[i];
  int kk[4] = {c.x, c.y, c.z, c.w};
  float lo[3][4], hi[3][4];
  int ref[4], cnt[4];
  for (int k = 0; k < 4; ++k) {
    if (kk[k] < 0) {
      for (int a = 0; a < 3; ++a) lo[a][k] = hi[a][k] = 0.0f;
      ref[k] = -1;
      cnt[k] = -1;
      continue;
    }
    const BNode& b = nodes[kk[k]];
    for (int a = 0; a < 3; ++a) {
      lo[a][k] = key2f(b.bkey[a]);
      hi[a][k] = key2f(b.bkey[3 + a]);
    }
    bool leaf = (b.flags & FLAG_LEAF) != 0;
    ref[k] = leaf ? (int)b.start : (int)map[kk[k]];
    cnt[k] = leaf ? (int)b.count : 0;
  }
  Node4 o;
  o.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
  o.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
  o.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
  o.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
  o.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
  o.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
  o.ref = make_int4(ref[0], ref[1], ref[2], ref[3]);
  o.cnt = make_int4(cnt[0], cnt[1], cnt[2], cnt[3]);
  out[map[i]] = o;
}

__global__ void k_collapse_leaf_root(const BNode* nodes, Node4* out) {
  const BNode& r = nodes[0];
  Node4 o;
  o.lox = make_float4(key2f(r.bkey[0]), 0, 0, 0);
  o.hix = make_float4(key2f(r.bkey[3]), 0, 0, 0);
  o.loy = make_float4(key2f(r.bkey[1]), 0, 0, 0);
  o.hiy = make_float4(key2f(r.bkey[4]), 0, 0, 0);
  o.loz = make_float4(key2f(r.bkey[2]), 0, 0, 0);
  o.hiz = make_float4(key2f(r.bkey[5]), 0, 0, 0);
  o.ref = make_int4((int)r.start, -1, -1, -1);
  o.cnt = make_int4((int)r.count, -1, -1, -1);
  out[0] = o;
}

// Outward-rounded cell of plane v on one grid axis: lo planes down, hi
// planes up, with a margin of 2^-20 cells, far above the double rounding of
// (v - g) * 2^-e, so g + q * 2^e <= v (lo) and >= v (hi) in exact arithmetic.
__device__ __forceinline__ uint32_t qcell(float v, float g, int e, bool up) {
  const double x = ((double)v - (double)g) * ldexp(1.0, -e);
  const double q = up ? ceil(x + 0x1p-20) : floor(x - 0x1p-20);
  return (uint32_t)fmin(fmax(q, 0.0), 65535.0);
}

// Node4 -> Node4Q, one node per thread.
__global__ void k_quantize4(const Node4* __restrict__ in, uint32_t n, QGrid grid, Node4Q* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Node4 nd = in[i];
  const float lo[3][4] = {{nd.lox.x, nd.lox.y, nd.lox.z, nd.lox.w},
                          {nd.loy.x, nd.loy.y, nd.loy.z, nd.loy.w},
                          {nd.loz.x, nd.loz.y, nd.loz.z, nd.loz.w}};
  const float hi[3][4] = {{nd.hix.x, nd.hix.y, nd.hix.z, nd.hix.w},
                          {nd.hiy.x, nd.hiy.y, nd.hiy.z, nd.hiy.w},
                          {nd.hiz.x, nd.hiz.y, nd.hiz.z, nd.hiz.w}};
  const int cnt[4] = {nd.cnt.x, nd.cnt.y, nd.cnt.z, nd.cnt.w};
  const int ref[4] = {nd.ref.x, nd.ref.y, nd.ref.z, nd.ref.w};
  uint32_t ql[3][4], qh[3][4];
  int r[4];
  for (int c = 0; c < 4; ++c) {
    for (int a = 0; a < 3; ++a) {
      ql[a][c] = cnt[c] < 0 ? 65535u : qcell(lo[a][c], grid.lo[a], grid.e[a], false);
      qh[a][c] = cnt[c] < 0 ? 0u : qcell(hi[a][c], grid.lo[a], grid.e[a], true);
    }
    r[c] = cnt[c] < 0 ? -1 : (cnt[c] > 0 ? (kQLeaf | ref[c]) : ref[c]);
  }
  Node4Q o;
  uint32_t w[3][4];
  for (int a = 0; a < 3; ++a) {
    w[a][0] = ql[a][0] | (ql[a][1] << 16);
    w[a][1] = ql[a][2] | (ql[a][3] << 16);
    w[a][2] = qh[a][0] | (qh[a][1] << 16);
    w[a][3] = qh[a][2] | (qh[a][3] << 16);
  }
  o.x = make_uint4(w[0][0], w[0][1], w[0][2], w[0][3]);
  o.y = make_uint4(w[1][0], w[1][1], w[1][2], w[1][3]);
  o.z = make_uint4(w[2][0], w[2][1], w[2][2], w[2][3]);
  o.ref = make_int4(r[0], r[1], r[2], r[3]);
  out[i] = o;
}

}  // namespace

// The quantised copy of the 4-wide tree: a grid over the root box (read back
// from the build), cells of 2^e with 65534 cells covering each axis' extent.
void bvh_quantize4(pyre_scene* s, hipStream_t st) {
  dev_free(s->d_nodes4q);
  int32_t key[6];
  PYRE_HIP_CHECK(hipMemcpyAsync(key, reinterpret_cast<const char*>(s->d_bnodes) + offsetof(BNode, bkey), sizeof(key),
                                hipMemcpyDeviceToHost, st));
  PYRE_HIP_CHECK(hipStreamSynchronize(st));
  QGrid g{};
  for (int a = 0; a < 3; ++a) {
    const float lo = key2f(key[a]), hi = key2f(key[3 + a]);
    const double ext = (double)hi - (double)lo;
    // cells of 2^e with 65534 of them spanning the extent (clamped to the
    // exponents whose cells are normal floats; a non-finite extent gets the
    // largest cell, and the outward rounding still covers every box)
    double le = ext > 0.0 ? std::ceil(std::log2(ext / 65534.0)) : 0.0;
    if (!(le >= -120.0)) le = -120.0;
    if (!(le <= 120.0)) le = 120.0;
    g.lo[a] = lo;
    g.e[a] = (int)le;
  }
  s->qgrid = g;
  s->d_nodes4q = dev_alloc<Node4Q>(s->n_nodes4);
  hipLaunchKernelGGL(k_quantize4, dim3(blocks(s->n_nodes4, 256)), dim3(256), 0, st, s->d_nodes4, s->n_nodes4, g,
                     s->d_nodes4q);
  PYRE_HIP_CHECK(hipGetLastError());
}

void bvh_collapse4(pyre_scene* s, Scratch& S, hipStream_t st) {
  const uint32_t nb = s->n_bnodes;
  dev_free(s->d_nodes4);
  if (s->n_interior == 0) {
    s->n_nodes4 = 1;
    s->d_nodes4 = dev_alloc<Node4>(1);
    hipLaunchKernelGGL(k_collapse_leaf_root, dim3(1), dim3(1), 0, st, s->d_bnodes, s->d_nodes4);
    s->depth4 = 1;
    return;
  }
  int4* cand = S.get<int4>(nb);
  uint32_t* used = S.get<uint32_t>(nb + 1);
  uint32_t* map = S.get<uint32_t>(nb + 1);
  uint32_t* total = S.get<uint32_t>(1);
  PYRE_HIP_CHECK(hipMemsetAsync(used, 0, 4 * (size_t)(nb + 1), st));
  hipLaunchKernelGGL(k_collapse_plan, dim3(blocks(nb, 256)), dim3(256), 0, st, s->d_bnodes, nb, cand);
  hipLaunchKernelGGL(k_collapse_root, dim3(1), dim3(1), 0, st, used);
  // every 4-wide level consumes at least one binary level
  for (uint32_t it = 0; it <= s->depth; ++it)
    hipLaunchKernelGGL(k_collapse_mark, dim3(blocks(nb, 256)), dim3(256), 0, st, s->d_bnodes, nb, cand, used);
  exclusive_scan(S, used, map, nb + 1, total, st);
  uint32_t h_total = 0;
  PYRE_HIP_CHECK(hipMemcpyAsync(&h_total, total, 4, hipMemcpyDeviceToHost, st));
  PYRE_HIP_CHECK(hipStreamSynchronize(st));
  s->n_nodes4 = h_total;
  s->d_nodes4 = dev_alloc<Node4>(h_total);
  hipLaunchKernelGGL(k_collapse_write, dim3(blocks(nb, 256)), dim3(256), 0, st, s->d_bnodes, nb, cand, used, map,
                     s->d_nodes4);
  // each 4-wide level spans at most two binary levels and at least one
  s->depth4 = s->depth;
}

void bvh_release(pyre_scene* s) {
  dev_free(s->d_nodes4);
  dev_free(s->d_nodes4q);
  dev_free(s->d_nodes);
  dev_free(s->d_tris);
  dev_free(s->d_prim_order);
  dev_free(s->d_bnodes);
  s->built = false;
}

void bvh_build(pyre_scene* s, const pyre_bvh_params& prm, hipStream_t st) {
  require(prm.bins >= 2 && prm.bins <= (uint32_t)kMaxBins, "bvh bins must be in [2, 64]");
  require(prm.max_leaf >= 1 && prm.max_leaf <= 65535, "bvh max_leaf must be in [1, 65535]");
  require(s->n_tris > 0, "scene has no triangles");
  bvh_release(s);
  auto t0 = std::chrono::steady_clock::now();
  const uint32_t n = s->n_tris;
  BuildParams p{prm.bins, prm.max_leaf, prm.traversal_cost, prm.intersect_cost};
  Scratch S;
  float4* plo = S.get<float4>(n);
  float4* phi = S.get<float4>(n);
  float4* pc = S.get<float4>(n);
  uint32_t* ref = dev_alloc<uint32_t>(n);
  s->d_prim_order = ref;
  uint32_t* tmp = S.get<uint32_t>(n);
  const size_t cap = 8 + 4 * ((size_t)n / kSubtreeMax + 1) + 2 * (size_t)n;
  BNode* nodes = dev_alloc<BNode>(cap);
  s->d_bnodes = nodes;
  PYRE_HIP_CHECK(hipMemsetAsync(nodes, 0, cap * sizeof(BNode), st));
  Counters* ctr = S.get<Counters>(1);

  // subtree roots: at most one per kSubtreeMax/2... bounded by the node count
  const size_t max_sub = 2 * ((size_t)n / 1 + 1);
  uint32_t* sub_roots = S.get<uint32_t>(std::min(max_sub, cap));
  uint32_t* sub_base = S.get<uint32_t>(std::min(max_sub, cap));
  const size_t max_lvl = 4 * ((size_t)n / kSubtreeMax + 1);
  uint32_t* lvl = S.get<uint32_t>(max_lvl);
  uint32_t* next_lvl = S.get<uint32_t>(max_lvl);
  const size_t max_chunks = (size_t)n / kChunk + max_lvl + 1;
  uint32_t* nch = S.get<uint32_t>(max_lvl + 1);
  uint32_t* nch_off = S.get<uint32_t>(max_lvl + 1);
  uint32_t* chunk_node = S.get<uint32_t>(max_chunks);
  uint32_t* chunk_first = S.get<uint32_t>(max_chunks);
  uint32_t* chunk_left = S.get<uint32_t>(max_chunks);
  uint32_t* chunk_loff = S.get<uint32_t>(max_chunks);
  int32_t* gbins = S.get<int32_t>(max_lvl * kMaxBins * 7);

  const bool small = n <= kSubtreeMax;
  hipLaunchKernelGGL(k_init_root, dim3(1), dim3(1), 0, st, nodes, n, ctr, lvl, sub_roots, sub_base,
                     small ? 1u : 0u);
  Counters h{};
  h.n_subtrees = small ? 1 : 0;
  h.region_cursor = small ? 1 + 2 * n : 0;
  h.node_cursor = 1;
  uint32_t L = small ? 0 : 1;
  hipLaunchKernelGGL(k_prep, dim3(std::min<unsigned>(blocks(n, 256), (unsigned)std::max(1, s->ctx->cus) * 8)),
                     dim3(256), 0, st, s->d_verts, n, plo, phi, pc, ref, nodes, s->d_motion, s->d_tri_offset,
                     s->n_meshes);
  uint32_t level = 0;
  while (L > 0) {
    // L + 1 threads: the last writes the scan sentinel nch[L] = 0
    hipLaunchKernelGGL(k_chunk_counts, dim3(blocks((size_t)L + 1, 256)), dim3(256), 0, st, nodes, lvl, L, nch);
    exclusive_scan(S, nch, nch_off, L + 1, &ctr->total_chunks, st);
    hipLaunchKernelGGL(k_chunk_fill, dim3(blocks(L, 256)), dim3(256), 0, st, nodes, lvl, L, nch_off, chunk_node,
                       chunk_first);
    uint32_t total_chunks = 0;
    PYRE_HIP_CHECK(hipMemcpyAsync(&total_chunks, &ctr->total_chunks, 4, hipMemcpyDeviceToHost, st));
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
    require(total_chunks <= max_chunks, "internal: chunk table overflow");
    hipLaunchKernelGGL(k_gbins_init, dim3(blocks((size_t)L * kMaxBins, 256)), dim3(256), 0, st, gbins, L);
    hipLaunchKernelGGL(k_wide_bin, dim3(total_chunks), dim3(kWideThreads), 0, st, nodes, lvl, chunk_node,
                       chunk_first, &ctr->total_chunks, ref, plo, phi, pc, gbins, p);
    hipLaunchKernelGGL(k_wide_decide, dim3(L), dim3(64), 0, st, nodes, lvl, gbins, p);
    hipLaunchKernelGGL(k_wide_count, dim3(total_chunks), dim3(kWideThreads), 0, st, nodes, lvl, chunk_node,
                       chunk_first, &ctr->total_chunks, ref, pc, chunk_left);
    hipLaunchKernelGGL(k_wide_finalize, dim3(1), dim3(1024), 0, st, nodes, lvl, L, nch_off, chunk_left, chunk_loff,
                       next_lvl, sub_roots, sub_base, ctr, p, level);
    hipLaunchKernelGGL(k_wide_scatter, dim3(total_chunks), dim3(kWideThreads), 0, st, nodes, lvl, chunk_node,
                       chunk_first, &ctr->total_chunks, chunk_loff, ref, tmp, plo, phi, pc);
    hipLaunchKernelGGL(k_wide_copyback, dim3(total_chunks), dim3(kWideThreads), 0, st, nodes, lvl, chunk_node,
                       chunk_first, &ctr->total_chunks, ref, tmp);
    PYRE_HIP_CHECK(hipMemcpyAsync(&h, ctr, sizeof(Counters), hipMemcpyDeviceToHost, st));
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
    L = h.next_big;
    require((size_t)L <= max_lvl, "internal: level list overflow");
    std::swap(lvl, next_lvl);
    level++;
  }
  if (!small) {
    // place the subtree regions after the wide-phase nodes
    const uint32_t region_start = h.node_cursor;
    if (h.n_subtrees > 0)
      hipLaunchKernelGGL(k_shift, dim3(blocks(h.n_subtrees, 256)), dim3(256), 0, st, sub_base, h.n_subtrees,
                         region_start);
    h.region_cursor += region_start;
  }
  require((size_t)h.region_cursor <= cap, "internal: build node overflow");
  if (h.n_subtrees > 0)
    hipLaunchKernelGGL(k_subtree, dim3(h.n_subtrees), dim3(64), 0, st, nodes, sub_roots, sub_base, h.n_subtrees,
                       ref, plo, phi, pc, p, ctr);
  const uint32_t nb = h.region_cursor;
  s->n_bnodes = nb;

  // flatten
  uint32_t* flag = S.get<uint32_t>(nb + 1);
  uint32_t* map = S.get<uint32_t>(nb + 1);
  uint32_t* total = S.get<uint32_t>(2);
  PYRE_HIP_CHECK(hipMemsetAsync(total, 0, 8, st));
  hipLaunchKernelGGL(k_interior_flags, dim3(blocks(nb, 256)), dim3(256), 0, st, nodes, nb, flag);
  exclusive_scan(S, flag, map, nb, total, st);
  hipLaunchKernelGGL(k_count_leaves, dim3(std::min<unsigned>(blocks(nb, 256), (unsigned)std::max(1, s->ctx->cus) * 8)),
                     dim3(256), 0, st, nodes, nb, total + 1);
  uint32_t ht[2];
  Counters hc{};
  PYRE_HIP_CHECK(hipMemcpyAsync(ht, total, 8, hipMemcpyDeviceToHost, st));
  PYRE_HIP_CHECK(hipMemcpyAsync(&hc, ctr, sizeof(Counters), hipMemcpyDeviceToHost, st));
  PYRE_HIP_CHECK(hipStreamSynchronize(st));
  s->n_interior = ht[0];
  s->n_leaves = ht[1];
  s->depth = hc.depth;
  s->n_nodes2 = std::max(ht[0], 1u);
  s->d_nodes = dev_alloc<Node2>(s->n_nodes2);
  if (ht[0] == 0)
    hipLaunchKernelGGL(k_flatten_leaf_root, dim3(1), dim3(1), 0, st, nodes, s->d_nodes);
  else
    hipLaunchKernelGGL(k_flatten, dim3(blocks(nb, 256)), dim3(256), 0, st, nodes, nb, flag, map, s->d_nodes);
  s->d_tris = dev_alloc<float4>(3 * (size_t)n);
  uint32_t* moving = nullptr;
  if (s->has_motion) {
    moving = S.get<uint32_t>(s->n_meshes);
    PYRE_HIP_CHECK(hipMemcpyAsync(moving, s->h_moving.data(), s->n_meshes * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  }
  hipLaunchKernelGGL(k_reorder_tris, dim3(blocks(n, 256)), dim3(256), 0, st, s->d_verts, ref, n, s->d_tris, moving,
                     s->d_tri_offset, s->n_meshes);
  hipLaunchKernelGGL(k_mark_leaf_ends, dim3(blocks(nb, 256)), dim3(256), 0, st, nodes, nb, s->d_tris);
  bvh_collapse4(s, S, st);
  bvh_quantize4(s, st);
  PYRE_HIP_CHECK(hipGetLastError());
  PYRE_HIP_CHECK(hipStreamSynchronize(st));
  s->built = true;
  s->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace pyre
