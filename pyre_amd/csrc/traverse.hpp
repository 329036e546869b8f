// BVH traversal core shared by the batched trace kernels and the
// integrator's extend/shadow stages.
//
// Semantics follow Bvh::intersect + moller_trumbore
// (crates/pyre/src/geometry/bvh.rs:86-153, mesh.rs:153-178) and
// Scene::intersect / occluded (scene.rs:57-140):
//   * the slab test is the reference's ray_aabb (bvh.rs:146-153) against the
//     current closest t (inclusive: tmin <= tmax);
//   * Moller-Trumbore runs the same f32 operations on (v0, e1, e2), so t/u/v
//     are bit-identical to the reference for the same triangle and ray;
//   * closest hit = minimum t over triangles with t in (t_min, t_max); an exact
//     tie keeps the lower scene-global triangle id, i.e. the lower primitive
//     (scene.rs:65 keeps the first primitive on ties);
//   * any hit = some triangle with t in (t_min, t_max), early exit.
// One thread per ray; 64-byte two-box nodes; leaf triangles as
// (v0|gid, e1, e2) float4 triples; the traversal stack is a per-thread column
// of an LDS array (or of a global array for very deep trees).
#pragma once

#include <type_traits>

#include "internal.hpp"

namespace pyre {

struct TRay {
  V3 o, d, inv;
  float t_min, t_max;
  float time;  // shutter time; read only by the traversals of scenes with moving meshes
};

PH TRay make_tray(V3 o, V3 d, float t_min, float t_max, float time = 0.0f) {
  TRay r;
  r.o = o;
  r.d = d;
  r.t_min = t_min;
  r.t_max = t_max;
  r.time = time;
  r.inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);  // bvh.rs:93
  return r;
}

// ray_aabb (bvh.rs:146-153) for one child of a two-box node.
__device__ __forceinline__ bool slab_test(const TRay& r, float lx, float hx, float ly, float hy, float lz, float hz,
                                          float t_hi, float* t_enter) {
  float t1x = (lx - r.o.x) * r.inv.x, t2x = (hx - r.o.x) * r.inv.x;
  float t1y = (ly - r.o.y) * r.inv.y, t2y = (hy - r.o.y) * r.inv.y;
  float t1z = (lz - r.o.z) * r.inv.z, t2z = (hz - r.o.z) * r.inv.z;
  float tmin = rmax(rmax(rmin(t1x, t2x), rmax(rmin(t1y, t2y), rmin(t1z, t2z))), r.t_min);
  float tmax = rmin(rmin(rmax(t1x, t2x), rmin(rmax(t1y, t2y), rmax(t1z, t2z))), t_hi);
  *t_enter = tmin;
  return tmin <= tmax;
}

struct THit {
  uint32_t gid;  // PYRE_MISS when nothing was hit
  float t, u, v;
};

// Returns false on stack overflow (result then incomplete).
template <bool ANY>
__device__ __forceinline__ bool traverse(const Node2* __restrict__ nodes, const float4* __restrict__ tris,
                                         const TRay& r, int32_t* stack, size_t stride, int cap, THit* out) {
  float best_t = r.t_max;
  uint32_t best_gid = PYRE_MISS;
  float bu = 0.0f, bv = 0.0f;
  bool found = false;
  int sp = 0;
  int32_t cur = 0;
  bool ok = true;
  for (;;) {
    const Node2 nd = nodes[cur];
    float e0 = 0.0f, e1 = 0.0f;
    bool h0 = slab_test(r, nd.x.x, nd.x.y, nd.y.x, nd.y.y, nd.z.x, nd.z.y, best_t, &e0);
    bool h1 = nd.meta.y >= 0 && slab_test(r, nd.x.z, nd.x.w, nd.y.z, nd.y.w, nd.z.z, nd.z.w, best_t, &e1);
    int lc0 = h0 ? nd.meta.z : 0, lc1 = h1 ? nd.meta.w : 0;
    if (lc0 | lc1) {
      // leaves first, the nearer one first
      int first = (lc0 && lc1 && e1 < e0) ? 1 : 0;
      for (int q = 0; q < 2; ++q) {
        int c = q == 0 ? first : 1 - first;
        int cnt = c == 0 ? lc0 : lc1;
        int32_t s = c == 0 ? nd.meta.x : nd.meta.y;
        for (int k = 0; k < cnt; ++k) {
          const float4 a = tris[3 * (size_t)(s + k)];
          const float4 b = tris[3 * (size_t)(s + k) + 1];
          const float4 cc = tris[3 * (size_t)(s + k) + 2];
          float t, u, v;
          if (ANY) {
            if (mt_intersect(r.o, r.d, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(cc.x, cc.y, cc.z), r.t_min, r.t_max,
                             false, &t, &u, &v)) {
              found = true;
              break;
            }
          } else if (mt_intersect(r.o, r.d, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(cc.x, cc.y, cc.z), r.t_min,
                                  best_t, true, &t, &u, &v)) {
            uint32_t gid = __builtin_bit_cast(uint32_t, a.w);
            if (t < best_t || (found && gid < best_gid)) {
              best_t = t;
              best_gid = gid;
              bu = u;
              bv = v;
              found = true;
            }
          }
        }
        if (ANY && found) break;
      }
      if (ANY && found) break;
      if (lc0) h0 = false;
      if (lc1) h1 = false;
    }
    if (h0 && h1) {
      int32_t nearc = e0 <= e1 ? nd.meta.x : nd.meta.y;
      int32_t farc = e0 <= e1 ? nd.meta.y : nd.meta.x;
      if (sp >= cap) {
        ok = false;
        break;
      }
      stack[(size_t)sp * stride] = farc;
      sp++;
      cur = nearc;
    } else if (h0) {
      cur = nd.meta.x;
    } else if (h1) {
      cur = nd.meta.y;
    } else {
      if (sp == 0) break;
      sp--;
      cur = stack[(size_t)sp * stride];
    }
  }
  out->gid = ANY ? (found ? 0u : PYRE_MISS) : (found ? best_gid : PYRE_MISS);
  out->t = found ? best_t : 0.0f;
  out->u = bu;
  out->v = bv;
  return ok;
}

// Per-thread traversal stack of (node, entry distance) pairs: a ring of SL
// entries in LDS (a column of a [SL][threads] array, so a wave's accesses at
// equal depth hit distinct banks) whose oldest entries spill to a per-thread
// column of a global array when the ring is full. Positions below `lo` live
// in global memory.
template <int SL>
struct HybridStack {
  int2* lds;
  int lstride;
  int2* glb;
  size_t gstride;
  int sp, lo;
  __device__ __forceinline__ void push(int2 v) {
    if (sp - lo == SL) {
      glb[(size_t)lo * gstride] = lds[(lo & (SL - 1)) * lstride];
      lo++;
    }
    lds[(sp & (SL - 1)) * lstride] = v;
    sp++;
  }
  __device__ __forceinline__ int2 pop() {
    sp--;
    if (sp < lo) {
      lo = sp;
      return glb[(size_t)sp * gstride];
    }
    return lds[(sp & (SL - 1)) * lstride];
  }
};

// A moving mesh's ray origin (mesh.rs:109-124): the ray shifted by
// -lerp(offset_t0, offset_t1, clamp(time, 0, 1)); a zero offset leaves the
// ray as it is. slot = the mesh's motion slot + 1 (0: static).
__device__ __forceinline__ V3 motion_origin(const float4* __restrict__ motion, uint32_t slot, const TRay& r) {
  if (slot == 0) return r.o;
  const float4 m0 = motion[2 * (slot - 1)], m1 = motion[2 * (slot - 1) + 1];
  const V3 off = lerp(v3(m0.x, m0.y, m0.z), v3(m1.x, m1.y, m1.z), rclamp(r.time, 0.0f, 1.0f));
  return (off.x == 0.0f && off.y == 0.0f && off.z == 0.0f) ? r.o : r.o - off;
}

// Leaf triangles [s, s + cnt): Moller-Trumbore against the current closest.
// MOTION: the triangle records carry their mesh's motion slot in e1.w and a
// moving mesh's triangles are tested with the shifted origin (the boxes
// above them are swept, so the slab tests use the unshifted ray).
// ENDMARK: the leaf runs from slot s to the slot marked kLeafEnd (cnt unused).
template <bool ANY, bool MOTION = false, bool ENDMARK = false>
__device__ __forceinline__ bool leaf_hits(const float4* __restrict__ tris, const TRay& r, int32_t s, int cnt,
                                          float& best_t, uint32_t& best_gid, float& bu, float& bv, bool& found,
                                          const float4* __restrict__ motion = nullptr) {
  for (int k = 0;; ++k) {
    if (!ENDMARK && k >= cnt) break;
    const float4 a = tris[3 * (size_t)(s + k)];
    const float4 b = tris[3 * (size_t)(s + k) + 1];
    const float4 c = tris[3 * (size_t)(s + k) + 2];
    V3 o = r.o;
    if constexpr (MOTION) o = motion_origin(motion, __builtin_bit_cast(uint32_t, b.w), r);
    float t, u, v;
    if (ANY) {
      if (mt_intersect(o, r.d, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), r.t_min, r.t_max, false, &t,
                       &u, &v)) {
        found = true;
        return true;
      }
    } else {
      // the closest-hit update as selects (no branch on the test's outcome)
      const bool ok = mt_intersect(o, r.d, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), r.t_min, best_t,
                                   true, &t, &u, &v);
      const uint32_t gid = __builtin_bit_cast(uint32_t, a.w);
      const bool take = ok && (t < best_t || (found && gid < best_gid));
      best_t = take ? t : best_t;
      best_gid = take ? gid : best_gid;
      bu = take ? u : bu;
      bv = take ? v : bv;
      found = found || take;
    }
    if (ENDMARK && __builtin_bit_cast(uint32_t, c.w) == kLeafEnd) break;
  }
  return false;
}

// Per-ray terms of the slab test on quantised nodes (Node4Q). A child plane
// of grid cell q on axis k lies at t = q * a_k + b_k along the ray, with
// a_k = 2^e_k * inv_k (exact) and b_k = (grid lo_k - o_k) * inv_k. That form
// and the reference's (plane - o) * inv (bvh.rs:146-153) each round within a
// few ulps of |q a_k| + |b_k| <= 65536 |a_k| + |b_k|; pad_k is 16 ulps of
// that bound, and it is folded into the offsets of the axis' lo and hi planes
// (moving the nearer plane nearer and the farther one farther, by the sign of
// a_k), so on boxes already rounded outward to the grid every child interval
// contains the reference's and the test accepts every box the reference's
// test on the exact box accepts. An axis whose terms are not finite (a
// direction component of 0 or near it) gets NaN planes, which the min/max
// ignore: no bound from that axis, a wider test again.
struct QRay {
  V3 a, blo, bhi;
};

__device__ __forceinline__ QRay make_qray(const TRay& r, const QGrid& g) {
  const float inv[3] = {r.inv.x, r.inv.y, r.inv.z}, o[3] = {r.o.x, r.o.y, r.o.z};
  float av[3], lo[3], hi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float a = ldexpf(inv[k], g.e[k]);
    const float b = (g.lo[k] - o[k]) * inv[k];
    const bool fin = __builtin_isfinite(a) && __builtin_isfinite(b);
    const float pad = 0x1p-20f * fmaf(65536.0f, fabsf(a), fabsf(b));
    const float sp = a >= 0.0f ? pad : -pad;  // lo plane nearer when a >= 0
    av[k] = fin ? a : 0.0f;
    lo[k] = fin ? b - sp : __builtin_nanf("");
    hi[k] = fin ? b + sp : __builtin_nanf("");
  }
  QRay q;
  q.a = v3(av[0], av[1], av[2]);
  q.blo = v3(lo[0], lo[1], lo[2]);
  q.bhi = v3(hi[0], hi[1], hi[2]);
  return q;
}

// The slab test of one quantised child box (cells lx..hz), widened per axis.
__device__ __forceinline__ bool qslab_test(const QRay& q, float t_min, uint32_t lx, uint32_t hx, uint32_t ly,
                                           uint32_t hy, uint32_t lz, uint32_t hz, float t_hi, float* t_enter) {
  const float t1x = fmaf((float)lx, q.a.x, q.blo.x), t2x = fmaf((float)hx, q.a.x, q.bhi.x);
  const float t1y = fmaf((float)ly, q.a.y, q.blo.y), t2y = fmaf((float)hy, q.a.y, q.bhi.y);
  const float t1z = fmaf((float)lz, q.a.z, q.blo.z), t2z = fmaf((float)hz, q.a.z, q.bhi.z);
  const float tmin = rmax(rmax(rmin(t1x, t2x), rmax(rmin(t1y, t2y), rmin(t1z, t2z))), t_min);
  const float tmax = rmin(rmin(rmax(t1x, t2x), rmin(rmax(t1y, t2y), rmax(t1z, t2z))), t_hi);
  *t_enter = tmin;
  return tmin <= tmax;
}

__device__ __forceinline__ void cswap(float& ka, int32_t& ra, int& ca, float& kb, int32_t& rb, int& cb) {
  if (kb < ka) {
    float tk = ka;
    ka = kb;
    kb = tk;
    int32_t tr = ra;
    ra = rb;
    rb = tr;
    int tc = ca;
    ca = cb;
    cb = tc;
  }
}

// Compare-exchange of (key, reference) pairs: the pair with the smaller key
// first, ties kept in place.
__device__ __forceinline__ void cswap2(float& ka, int32_t& ra, float& kb, int32_t& rb) {
  if (kb < ka) {
    const float tk = ka;
    ka = kb;
    kb = tk;
    const int32_t tr = ra;
    ra = rb;
    rb = tr;
  }
}

// Traversal of the 4-wide tree as a resumable state machine, one node visit
// per step() so a persistent kernel can hand a finished lane a new ray at
// any step. Same hit semantics as traverse(): children are visited
// nearest-first by slab entry distance, leaves as soon as their box is hit;
// a stacked child whose entry distance lies beyond the closest hit found
// since it was pushed is dropped without fetching it (the reference's slab
// test against the tightened t_max would reject it, bvh.rs:146-153).
//
// QNODE: the tree of quantised nodes (Node4Q) with the widened slab test of
// QRay: every box the exact test accepts is visited, so the set of
// triangles that can hold the closest hit is tested as before and the hit
// (least t, then least scene-global id) is the same.
template <bool ANY, int SL, bool MOTION = false, bool QNODE = false>
struct Trav4 {
  using NodeT = std::conditional_t<QNODE, Node4Q, Node4>;
  TRay r;
  QRay qr;  // QNODE only
  float best_t, bu, bv;
  uint32_t best_gid;
  bool found;
  int32_t cur;

  // grid: the tree's quantisation grid (QNODE only)
  __device__ __forceinline__ void start(const TRay& ray, HybridStack<SL>& st, const QGrid* grid = nullptr) {
    r = ray;
    if constexpr (QNODE) qr = make_qray(ray, *grid);
    best_t = ray.t_max;
    bu = bv = 0.0f;
    best_gid = PYRE_MISS;
    found = false;
    cur = 0;
    st.sp = 0;
    st.lo = 0;
  }

  // Visits node `cur`; true when the ray is finished. motion: the scene's
  // motion table (MOTION only).
  __device__ __forceinline__ bool step(const NodeT* __restrict__ nodes, const float4* __restrict__ tris,
                                       HybridStack<SL>& st, const float4* __restrict__ motion = nullptr) {
    const NodeT nd = nodes[cur];
    // missed children: key +inf and count -1 (a hit child can have key +inf
    // only if its box is at infinity, so the count marks misses)
    // all four slab tests without branches (a missing child's box is
    // tested too and masked out), so the lanes of a wave stay converged
    float e0, e1, e2, e3;
    bool h0, h1, h2, h3;
    int32_t r0, r1, r2, r3;
    if constexpr (QNODE) {
      // child i's cells: lo/hi words (i / 2) and (2 + i / 2), half i % 2;
      // the reference word tells leaves (kQLeaf set; the end mark in the
      // triangle records ends them) from inner nodes, and -1 marks a miss
      constexpr uint32_t m = 0xFFFFu;
      h0 = (nd.ref.x >= 0) & qslab_test(qr, r.t_min, nd.x.x & m, nd.x.z & m, nd.y.x & m, nd.y.z & m, nd.z.x & m,
                                        nd.z.z & m, best_t, &e0);
      h1 = (nd.ref.y >= 0) & qslab_test(qr, r.t_min, nd.x.x >> 16, nd.x.z >> 16, nd.y.x >> 16, nd.y.z >> 16,
                                        nd.z.x >> 16, nd.z.z >> 16, best_t, &e1);
      h2 = (nd.ref.z >= 0) & qslab_test(qr, r.t_min, nd.x.y & m, nd.x.w & m, nd.y.y & m, nd.y.w & m, nd.z.y & m,
                                        nd.z.w & m, best_t, &e2);
      h3 = (nd.ref.w >= 0) & qslab_test(qr, r.t_min, nd.x.y >> 16, nd.x.w >> 16, nd.y.y >> 16, nd.y.w >> 16,
                                        nd.z.y >> 16, nd.z.w >> 16, best_t, &e3);
      float k0 = h0 ? e0 : kInf, k1 = h1 ? e1 : kInf, k2 = h2 ? e2 : kInf, k3 = h3 ? e3 : kInf;
      r0 = h0 ? nd.ref.x : -1;
      r1 = h1 ? nd.ref.y : -1;
      r2 = h2 ? nd.ref.z : -1;
      r3 = h3 ? nd.ref.w : -1;
      cswap2(k0, r0, k1, r1);
      cswap2(k2, r2, k3, r3);
      cswap2(k0, r0, k2, r2);
      cswap2(k1, r1, k3, r3);
      cswap2(k1, r1, k2, r2);
      if (qleaf(tris, k0, r0, motion) || qleaf(tris, k1, r1, motion) || qleaf(tris, k2, r2, motion) ||
          qleaf(tris, k3, r3, motion))
        return true;
      int2 next = make_int2(-1, 0);
      qinner(st, next, k3, r3);
      qinner(st, next, k2, r2);
      qinner(st, next, k1, r1);
      qinner(st, next, k0, r0);
      return advance(st, next);
    } else {
      int c0, c1, c2, c3;
      h0 = (nd.cnt.x >= 0) & slab_test(r, nd.lox.x, nd.hix.x, nd.loy.x, nd.hiy.x, nd.loz.x, nd.hiz.x, best_t, &e0);
      h1 = (nd.cnt.y >= 0) & slab_test(r, nd.lox.y, nd.hix.y, nd.loy.y, nd.hiy.y, nd.loz.y, nd.hiz.y, best_t, &e1);
      h2 = (nd.cnt.z >= 0) & slab_test(r, nd.lox.z, nd.hix.z, nd.loy.z, nd.hiy.z, nd.loz.z, nd.hiz.z, best_t, &e2);
      h3 = (nd.cnt.w >= 0) & slab_test(r, nd.lox.w, nd.hix.w, nd.loy.w, nd.hiy.w, nd.loz.w, nd.hiz.w, best_t, &e3);
      c0 = h0 ? nd.cnt.x : -1;
      c1 = h1 ? nd.cnt.y : -1;
      c2 = h2 ? nd.cnt.z : -1;
      c3 = h3 ? nd.cnt.w : -1;
      r0 = nd.ref.x;
      r1 = nd.ref.y;
      r2 = nd.ref.z;
      r3 = nd.ref.w;
      float k0 = h0 ? e0 : kInf, k1 = h1 ? e1 : kInf, k2 = h2 ? e2 : kInf, k3 = h3 ? e3 : kInf;
      // sort the four children by entry distance (misses sort last as +inf)
      cswap(k0, r0, c0, k1, r1, c1);
      cswap(k2, r2, c2, k3, r3, c3);
      cswap(k0, r0, c0, k2, r2, c2);
      cswap(k1, r1, c1, k3, r3, c3);
      cswap(k1, r1, c1, k2, r2, c2);
      // leaves first, nearest first; a child entered beyond the closest hit
      // found meanwhile is skipped
      if (leaf(tris, k0, r0, c0, motion) || leaf(tris, k1, r1, c1, motion) || leaf(tris, k2, r2, c2, motion) ||
          leaf(tris, k3, r3, c3, motion))
        return true;
      // inner children: the nearest becomes the next node, the others are
      // stacked farthest first
      int2 next = make_int2(-1, 0);
      inner(st, next, k3, r3, c3);
      inner(st, next, k2, r2, c2);
      inner(st, next, k1, r1, c1);
      inner(st, next, k0, r0, c0);
      return advance(st, next);
    }
  }

  // The next node: the nearest hit inner child, else the nearest stacked
  // entry still in range; true when there is none (the ray is finished).
  __device__ __forceinline__ bool advance(HybridStack<SL>& st, int2 next) {
    if (next.x >= 0) {
      cur = next.x;
      return false;
    }
    while (st.sp > 0) {
      const int2 en = st.pop();
      if (ANY || __builtin_bit_cast(float, en.y) <= best_t) {
        cur = en.x;
        return false;
      }
    }
    return true;
  }

  // QNODE: a hit leaf child (reference word with kQLeaf set); true when an
  // any-hit query is answered.
  __device__ __forceinline__ bool qleaf(const float4* __restrict__ tris, float k, int32_t ref,
                                        const float4* __restrict__ motion) {
    if (ref < kQLeaf || k > best_t) return false;
    return leaf_hits<ANY, MOTION, true>(tris, r, ref & (kQLeaf - 1), 0, best_t, best_gid, bu, bv, found, motion);
  }

  // QNODE: a hit inner child (reference word in [0, kQLeaf)).
  __device__ __forceinline__ void qinner(HybridStack<SL>& st, int2& next, float k, int32_t ref) const {
    if ((uint32_t)ref >= kQLeaf || k > best_t) return;
    if (next.x >= 0) st.push(next);
    next = make_int2(ref, __builtin_bit_cast(int32_t, k));
  }

  // A hit leaf child: its triangles; true when an any-hit query is answered.
  __device__ __forceinline__ bool leaf(const float4* __restrict__ tris, float k, int32_t ref, int cnt,
                                       const float4* __restrict__ motion) {
    if (cnt <= 0 || k > best_t) return false;
    return leaf_hits<ANY, MOTION, QNODE>(tris, r, ref, cnt, best_t, best_gid, bu, bv, found, motion);
  }

  __device__ __forceinline__ void inner(HybridStack<SL>& st, int2& next, float k, int32_t ref, int cnt) const {
    if (cnt != 0 || k > best_t) return;
    if (next.x >= 0) st.push(next);
    next = make_int2(ref, __builtin_bit_cast(int32_t, k));
  }

  __device__ __forceinline__ void result(THit* out) const {
    out->gid = ANY ? (found ? 0u : PYRE_MISS) : (found ? best_gid : PYRE_MISS);
    out->t = found ? best_t : 0.0f;
    out->u = bu;
    out->v = bv;
  }
};

// Traversal of the 4-wide tree with leaf tests decoupled from node visits:
// each step() visits at most one node and tests at most one triangle. A
// lane keeps one pending leaf (the next slot to test); further hit leaves
// wait on the traversal stack beside the inner nodes, marked by a negative
// reference (~slot). A wave thus runs the triangle test once per step for
// all lanes that have one pending, instead of the longest leaf loop of any
// lane at every node visit.
//
// Hit semantics as Trav4: the closest hit is the least (t, scene-global id)
// over the triangles tested with t in (t_min, t_max), which does not depend
// on the order triangles are tested in; children and stacked entries
// (nodes and leaves) whose slab entry distance lies beyond the closest hit
// found so far are dropped, as the reference's slab test against the
// tightened t_max would reject them (bvh.rs:146-153).
template <bool ANY, int SL, bool MOTION = false>
struct Trav4Q {
  using NodeT = Node4;
  TRay r;
  float best_t, bu, bv;
  uint32_t best_gid;
  bool found;
  int32_t cur;  // next inner node to visit; -1: none
  int32_t tri;  // next triangle slot of the pending leaf; -1: none

  __device__ __forceinline__ void start(const TRay& ray, HybridStack<SL>& st, const QGrid* = nullptr) {
    r = ray;
    best_t = ray.t_max;
    bu = bv = 0.0f;
    best_gid = PYRE_MISS;
    found = false;
    cur = 0;
    tri = -1;
    st.sp = 0;
    st.lo = 0;
  }

  // One step; true when the ray is finished. motion: the scene's motion
  // table (MOTION only).
  __device__ __forceinline__ bool step(const Node4* __restrict__ nodes, const float4* __restrict__ tris,
                                       HybridStack<SL>& st, const float4* __restrict__ motion = nullptr) {
    if (cur < 0 && tri < 0) {
      // nothing in hand: the nearest stacked entry still in range
      bool got = false;
      while (st.sp > 0) {
        const int2 en = st.pop();
        if (ANY || __builtin_bit_cast(float, en.y) <= best_t) {
          if (en.x >= 0)
            cur = en.x;
          else
            tri = ~en.x;
          got = true;
          break;
        }
      }
      if (!got) return true;
    }
    if (tri >= 0) {
      const float4 a = tris[3 * (size_t)tri];
      const float4 b = tris[3 * (size_t)tri + 1];
      const float4 c = tris[3 * (size_t)tri + 2];
      V3 o = r.o;
      if constexpr (MOTION) o = motion_origin(motion, __builtin_bit_cast(uint32_t, b.w), r);
      float t, u, v;
      if (ANY) {
        if (mt_intersect(o, r.d, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), r.t_min, r.t_max, false,
                         &t, &u, &v)) {
          found = true;
          return true;
        }
      } else {
        const bool ok = mt_intersect(o, r.d, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), r.t_min,
                                     best_t, true, &t, &u, &v);
        const uint32_t gid = __builtin_bit_cast(uint32_t, a.w);
        const bool take = ok && (t < best_t || (found && gid < best_gid));
        best_t = take ? t : best_t;
        best_gid = take ? gid : best_gid;
        bu = take ? u : bu;
        bv = take ? v : bv;
        found = found || take;
      }
      tri = __builtin_bit_cast(uint32_t, c.w) == kLeafEnd ? -1 : tri + 1;
    }
    if (cur >= 0) {
      const Node4 nd = nodes[cur];
      float e0, e1, e2, e3;
      const bool h0 = (nd.cnt.x >= 0) & slab_test(r, nd.lox.x, nd.hix.x, nd.loy.x, nd.hiy.x, nd.loz.x, nd.hiz.x, best_t, &e0);
      const bool h1 = (nd.cnt.y >= 0) & slab_test(r, nd.lox.y, nd.hix.y, nd.loy.y, nd.hiy.y, nd.loz.y, nd.hiz.y, best_t, &e1);
      const bool h2 = (nd.cnt.z >= 0) & slab_test(r, nd.lox.z, nd.hix.z, nd.loy.z, nd.hiy.z, nd.loz.z, nd.hiz.z, best_t, &e2);
      const bool h3 = (nd.cnt.w >= 0) & slab_test(r, nd.lox.w, nd.hix.w, nd.loy.w, nd.hiy.w, nd.loz.w, nd.hiz.w, best_t, &e3);
      float k0 = h0 ? e0 : kInf, k1 = h1 ? e1 : kInf, k2 = h2 ? e2 : kInf, k3 = h3 ? e3 : kInf;
      int c0 = h0 ? nd.cnt.x : -1, c1 = h1 ? nd.cnt.y : -1, c2 = h2 ? nd.cnt.z : -1, c3 = h3 ? nd.cnt.w : -1;
      int32_t r0 = nd.ref.x, r1 = nd.ref.y, r2 = nd.ref.z, r3 = nd.ref.w;
      cswap(k0, r0, c0, k1, r1, c1);
      cswap(k2, r2, c2, k3, r3, c3);
      cswap(k0, r0, c0, k2, r2, c2);
      cswap(k1, r1, c1, k3, r3, c3);
      cswap(k1, r1, c1, k2, r2, c2);
      // farthest first: the nearest inner child becomes the next node and
      // the nearest leaf the pending one (or, with a leaf pending, the
      // nearest stacked entry); the rest are stacked
      int2 next = make_int2(-1, 0), leaf = make_int2(0, 0);  // leaf.x < 0: ~slot of a leaf
      child(st, next, leaf, k3, r3, c3);
      child(st, next, leaf, k2, r2, c2);
      child(st, next, leaf, k1, r1, c1);
      child(st, next, leaf, k0, r0, c0);
      if (leaf.x < 0) {
        if (tri < 0)
          tri = ~leaf.x;
        else
          st.push(leaf);
      }
      cur = next.x;
    }
    return false;
  }

  __device__ __forceinline__ void child(HybridStack<SL>& st, int2& next, int2& leaf, float k, int32_t ref,
                                        int cnt) const {
    if (cnt < 0 || k > best_t) return;
    if (cnt > 0) {
      if (leaf.x < 0) st.push(leaf);
      leaf = make_int2(~ref, __builtin_bit_cast(int32_t, k));
    } else {
      if (next.x >= 0) st.push(next);
      next = make_int2(ref, __builtin_bit_cast(int32_t, k));
    }
  }

  __device__ __forceinline__ void result(THit* out) const {
    out->gid = ANY ? (found ? 0u : PYRE_MISS) : (found ? best_gid : PYRE_MISS);
    out->t = found ? best_t : 0.0f;
    out->u = bu;
    out->v = bv;
  }
};

// Whole traversal of one ray.
template <bool ANY, int SL>
__device__ __forceinline__ void traverse4(const Node4* __restrict__ nodes, const float4* __restrict__ tris,
                                          const TRay& r, HybridStack<SL>& st, THit* out) {
  Trav4<ANY, SL> tv;
  tv.start(r, st);
  while (!tv.step(nodes, tris, st)) {
  }
  tv.result(out);
}

// Lanes of a wave refill once at least this many are idle.
constexpr int kRefillLanes = 32;
// Items each wave takes without any atomic (its static share), and items a
// wave claims from the shared counter at a time once that is used up: large
// chunks first, small ones over the last stretch so the waves run dry
// together.
constexpr uint32_t kStaticChunk = 64;

// Work distribution of the persistent kernels over n items: wave w first
// takes a static share [w * share, (w + 1) * share) with no atomic (share =
// half of an even split, at least kStaticChunk), so a short queue costs no
// atomics at all and a long one few. The items past the static shares are
// claimed a chunk at a time from one counter of chunk numbers, which
// balances expensive and cheap rays across the GPU: CHUNK items per claim up
// to the last 2 * waves * TAIL items, TAIL items per claim after that
// (guided self-scheduling in two sizes; one counter keeps the claims few).
// A wave that sees the counter pass the end stops claiming. All members are
// wave-uniform.
template <uint32_t CHUNK, uint32_t TAIL>
struct WaveFeed {
  uint32_t n, next, end, dyn0, tail0, n_head;
  uint32_t* ctr;
  bool dry;
  __device__ __forceinline__ WaveFeed(uint32_t n_, uint32_t* counter) {
    n = n_;
    ctr = counter;
    const uint32_t waves_per_block = blockDim.x / 64u;
    const uint64_t wave = (uint64_t)blockIdx.x * waves_per_block + (threadIdx.x >> 6);
    const uint64_t waves = (uint64_t)gridDim.x * waves_per_block;
    // static share: half of an even split (at least kStaticChunk), the rest
    // is claimed dynamically
    const uint64_t even = (uint64_t)n / waves / 2;
    const uint64_t share = even > kStaticChunk ? (even & ~63ull) : kStaticChunk;
    const uint64_t s0 = wave * share, d0 = waves * share;
    next = s0 < n ? (uint32_t)s0 : n;
    end = n - next > share ? (uint32_t)(next + share) : n;
    dyn0 = d0 < n ? (uint32_t)d0 : n;
    dry = dyn0 >= n;
    // dynamic range [dyn0, n): large chunks over [dyn0, tail0), small ones
    // over [tail0, n)
    const uint64_t tail = 2 * waves * TAIL;
    const uint64_t head = n - dyn0 > tail ? n - dyn0 - tail : 0;
    tail0 = dyn0 + (uint32_t)head;
    n_head = (uint32_t)((head + CHUNK - 1) / CHUNK);
  }
  // Gives the idle lanes (ballot `idle`) the next items; true for an idle
  // lane that got one (in `item`). Called by the whole wave.
  __device__ __forceinline__ bool take(uint64_t idle, bool is_idle, uint32_t& item) {
    if (next >= end) {
      if (dry) return false;
      uint32_t c = 0;
      if (__lane_id() == 0) c = atomicAdd(ctr, 1u);
      c = __builtin_amdgcn_readfirstlane(c);
      uint64_t s, e;
      if (c < n_head) {
        s = dyn0 + (uint64_t)c * CHUNK;
        e = s + CHUNK < tail0 ? s + CHUNK : tail0;
      } else {
        s = tail0 + (uint64_t)(c - n_head) * TAIL;
        e = s + TAIL < n ? s + TAIL : n;
      }
      if (s >= n) {
        dry = true;
        return false;
      }
      next = (uint32_t)s;
      end = (uint32_t)e;
      dry = end >= n;
    }
    const uint32_t k = next + (uint32_t)__popcll(idle & ((1ull << __lane_id()) - 1ull));
    const uint32_t cnt = (uint32_t)__popcll(idle);
    next = end - next > cnt ? next + cnt : end;
    if (is_idle && k < end) {
      item = k;
      return true;
    }
    return false;
  }
};

}  // namespace pyre
