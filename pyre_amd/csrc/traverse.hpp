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

#include "internal.hpp"

namespace pyre {

struct TRay {
  V3 o, d, inv;
  float t_min, t_max;
};

PH TRay make_tray(V3 o, V3 d, float t_min, float t_max) {
  TRay r;
  r.o = o;
  r.d = d;
  r.t_min = t_min;
  r.t_max = t_max;
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

}  // namespace pyre
