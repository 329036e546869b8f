// Internal types of the gfx950 backend (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pyre_hip.h"
#include "pmath.hpp"

struct pyre_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::string name;
  int cus = 0;
  pyre_tuning tune{};  // launch settings (0 fields = defaults)
  // traversal scratch of the trace entry points: work counter, stack-overflow
  // flag, spill columns for the largest persistent grid (8 blocks per CU) on
  // trees up to trace_depth; grown by pyre_bvh_build, never by a trace call.
  // trace_ev orders calls that share it across streams.
  uint32_t* trace_fetch = nullptr;
  uint32_t* trace_flag = nullptr;
  int2* trace_gstack = nullptr;
  uint32_t trace_depth = 0;
  hipEvent_t trace_ev = nullptr;
  // the device address of the no-allocation trace scratch (a fixed array in
  // trace.hip) on the context's device
  void* noalloc_scratch = nullptr;
  // path buffers kept from the last one-shot render (owned; freed with the
  // context, on its device)
  std::shared_ptr<void> render_cache;
  // every job's film sums of pyre_render_batch (10 floats and 3 doubles per
  // pixel of the batch), kept for the next batch and grown, never shrunk
  float* batch_acc = nullptr;
  double* batch_acc64 = nullptr;
  size_t batch_npix = 0;
  // the host libm's sinf/cosf on the BSDF samplers' angle grid (PhiTable in
  // shade.hpp), uploaded at context creation
  uint4* d_phi_rank = nullptr;
  float2* d_phi_val = nullptr;
};

namespace pyre {

// Failure carried to the C ABI boundary as a status code.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PYRE_HIP_CHECK(expr)                                                                        \
  do {                                                                                              \
    hipError_t e_ = (expr);                                                                         \
    if (e_ != hipSuccess)                                                                           \
      throw ::pyre::Error(e_ == hipErrorOutOfMemory ? PYRE_ENOMEM : PYRE_EHIP,                      \
                          std::string(#expr) + ": " + hipGetErrorString(e_));                      \
  } while (0)

inline void require(bool ok, const std::string& msg) {
  if (!ok) throw Error(PYRE_EINVAL, msg);
}

template <class T>
T* dev_alloc(size_t count) {
  T* p = nullptr;
  if (count == 0) count = 1;
  PYRE_HIP_CHECK(hipMalloc((void**)&p, count * sizeof(T)));
  return p;
}
template <class T>
void dev_free(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

// Traversal node: both children's boxes live in the parent so one 64-byte
// fetch (four dwordx4 loads) tests two boxes.
//   x = (c0.lo.x, c0.hi.x, c1.lo.x, c1.hi.x), y and z likewise,
//   meta = (c0 ref, c1 ref, c0 count, c1 count):
//   count == 0 -> ref is a node index; count > 0 -> ref is the first slot of a
//   leaf in the triangle array; ref < 0 -> no child.
struct alignas(16) Node2 {
  float4 x, y, z;
  int4 meta;
};

// 4-wide traversal node, 128 bytes (one L2 line): child boxes as
// per-axis lanes, child refs and counts as in Node2 (count -1 = no child).
struct alignas(16) Node4 {
  float4 lox, hix, loy, hiy, loz, hiz;
  int4 ref;
  int4 cnt;
};
static_assert(sizeof(Node4) == 128, "Node4 layout");

// Mark in a triangle record's e2.w: the last slot of its leaf (set by the
// build on every leaf), so a traversal can hold a pending leaf as one slot.
constexpr uint32_t kLeafEnd = 1u;
// Triangles per entry of a scene's mesh lookup table (d_mesh_blk), log2.
constexpr uint32_t kMeshBlkShift = 12;

// 4-wide node with its child boxes quantised to a scene-wide grid, 64 bytes
// (four dwordx4 loads instead of eight). Each child plane is a 16-bit cell
// index q of the grid, rounded outward (plane = QGrid.lo + q * 2^e, lo planes
// rounded down, hi planes up), two children per word:
//   x = (lo0 | lo1 << 16, lo2 | lo3 << 16, hi0 | hi1 << 16, hi2 | hi3 << 16),
//   y and z likewise;
//   ref = -1 (no child), an inner child's node index, or kQLeaf | the first
//   slot of a leaf (its last slot carries kLeafEnd).
struct alignas(16) Node4Q {
  uint4 x, y, z;
  int4 ref;
};
static_assert(sizeof(Node4Q) == 64, "Node4Q layout");
constexpr int32_t kQLeaf = 1 << 30;

// The quantisation grid of a scene's tree: per axis, lo[a] + q * 2^e[a] for
// q in [0, 65535] spans the root box.
struct QGrid {
  float lo[3];
  int e[3];
};

// Build node (binned-SAH builder). Bounds are kept as order-preserving int
// keys so they can be accumulated with integer atomics.
struct BNode {
  int32_t bkey[6];  // node bounds min xyz, max xyz
  int32_t ckey[6];  // centroid bounds min xyz, max xyz
  uint32_t start, count;
  int32_t left, right;  // child build-node ids (interior), -1 for leaves
  uint32_t axis;
  uint32_t flags;  // bit0 leaf, bit1 used
  float split_pos;
  uint32_t mode;  // decide result for the wide-node path
};
static_assert(sizeof(BNode) == 80, "BNode layout");

PH int32_t f2key(float f) {
  int32_t i = __builtin_bit_cast(int32_t, f);
  return i >= 0 ? i : (i ^ 0x7FFFFFFF);
}
PH float key2f(int32_t k) {
  int32_t i = k >= 0 ? k : (k ^ 0x7FFFFFFF);
  return __builtin_bit_cast(float, i);
}

// DiffuseAreaQuadLight with the fields its constructor derives (light.rs:61-76).
struct QuadLightDev {
  V3 p0, eu, ev, cross, normal, emission;
  float inv_cross_len_sq, area;
};

// An emissive triangle mesh (pyre_mesh_light): its triangles are tri0 ..
// tri0 + ntri - 1 of the scene's emitter arrays (4 float4 each: v0, e1, e2,
// normal), its Distribution1D cdf (ntri + 1 entries) starts at cdf0.
struct MeshLightDev {
  uint32_t tri0, ntri, cdf0;
  float area;  // the cdf's integral: the mesh's total area
  V3 emission;
  V3 bmin, bmax;  // the triangles' box, padded far beyond rounding (a cull test only)
};

struct MatDev {
  uint32_t kind;
  V3 base;
  float metallic, roughness, specular;
};

// HdriEnvironmentLight on the device (light.rs:205-211): texels as float4,
// conditional CDFs (height rows of width + 1), marginal CDF (height + 1).
// Guide tables over a CDF of n + 1 entries (cutpoint method, Chen & Asau
// 1974): guide[k], k in [0, kEnvGuide], is the first index i in [1, n) with
// cdf[i] > k / kEnvGuide, or n. A sample u in [k, k + 1) / kEnvGuide has its
// bin search confined to [guide[k], guide[k + 1]].
constexpr uint32_t kEnvGuide = 256;

struct EnvDev {
  uint32_t width = 0, height = 0;
  float intensity = 1.0f;
  const float4* texels = nullptr;
  const float* cond_cdf = nullptr;
  const float* marg_cdf = nullptr;
  const uint32_t* cond_guide = nullptr;  // kEnvGuide + 1 per row
  const uint32_t* marg_guide = nullptr;  // kEnvGuide + 1
};

}  // namespace pyre

struct pyre_scene {
  pyre_ctx* ctx = nullptr;
  uint32_t n_meshes = 0, n_tris = 0, n_materials = 0, n_lights = 0;
  std::vector<uint32_t> h_tri_offset;  // n_meshes + 1
  std::vector<uint32_t> h_mesh_info;   // material id | has_normals << 31
  std::vector<pyre::MatDev> h_mats;
  std::vector<pyre::QuadLightDev> h_lights;
  uint32_t n_mlights = 0, n_etris = 0;  // emissive meshes and their triangles
  std::vector<pyre::MeshLightDev> h_mlights;
  pyre::MeshLightDev* d_mlights = nullptr;
  float4* d_etris = nullptr;  // 4 per emitter triangle
  float* d_ecdf = nullptr;    // the meshes' cdfs, one after another
  float* d_verts = nullptr;    // 9 floats per scene triangle (v0, v1, v2)
  float* d_normals = nullptr;  // 9 floats per triangle (vertex normals) or null
  bool has_motion = false;     // some mesh moves over the shutter
  std::vector<float4> h_motion;  // 2 per mesh: (offset_t0, 0), (offset_t1, 0)
  std::vector<uint32_t> h_moving;  // per mesh: 1 = moving
  float4* d_motion = nullptr;    // h_motion on the device (null for static scenes)
  uint32_t* d_tri_offset = nullptr;
  uint32_t* d_mesh_blk = nullptr;  // per block of 2^kMeshBlkShift triangles: the last mesh starting at or before it
  uint32_t* d_mesh_info = nullptr;
  pyre::MatDev* d_mats = nullptr;
  pyre::QuadLightDev* d_lights = nullptr;
  bool has_env = false;
  pyre::EnvDev env;
  float4* d_env_texels = nullptr;
  float* d_env_cond = nullptr;
  float* d_env_marg = nullptr;
  uint32_t* d_env_guide = nullptr;  // conditional guides (rows x (kEnvGuide + 1)), then the marginal's
  float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};

  // BVH
  bool built = false;
  pyre::Node2* d_nodes = nullptr;
  uint32_t n_nodes2 = 0;
  pyre::Node4* d_nodes4 = nullptr;
  uint32_t n_nodes4 = 0, depth4 = 0;
  pyre::Node4Q* d_nodes4q = nullptr;  // the same tree with quantised child boxes
  pyre::QGrid qgrid{};
  float4* d_tris = nullptr;        // 3 float4 per slot: (v0, gid bits), (e1, 0), (e2, 0)
  uint32_t* d_prim_order = nullptr;  // scene-global triangle id per slot
  pyre::BNode* d_bnodes = nullptr;
  uint32_t n_bnodes = 0;
  uint32_t n_leaves = 0, n_interior = 0, depth = 0;
  double build_ms = 0.0;
  uint64_t last_extension_rays = 0, last_shadow_rays = 0;  // of the last pyre_render
};

namespace pyre {
// bvh_build.hip
void bvh_build(pyre_scene* s, const pyre_bvh_params& p, hipStream_t st);
void bvh_release(pyre_scene* s);
// trace.hip
// check = synchronise and turn a stack overflow into PYRE_EOVERFLOW. times:
// per-ray shutter times for scenes with moving meshes (null: time 0).
void trace_closest(pyre_scene* s, const float4* rays, size_t n, pyre_hit8* hits, float2* uv, hipStream_t st,
                   bool check, const float* times = nullptr);
void trace_any(pyre_scene* s, const float4* rays, size_t n, uint32_t* occluded, hipStream_t st, bool check,
               const float* times = nullptr);
// The context's trace scratch sized for trees of binary depth `depth`
// (synchronises the device when it has to grow; called by pyre_bvh_build).
void trace_scratch_reserve(pyre_ctx* ctx, uint32_t depth);
void trace_scratch_release(pyre_ctx* ctx);
// pyre_trace_closest_dev_noalloc: the same kernel on the fixed scratch array.
void noalloc_scratch_init(pyre_ctx* ctx);
void trace_closest_noalloc(pyre_scene* s, const float4* rays, size_t n, pyre_hit8* hits, hipStream_t st);

// Resolved tuning values.
inline uint32_t tune_blocks_per_cu(const pyre_ctx* c) {
  const uint32_t v = c->tune.trace_blocks_per_cu;
  return v == 0 || v > 8 ? 8u : v;
}
inline uint32_t tune_ring(const pyre_ctx* c) {
  const uint32_t v = c->tune.stack_ring;
  return v == 4 || v == 16 ? v : 8u;
}
inline float tune_memory_fraction(const pyre_ctx* c) {
  const float f = c->tune.memory_fraction;
  return f > 0.0f && f <= 1.0f ? f : 0.5f;
}
}  // namespace pyre
