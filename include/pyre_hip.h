/*
 * pyre_hip.h -- C ABI of the MI355X (gfx950) path-tracing backend for Pyre.
 *
 * The backend replaces the body of the reference CLI's render closure
 * (crates/pyre-cli/src/main.rs:253-287) and the hot-path traits it calls:
 *   Shape::intersect          crates/pyre/src/geometry/mod.rs:19-21
 *   Bvh::build / intersect    crates/pyre/src/geometry/bvh.rs:45-64, 86-143
 *   Scene::intersect/occluded crates/pyre/src/scene.rs:57-140
 *   Bsdf eval/pdf/sample      crates/pyre/src/material.rs:23-36
 *   Light / EnvironmentLight  crates/pyre/src/light.rs:36-42, 185-193
 *   PathIntegrator::integrate crates/pyre/src/integrator.rs:57-231
 *   Film::render              crates/pyre/src/film.rs:96-109
 *
 * Conventions
 *   - Every entry point returns PYRE_OK (0) or a negative PYRE_E* status and
 *     never unwinds; pyre_last_error() returns the text of the last failure.
 *     The reference panics where this library returns a status
 *     (bvh.rs:46 length assert, bvh.rs:99 stack overflow,
 *     distribution.rs:131 and light.rs:215 buffer-size asserts).
 *   - Host pointers are owned by the caller and copied in/out. Entry points
 *     whose name ends in _dev take device pointers (memory on the context's
 *     device) and enqueue on the given HIP stream (NULL = the context stream)
 *     without synchronising and without allocating.
 *   - A context is used by one host thread at a time. A scene is read-only
 *     after upload + build. The trace entry points of a context share one
 *     traversal scratch (work counter, stack spill columns), sized when a BVH
 *     is built; calls enqueued on different streams are ordered on the device
 *     in the order they were made (each waits for the previous one).
 *   - Library behaviour depends on arguments only: launch settings live in
 *     the context's pyre_tuning, never in environment variables.
 *   - All arithmetic on the hot path is IEEE FP32 without contraction
 *     (-ffp-contract=off), matching the reference's x86-64 build.
 */
#ifndef PYRE_HIP_H
#define PYRE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PYRE_OK 0
#define PYRE_EINVAL (-1)    /* bad argument or descriptor */
#define PYRE_EHIP (-2)      /* HIP runtime failure */
#define PYRE_ENOMEM (-3)    /* device allocation failed */
#define PYRE_EOVERFLOW (-4) /* traversal stack or queue could not be grown */
#define PYRE_ESTATE (-5)    /* call out of order (e.g. trace before build) */
#define PYRE_EIO (-6)       /* a file could not be opened or read */

#define PYRE_MISS 0xFFFFFFFFu

typedef struct pyre_ctx pyre_ctx;
typedef struct pyre_scene pyre_scene;

/* Ray (crates/pyre/src/math.rs:3-13), 36 bytes. */
typedef struct pyre_ray {
  float o[3];
  float d[3];
  float t_min;
  float t_max;
  float time;
} pyre_ray;

/* Device ray record for the _dev trace entry points, 32 bytes:
 * {o.x, o.y, o.z, t_min, d.x, d.y, d.z, t_max}. */
typedef struct pyre_ray32 {
  float o[3];
  float t_min;
  float d[3];
  float t_max;
} pyre_ray32;

/* Closest hit. inst = primitive index in the scene (scene.rs:37-40
 * HitKind::Surface primitive_id), tri = triangle index within that
 * primitive's mesh (mesh.rs:124, dropped by the reference), t/u/v as
 * returned by moller_trumbore (mesh.rs:153-178). inst = PYRE_MISS on a miss. */
typedef struct pyre_hit {
  uint32_t inst;
  uint32_t tri;
  float t;
  float u;
  float v;
} pyre_hit;

/* Compact device hit for the _dev closest-hit entry point, 8 bytes:
 * prim = scene-global triangle id (PYRE_MISS on a miss), t. */
typedef struct pyre_hit8 {
  uint32_t prim;
  float t;
} pyre_hit8;

#define PYRE_MAT_LAMBERTIAN 0u /* material.rs:42-77 (base_color = albedo) */
#define PYRE_MAT_DISNEY 1u     /* material.rs:83-226 */

typedef struct pyre_material {
  uint32_t kind;
  float base_color[3];
  float metallic;
  float roughness;
  float specular;
} pyre_material;

/* DiffuseAreaQuadLight::new arguments (light.rs:61-76). */
typedef struct pyre_quad_light {
  float p0[3];
  float edge_u[3];
  float edge_v[3];
  float emission[3];
} pyre_quad_light;

/* One scene primitive (scene.rs:19-22): an indexed triangle mesh in world
 * space (TriangleMesh, geometry/mesh.rs:5-11) and its material id.
 * normals may be NULL (face normals are used, mesh.rs:131-136). motion may
 * be NULL (a static mesh) or point at InstanceMotion's six floats
 * {offset_t0[3], offset_t1[3]} (mesh.rs:46-60): the mesh is hit by rays
 * shifted by -lerp(offset_t0, offset_t1, clamp(ray.time, 0, 1)), and the hit
 * position is shifted back (mesh.rs:109-148). */
typedef struct pyre_mesh_desc {
  const float* positions;  /* 3 * n_vertices */
  const uint32_t* indices; /* 3 * n_triangles */
  const float* normals;    /* 3 * n_vertices or NULL */
  uint32_t n_vertices;
  uint32_t n_triangles;
  uint32_t material_id;
  const float* motion; /* 6 floats or NULL */
} pyre_mesh_desc;

/* An emissive triangle mesh as one light: the general emitter the reference's
 * PBRT builder drops (io/pbrt/builder.rs:463-485,513-516), behind the Light
 * trait (light.rs:36-42). World-space positions and indices; each triangle
 * emits `emission` along normalize(e1 x e2) (one-sided, as
 * DiffuseAreaQuadLight). NEE draws one point per mesh: a triangle by area
 * (Distribution1D, distribution.rs:39-101), then a uniform point on it; its
 * pdf is dist^2 / (cos_light * total area). Like the quads, its triangles are
 * not occluders (scene.rs:118-119) and are hit after the meshes with a strict
 * t < t_max; light ids run over the quads, then these. */
typedef struct pyre_mesh_light {
  const float* positions;  /* 3 * n_vertices */
  const uint32_t* indices; /* 3 * n_triangles */
  uint32_t n_vertices;
  uint32_t n_triangles; /* 1 .. PYRE_MESH_LIGHT_MAX_TRIS over the scene */
  float emission[3];
} pyre_mesh_light;
#define PYRE_MESH_LIGHT_MAX_TRIS 65536u

/* HdriEnvironmentLight::new arguments (light.rs:214-242): row-major RGB32F. */
typedef struct pyre_env_desc {
  uint32_t width;
  uint32_t height;
  const float* rgb; /* 3 * width * height */
  float intensity;
} pyre_env_desc;

typedef struct pyre_scene_desc {
  const pyre_mesh_desc* meshes;
  uint32_t n_meshes;
  const pyre_material* materials;
  uint32_t n_materials;
  const pyre_quad_light* lights;
  uint32_t n_lights;
  const pyre_env_desc* env; /* NULL = no environment (scene.rs:33) */
  const pyre_mesh_light* mesh_lights; /* may be NULL when n_mesh_lights = 0 */
  uint32_t n_mesh_lights;
} pyre_scene_desc;

/* Binned-SAH parameters (bvh.rs:13-16): bins (2..64), max primitives per
 * leaf, traversal and intersection costs. The reference uses 12/8/1/2. */
typedef struct pyre_bvh_params {
  uint32_t bins;
  uint32_t max_leaf;
  float traversal_cost;
  float intersect_cost;
} pyre_bvh_params;

/* Camera (camera.rs:29-143). aperture_radius > 0 selects the thin lens
 * (main.rs:206-224); focus_distance <= 0 means |target - origin|
 * (main.rs:203-205). */
typedef struct pyre_camera_desc {
  float origin[3];
  float target[3];
  float up[3];
  float vfov_deg;
  float aspect;
  float aperture_radius;
  float focus_distance;
} pyre_camera_desc;

/* PathIntegrator (integrator.rs:36-48). */
typedef struct pyre_integrator_params {
  uint32_t max_depth;
  uint32_t min_rr_depth;
} pyre_integrator_params;

/* PixelSample (integrator.rs:14-20), 40 bytes, same field order. */
typedef struct pyre_pixel_sample {
  float radiance[3];
  float albedo[3];
  float normal[3];
  float depth;
} pyre_pixel_sample;

/* ---------------------------------------------------------------- context */
int pyre_ctx_create(int device, pyre_ctx** out);
void pyre_ctx_destroy(pyre_ctx* ctx);
const char* pyre_last_error(const pyre_ctx* ctx);
/* Device name and CU count of the context's GPU (for reports). */
int pyre_ctx_device_info(pyre_ctx* ctx, char* name, size_t name_len, int* cu_count);

/* Launch configuration of a context's kernels (no counterpart in the
 * reference). Results do not depend on it: every setting gives the same
 * bits, except that the trace entry points' quantised and exact box forms
 * (traversal 0 and 4) can differ where two triangles' hits tie within
 * rounding (rays within rounding of a shared edge). A 0 field selects the
 * default; settings apply to later calls. */
typedef struct pyre_tuning {
  uint32_t pool_paths;          /* path-pool paths per render pass; 0 = 2^27. Never more than
                                   memory_fraction of the device's free memory holds */
  uint32_t trace_blocks_per_cu; /* persistent traversal blocks per CU, 1..8; 0 = 8 */
  uint32_t stack_ring;          /* LDS traversal stack entries per thread: 4, 8 or 16; 0 = 8 */
  uint32_t shade_waves;         /* 0 = shade kernel built for 4 waves per SIMD; 1 = no occupancy target */
  uint32_t traversal;           /* trace entry points: 0 = 4-wide tree with child boxes quantised to 16
                                   bits, persistent threads; 1 = binary tree, LDS stack (trees of depth
                                   <= 47); 2 = 4-wide tree, grid-stride; 3 = 4-wide tree, persistent, leaf
                                   triangles tested one per step beside the node visits; 4 = as 0 on the
                                   exact float boxes. Scenes with moving meshes take 4 for 0 */
  uint32_t eager;               /* 1 = render passes launched kernel by kernel instead of replaying a
                                   captured hipGraph */
  float memory_fraction;        /* share of the device's free memory a path pool may take, (0, 1];
                                   0 = 0.5 */
} pyre_tuning;
int pyre_ctx_set_tuning(pyre_ctx* ctx, const pyre_tuning* tuning);
int pyre_ctx_get_tuning(const pyre_ctx* ctx, pyre_tuning* tuning);

/* ------------------------------------------------------------------ scene */
/* Copies a flattened Scene (scene.rs:24-34) to the device. The triangle
 * arrays become one scene-global triangle list (primitive by primitive, in
 * order); the environment's marginal/conditional CDFs
 * (distribution.rs:39-62, 130-147) are built on the device in the
 * reference's sequential f32 order. */
int pyre_scene_upload(pyre_ctx* ctx, const pyre_scene_desc* desc, pyre_scene** out);
void pyre_scene_destroy(pyre_scene* scene);
/* Scene::bounds (scene.rs:146-152) over the uploaded geometry. */
int pyre_scene_bounds(pyre_ctx* ctx, pyre_scene* scene, float out_min[3], float out_max[3]);

/* GPU binned-SAH build over all triangles of the scene (replaces
 * Bvh::build, bvh.rs:45-64, called per mesh by MeshInstance::build,
 * mesh.rs:77-87). Split decisions follow bvh.rs:180-346; nodes are emitted
 * breadth-first. */
int pyre_bvh_build(pyre_ctx* ctx, pyre_scene* scene, const pyre_bvh_params* params);
/* The built tree: n_nodes = build-node slots (the length of
 * pyre_bvh_export's node arrays; slots the build reserved but did not use
 * are exported as unused), n_leaves, depth, and the build's wall time. */
int pyre_bvh_info(pyre_ctx* ctx, pyre_scene* scene, uint32_t* n_nodes, uint32_t* n_leaves, uint32_t* depth,
                  double* build_ms);
/* Exports the built tree for checking against the reference build:
 * per node bounds (6 floats: min, max), left child (interior) or first
 * primitive slot (leaf), right child (interior) or primitive count (leaf),
 * is_leaf, split axis; prim_order = scene-global triangle ids in leaf order. */
int pyre_bvh_export(pyre_ctx* ctx, pyre_scene* scene, float* bounds6, uint32_t* a, uint32_t* b, uint32_t* is_leaf,
                    uint32_t* axis, uint32_t* prim_order);

/* ---------------------------------------------------------------- tracing */
/* Batched Scene::intersect over the scene's triangles (scene.rs:57-94 minus
 * the light loop): closest hit with t in (t_min, t_max); exact ties in t
 * resolve to the lower primitive index like scene.rs:65. */
int pyre_intersect_closest(pyre_ctx* ctx, pyre_scene* scene, const pyre_ray* rays, size_t n, pyre_hit* hits);
/* Batched occlusion (scene.rs:102-140): occluded[i] = 1 iff any triangle has
 * a hit with t in (t_min, t_max). */
int pyre_intersect_any(pyre_ctx* ctx, pyre_scene* scene, const pyre_ray* rays, size_t n, uint32_t* occluded);
/* Device-resident forms: rays = pyre_ray32[n], hits = pyre_hit8[n] /
 * occluded = uint32_t[n]; stream = hipStream_t or NULL. pyre_ray32 has no
 * time field: moving meshes are seen at shutter time 0 (Ray::new's time,
 * math.rs:16-24); pyre_intersect_* honour pyre_ray.time. */
int pyre_trace_closest_dev(pyre_ctx* ctx, pyre_scene* scene, const void* rays, size_t n, void* hits, void* stream);
int pyre_trace_any_dev(pyre_ctx* ctx, pyre_scene* scene, const void* rays, size_t n, void* occluded, void* stream);
/* No counterpart in the reference: pyre_trace_closest_dev for callers that
 * cannot allocate device memory (code under hipGraph capture, tight serving
 * loops, harnesses that forbid hipMalloc). The same kernel and results, bit
 * for bit; its traversal scratch (work counter, stack spill columns) is a
 * fixed 64 MiB array in the library's device code, one per GPU, held until
 * the process ends, so the call allocates and frees nothing. The persistent
 * grid has as many 256-thread blocks (at most the tuning's blocks per CU) as
 * the array holds spill columns for on the scene's tree; a tree too deep for
 * one block is refused with PYRE_EINVAL. Every call on one GPU shares that
 * GPU's array: calls that overlap in time on one GPU (other streams,
 * contexts or threads) race, and the caller must order them. */
int pyre_trace_closest_dev_noalloc(pyre_ctx* ctx, pyre_scene* scene, const void* rays, size_t n, void* hits,
                                   void* stream);

/* -------------------------------------------------- shade-stage probes */
/* Bsdf::eval + Bsdf::pdf (material.rs:23-36) on local-frame directions:
 * wo, wi = 3n floats; f = 3n, pdf = n. */
int pyre_bsdf_eval_batch(pyre_ctx* ctx, const pyre_material* m, const float* wo, const float* wi, size_t n, float* f,
                         float* pdf);
/* Bsdf::sample: u = 2n floats; ok[i] = 0 where the reference returns None. */
int pyre_bsdf_sample_batch(pyre_ctx* ctx, const pyre_material* m, const float* wo, const float* u, size_t n,
                           float* wi, float* f, float* pdf, uint32_t* ok);
/* EnvironmentLight::sample / pdf / le of the scene's HDRI (light.rs:317-350). */
int pyre_env_sample_batch(pyre_ctx* ctx, pyre_scene* scene, const float* u, size_t n, float* wi, float* li,
                          float* pdf);
int pyre_env_pdf_batch(pyre_ctx* ctx, pyre_scene* scene, const float* wi, size_t n, float* pdf);
int pyre_env_le_batch(pyre_ctx* ctx, pyre_scene* scene, const float* wi, size_t n, float* le);
/* Light::sample / pdf of quad light `light` (light.rs:110-144):
 * from = 3n, u = 2n; outputs pos 3n, wi 3n, dist n, pdf n, ok n. */
int pyre_light_sample_batch(pyre_ctx* ctx, pyre_scene* scene, uint32_t light, const float* from, const float* u,
                            size_t n, float* pos, float* wi, float* dist, float* pdf, uint32_t* ok);
int pyre_light_pdf_batch(pyre_ctx* ctx, pyre_scene* scene, uint32_t light, const float* from, const float* wi,
                         size_t n, float* pdf);

/* ----------------------------------------------------------------- render */
/* Replaces main.rs:253-287: renders samples [spp_offset, spp_offset + spp)
 * of every pixel with the per-sample streams pixel_seed(x, y, s)
 * (sampler.rs:41-49) and writes the per-pixel mean (PixelSample) to out
 * (width*height, row-major, (0,0) top-left) if out != NULL. If accum64 !=
 * NULL it receives, added to its current contents, the FP64 sum of sample
 * radiance (3 doubles per pixel) for reference-quality accumulation. */
int pyre_render(pyre_ctx* ctx, pyre_scene* scene, const pyre_camera_desc* camera,
                const pyre_integrator_params* integrator, uint32_t width, uint32_t height, uint32_t spp,
                uint32_t spp_offset, pyre_pixel_sample* out, double* accum64);

/* The radiance cache's best case (no counterpart in the reference): the
 * samples of pyre_render, except that a path reaching the vertex where a
 * cache render would end it and query the cache (the first surface reached
 * after a bounce off a diffuse material: Lambertian, or Disney roughness >=
 * diffuse_roughness) continues from there on a stream of its own, numbered
 * `continuation`. Paths are identical up to that vertex for every
 * continuation, so the mean of K renders with continuations 0..K-1 is the
 * image a perfect cache would give these samples, up to the suffix variance
 * / K. accum64 (3 doubles per pixel, required) is added to. */
int pyre_render_split(pyre_ctx* ctx, pyre_scene* scene, const pyre_camera_desc* camera,
                      const pyre_integrator_params* integrator, uint32_t width, uint32_t height, uint32_t spp,
                      uint32_t spp_offset, float diffuse_roughness, uint32_t continuation, double* accum64);

/* Rays traced by the last pyre_render (or session render call) on this
 * scene: closest-hit extension rays (camera rays included) and any-hit
 * shadow rays. */
int pyre_render_stats(pyre_ctx* ctx, pyre_scene* scene, uint64_t* extension_rays, uint64_t* shadow_rays);

/* ------------------------------------------------------------- batches */
/* Many main.rs:253-287 renders in one go: the jobs (different scenes,
 * cameras, resolutions, spp) share one path pool; every pass of the wavefront
 * stages carries whole samples of several jobs (largest jobs first), and one
 * captured hipGraph per batch replays the pass. Each job's output equals
 * pyre_render of that job alone (same per-sample streams, same per-pixel
 * summation order). Scenes must belong to ctx and have a built BVH. */
typedef struct pyre_job {
  pyre_scene* scene;
  pyre_camera_desc camera;
  pyre_integrator_params integrator;
  uint32_t width, height, spp, spp_offset;
  pyre_pixel_sample* out; /* width*height means, or NULL */
  double* accum64;        /* FP64 radiance sums, added to (as pyre_render), or NULL */
} pyre_job;
/* A job's out and accum64 may be host memory or device memory on the
 * context's device (device buffers keep the results in HBM: no PCIe copy). */

/* Launch configuration; 0 fields keep the context's pyre_tuning. A batch
 * never uses more paths per pass than the sum of its jobs' width * height *
 * spp (nor fewer than one image of its largest job), nor more than the
 * tuning's share of free device memory holds. autotune != 0 first times
 * trial passes of the batch over pool sizes {2^22, 2^24, 2^26}, blocks per
 * CU {4, 8} and stack rings {4, 8, 16} and keeps the fastest. */
typedef struct pyre_batch_params {
  uint32_t pool_paths;
  uint32_t trace_blocks_per_cu;
  uint32_t stack_ring;
  uint32_t autotune;
} pyre_batch_params;

typedef struct pyre_batch_stats {
  uint64_t extension_rays, shadow_rays;
  uint32_t passes;
  float gpu_ms; /* GPU time of the passes (autotuning trials excluded) */
  uint32_t pool_paths, trace_blocks_per_cu, stack_ring; /* configuration used */
} pyre_batch_stats;

/* params and stats may be NULL. */
int pyre_render_batch(pyre_ctx* ctx, const pyre_job* jobs, size_t n, const pyre_batch_params* params,
                      pyre_batch_stats* stats);

/* ------------------------------------------------ neural radiance cache */
/* No counterpart in the reference. A 5x64 ReLU MLP on BF16 MFMA (FP32
 * accumulation, FP32 master weights, Adam) mapping a query (13 floats:
 * position xyz, outgoing direction xyz, normal xyz, roughness, albedo rgb)
 * to outgoing radiance (RGB). Positions are normalised by the given box. */
typedef struct pyre_nrc pyre_nrc;

typedef struct pyre_nrc_params {
  uint64_t seed;
  float learning_rate; /* Adam, e.g. 1e-2 */
  float beta1;         /* 0.9 */
  float beta2;         /* 0.99 */
  float epsilon;       /* 1e-15 */
  float bounds_min[3];
  float bounds_max[3];
} pyre_nrc_params;

int pyre_nrc_create(pyre_ctx* ctx, const pyre_nrc_params* params, pyre_nrc** out);
void pyre_nrc_destroy(pyre_nrc* nrc);
/* Number of FP32 parameters (layer by layer, [in][out] row-major). */
size_t pyre_nrc_param_count(void);
int pyre_nrc_get_weights(pyre_ctx* ctx, pyre_nrc* nrc, float* out);
/* Replaces the weights and resets the Adam moments. */
int pyre_nrc_set_weights(pyre_ctx* ctx, pyre_nrc* nrc, const float* weights);
/* rgb = 3n floats for feats = 13n floats. */
int pyre_nrc_infer(pyre_ctx* ctx, pyre_nrc* nrc, const float* feats, size_t n, float* rgb);
/* Loss and FP32 gradient (param_count floats) of one batch, no update: the
 * sum, in FP32 and in a fixed order, of per-CU partial gradients stored in
 * bf16 (the products' operands are bf16); the Adam step uses the same sum. */
int pyre_nrc_gradients(pyre_ctx* ctx, pyre_nrc* nrc, const float* feats, const float* targets, size_t n,
                       float* grad, float* loss);
/* One Adam step on the batch; loss receives the pre-step loss. */
int pyre_nrc_train_step(pyre_ctx* ctx, pyre_nrc* nrc, const float* feats, const float* targets, size_t n,
                        float* loss);
/* Device-resident forms (no counterpart in the reference): feats, rgb,
 * targets and loss are device arrays; the work is queued on stream
 * (hipStream_t, or NULL for the context's stream) without synchronising or
 * allocating (a training step's scratch is fixed at pyre_nrc_create). loss,
 * when not NULL, has the batch's mean loss added to it. */
int pyre_nrc_infer_dev(pyre_ctx* ctx, pyre_nrc* nrc, const float* feats, size_t n, float* rgb, void* stream);
int pyre_nrc_train_step_dev(pyre_ctx* ctx, pyre_nrc* nrc, const float* feats, const float* targets, size_t n,
                            float* loss, void* stream);
/* Sizes the staging buffers of the host-array entry points
 * (pyre_nrc_gradients, pyre_nrc_train_step) for batches of up to max_batch
 * samples now, so that no later call of at most that many allocates. */
int pyre_nrc_reserve(pyre_ctx* ctx, pyre_nrc* nrc, size_t max_batch);
/* Adam steps taken so far and the loss of the last training batch. */
int pyre_nrc_stats(pyre_ctx* ctx, pyre_nrc* nrc, uint64_t* steps, float* last_loss);

/* -------------------------------------------------- progressive session */
/* The render loop of main.rs:253-287 (and the viewer's progressive passes,
 * viewer.rs:114-184) with device-resident state: path buffers, the f32
 * PixelSample sums and FP64 radiance sums stay on the GPU between calls, and
 * each pass is a captured hipGraph replayed with new sample indices.
 * Samples are taken in order s = 0, 1, 2, ... per pixel with the streams
 * pixel_seed(x, y, s), so N spp from a session equal pyre_render(spp = N). */
typedef struct pyre_session pyre_session;

/* nrc_train_paths > 0 reserves the radiance-cache buffers (training paths
 * per frame); 0 for plain rendering only. */
int pyre_session_create(pyre_ctx* ctx, pyre_scene* scene, const pyre_camera_desc* camera,
                        const pyre_integrator_params* integrator, uint32_t width, uint32_t height,
                        uint32_t nrc_train_paths, pyre_session** out);
void pyre_session_destroy(pyre_session* session);
/* Adds spp samples per pixel (plain path tracer). A call renders its
 * samples in passes over the session's path pool, which is sized on the
 * first call (and grown by later calls that ask for more samples) to hold
 * the call's samples, within the context's pyre_tuning pool size and memory
 * share, and at least one sample per pixel. */
int pyre_session_render(pyre_ctx* ctx, pyre_session* session, uint32_t spp);
/* Sizes the session's path pool for passes of up to spp_per_pass samples per
 * pixel now, so that no later render call of at most that many samples
 * allocates (for timing, or for callers that must not allocate later). */
int pyre_session_reserve(pyre_ctx* ctx, pyre_session* session, uint32_t spp_per_pass);

/* Radiance-cache rendering (no counterpart in the reference). Per frame, in
 * one pass over the path pool: spp_per_frame samples per pixel are traced,
 * paths stopping at the first surface reached after a bounce off a diffuse
 * material (Lambertian, or Disney roughness >= diffuse_roughness); beside
 * them train_paths full unbiased paths from scattered pixels are traced, every
 * surface vertex of which yields a training sample (the query there and the
 * radiance the rest of the path carried back to it). Then up to train_steps
 * Adam steps run on consecutive batches of train_batch of the frame's
 * samples, and the stopped paths add the updated cache's radiance there.
 * Frames before warmup_frames only train (no image samples); the next
 * plain_frames frames trace their image samples to the end, as
 * pyre_session_render does (the same samples, bit for bit), while their
 * training paths train the cache; the frames after them stop paths in the
 * cache. */
typedef struct pyre_nrc_render_params {
  uint32_t frames;
  uint32_t spp_per_frame;
  uint32_t train_paths;
  uint32_t train_batch;
  uint32_t train_steps;
  float diffuse_roughness;
  uint32_t warmup_frames;
  uint32_t plain_frames;
} pyre_nrc_render_params;

/* losses: frames floats (mean training loss per frame) or NULL. */
int pyre_session_render_nrc(pyre_ctx* ctx, pyre_session* session, pyre_nrc* nrc, const pyre_nrc_render_params* p,
                            float* losses);
/* Per-pixel mean over the samples so far (PixelSample, width*height) and/or
 * the FP64 radiance sums (3 doubles per pixel), and the samples per pixel. */
int pyre_session_read(pyre_ctx* ctx, pyre_session* session, pyre_pixel_sample* mean, double* radiance_sum64,
                      uint32_t* spp_done);

/* Film sinks (film.rs:117-264). What pyre_film_pack / pyre_session_pack
 * produce: */
#define PYRE_PACK_BEAUTY_PNG 0 /* w*h*3 bytes: clamp(radiance, 0, 1)^(1/2.2) * 255, truncated (save_png) */
#define PYRE_PACK_ALBEDO_PNG 1 /* w*h*3 bytes: the same of the albedo (<stem>.albedo.png) */
#define PYRE_PACK_NORMAL_PNG 2 /* w*h*3 bytes: clamp(0.5 * n + 0.5, 0, 1) * 255 (<stem>.normal.png) */
#define PYRE_PACK_DEPTH_PNG 3  /* w*h bytes: near = bright over the depth range of the hits (<stem>.depth.png) */
#define PYRE_PACK_EXR 4        /* 10 * w*h floats: the planes B, G, N.X, N.Y, N.Z, R, Z, albedo.B, albedo.G,
                                  albedo.R (save_exr's channels in its sorted order) */
/* AOV groups of an EXR (film.rs:229-241; beauty R/G/B is always written). */
#define PYRE_AOV_ALBEDO 1u
#define PYRE_AOV_NORMAL 2u
#define PYRE_AOV_DEPTH 4u

/* The session's film (its means so far) packed on the device for the sinks
 * of pyre_gen.h: what = PYRE_PACK_*, out = host memory of that size. The
 * conversions of pyre_film_pack on pyre_session_read's means, byte for byte
 * and float for float (gamma bytes from thresholds of the host's powf). */
int pyre_session_pack(pyre_ctx* ctx, pyre_session* session, uint32_t what, void* out);
/* The viewer's displayed image (viewer.rs:114-184): the per-pixel running
 * mean of sample radiance updated sample by sample as d += (s - d) / n, in
 * f32, 3 floats per pixel. */
int pyre_session_read_viewer(pyre_ctx* ctx, pyre_session* session, float* rgb);
/* Totals over the session's lifetime: out[0] extension rays, out[1] shadow
 * rays, out[2] paths ended in the radiance cache, out[3] cache training
 * samples produced. */
int pyre_session_stats(pyre_ctx* ctx, pyre_session* session, uint64_t out[4]);
/* Clears the sums and the running mean; the sample index keeps advancing. */
int pyre_session_reset(pyre_ctx* ctx, pyre_session* session);

#ifdef __cplusplus
}
#endif

#endif /* PYRE_HIP_H */
