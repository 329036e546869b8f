/*
 * pyre_gen.h -- host-side scene and workload builders of the pyre_hip
 * library (no device work). They produce the inputs the C ABI in
 * pyre_hip.h consumes.
 *
 * Restated from the reference CLI:
 *   pyre_preset_cornell    cornell_box()       crates/pyre-cli/src/main.rs:365-496
 *                          quad_mesh()         main.rs:353-360
 *                          make_uv_sphere()    main.rs:565-599
 *   pyre_auto_frame        auto_frame_params() main.rs:337-351
 * Synthetic workloads with no reference counterpart (the reference has no
 * benchmark inputs); all randomness is xoshiro256++ seeded with SplitMix64,
 * transcendental/Gaussian draws are evaluated in double and rounded to f32:
 *   pyre_gen_random_triangles / pyre_gen_random_rays   1M-triangle traversal
 *   pyre_preset_production                             instanced icospheres +
 *                                                       procedural sun/sky HDRI
 */
#ifndef PYRE_GEN_H
#define PYRE_GEN_H

#include "pyre_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Triangle soup: centroid c ~ U[-1,1]^3, each vertex = c + N(0, sigma^2) per
 * axis. verts = 9 floats per triangle (v0, v1, v2). */
int pyre_gen_random_triangles(uint64_t seed, uint32_t n, float sigma, float* verts);

/* Rays with origins ~ U[-1.5,1.5]^3 and directions uniform on the sphere
 * (z = 1 - 2u, phi = 2 pi v), normalised like Ray::new (math.rs:16-24),
 * t_min = 1e-4. Rays [0, n_closest) have t_max = inf; the rest have
 * t_max ~ U[0.1, 3]. Either output may be NULL. */
int pyre_gen_random_rays(uint64_t seed, uint32_t n, uint32_t n_closest, pyre_ray* rays, pyre_ray32* rays32);

typedef struct pyre_preset pyre_preset;

/* The built-in Cornell box (motion_blur must be 0: the device path has no
 * moving instances). */
int pyre_preset_cornell(int motion_blur, pyre_preset** out);

/* Production scene: n_instances level-`level` icospheres with random
 * similarity transforms (scale U[0.2,1], uniform rotation, translation
 * U[-10,10]^3) baked to world space, per-instance Disney materials
 * (base U[0,1]^3, roughness U[0.05,1], metallic Bernoulli(0.3), specular
 * 0.5), a Lambertian 0.55 ground quad under them, and an env_w x env_h
 * sky-gradient HDRI with a Gaussian sun lobe (sigma 0.5 deg, peak 1e4 x the
 * zenith luminance). */
int pyre_preset_production(uint64_t seed, uint32_t n_instances, uint32_t level, uint32_t env_w, uint32_t env_h,
                           pyre_preset** out);
/* The same scene family with a Lambertian material (albedo = the drawn base
 * colour) on a lambert_fraction share of the instances; production is
 * lambert_fraction = 0. */
int pyre_preset_instances(uint64_t seed, uint32_t n_instances, uint32_t level, uint32_t env_w, uint32_t env_h,
                          float lambert_fraction, pyre_preset** out);

/* Points desc at the preset's arrays (valid until pyre_preset_destroy). */
int pyre_preset_desc(const pyre_preset* p, pyre_scene_desc* desc);
/* Camera the reference CLI would use for the preset at this aspect ratio:
 * Cornell (0,0,3) -> 0, vfov 45 (main.rs:199-201); production:
 * auto_frame_params over the instance bounds, vfov 45 (main.rs:186-194). */
int pyre_preset_camera(const pyre_preset* p, float aspect, float aperture, pyre_camera_desc* cam);
void pyre_preset_destroy(pyre_preset* p);

/* auto_frame_params (main.rs:337-351). */
void pyre_auto_frame(const float bmin[3], const float bmax[3], float vfov_deg, float aspect, float origin[3],
                     float target[3]);

#ifdef __cplusplus
}
#endif

#endif /* PYRE_GEN_H */
