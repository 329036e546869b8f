/*
 * pyre_gen.h -- host-side scene and workload builders of the pyre_hip
 * library (no device work). They produce the inputs the C ABI in
 * pyre_hip.h consumes.
 *
 * Restated from the reference CLI:
 *   pyre_preset_cornell    cornell_box()       crates/pyre-cli/src/main.rs:365-496
 *   pyre_preset_studio     studio_scene()      main.rs:502-562 (+ the default sky, main.rs:167-174)
 *                          quad_mesh()         main.rs:353-360
 *                          make_uv_sphere()    main.rs:565-599
 *   pyre_auto_frame        auto_frame_params() main.rs:337-351
 * Scene files, as the reference CLI reads them (main.rs:114-166):
 *   pyre_preset_load_pbrt  load_pbrt()         crates/pyre/src/io/pbrt/{mod,lexer,parser,builder}.rs
 *   pyre_preset_load_gltf  load_gltf()         crates/pyre/src/io/gltf.rs:20-92, main.rs:128-141
 *   pyre_load_hdr          load_hdri()         crates/pyre/src/io/hdr.rs:24-50
 * Film sinks: pyre_film_pack, pyre_write_png, pyre_write_exr -- Film::save_png,
 * save_aov_pngs, save_exr (crates/pyre/src/film.rs:117-264).
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

/* The built-in Cornell box; motion_blur != 0 animates the gold sphere from
 * offset (0, 0, 0) to (0, 0.25, 0) over the shutter (main.rs:455-465). */
int pyre_preset_cornell(int motion_blur, pyre_preset** out);

/* The CLI's studio preset (studio_scene, main.rs:502-562): a 100 x 100
 * Lambertian 0.55 ground quad and the gold / plastic spheres (64 x 32);
 * motion_blur != 0 moves the gold sphere from x - 0.15 to x + 0.15. Its sky
 * is the CLI's default for the studio without an HDRI (main.rs:167-174):
 * HdriEnvironmentLight::gradient (light.rs:252-272, 128 x 64) at
 * env_intensity; pyre_preset_set_env_hdr replaces it like --env. Camera
 * (1.6, 0.5, 2.6) -> (0, 0.05, 0), vfov 38 (main.rs:195-198). */
int pyre_preset_studio(int motion_blur, float env_intensity, pyre_preset** out);

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

/* What a scene file says about the render (io/pbrt/builder.rs:27-43):
 * LookAt + Camera "perspective" fov (both needed for a camera) and Film
 * xresolution / yresolution (defaults 800 x 600). Presets have neither. */
typedef struct pyre_view {
  int has_camera;
  float origin[3];
  float target[3];
  float up[3];
  float vfov_deg;
  int has_film;
  uint32_t width;
  uint32_t height;
} pyre_view;

/* load_pbrt (io/pbrt/mod.rs:56-76): the reference's PBRT subset -- LookAt,
 * Camera, Film; Identity / Translate / Scale / Rotate / Transform /
 * ConcatTransform with Attribute and Transform blocks; Material,
 * MakeNamedMaterial, NamedMaterial (matte -> Lambertian, plastic / metal /
 * conductor -> Disney, others -> grey Lambertian); AreaLightSource "diffuse"
 * on a four-vertex two-triangle trianglemesh -> quad light; LightSource
 * "infinite" (constant L or a mapname .hdr relative to the file);
 * Shape "sphere" (48 x 24 uv sphere) and "trianglemesh", baked to world
 * space. Other directives are skipped, like the reference's. Errors: PYRE_EIO
 * when a file cannot be read; PYRE_EINVAL for lex / parse / build / decode
 * errors, with the reference's message ("lex: ...", "parse: ...",
 * "build: ...") in err (err may be NULL). The camera of pyre_preset_camera
 * is the file's LookAt/fov when it has both, else auto_frame_params over the
 * scene bounds (main.rs:184-194). */
int pyre_preset_load_pbrt(const char* path, pyre_preset** out, char* err, size_t err_len);

/* pyre_preset_load_pbrt with flags. PYRE_PBRT_MESH_EMITTERS: an
 * AreaLightSource on a shape that is not a four-vertex two-triangle
 * trianglemesh (which stays a quad light) makes the shape an emissive mesh
 * light (pyre_mesh_light) instead of the non-emissive primitive the
 * reference's builder makes of it (io/pbrt/builder.rs:463-485,513-516).
 * flags = 0 is pyre_preset_load_pbrt; unknown flags give PYRE_EINVAL. */
#define PYRE_PBRT_MESH_EMITTERS 1u
int pyre_preset_load_pbrt_ex(const char* path, uint32_t flags, pyre_preset** out, char* err, size_t err_len);

/* main.rs:128-141 over load_gltf (io/gltf.rs:20-92): every triangle
 * primitive of the default scene (else the first scene), node transforms
 * accumulated parent * local and baked into the positions, normals by the
 * inverse transpose of the upper 3x3 and normalised (zero stays zero), one
 * Lambertian 0.73 material. .gltf with .bin or base64 data-URI buffers and
 * binary .glb. Non-triangle modes and primitives without positions are
 * PYRE_EINVAL, like the reference's GltfError. */
int pyre_preset_load_gltf(const char* path, pyre_preset** out, char* err, size_t err_len);

/* load_hdri (io/hdr.rs:24-50) for .hdr / .pic: Radiance RGBE, "-Y H +X W"
 * scanlines (flat, old run-length or new run-length), top row first, each
 * pixel R * 2^(E - 136) (E = 0: black), as the decoder the reference calls
 * (crate image 0.25.9, hdr::Rgbe8Pixel::to_hdr). Call with rgb = NULL for
 * the size; then rgb must hold 3 * width * height floats (rgb_len). */
int pyre_load_hdr(const char* path, uint32_t* width, uint32_t* height, float* rgb, size_t rgb_len, char* err,
                  size_t err_len);

/* The CLI's --env (main.rs:162-166): replaces the preset's environment by
 * the HDRI file at this intensity. */
int pyre_preset_set_env_hdr(pyre_preset* p, const char* path, float intensity, char* err, size_t err_len);

/* The file's render settings (zeroed for presets). */
int pyre_preset_view(const pyre_preset* p, pyre_view* view);

/* Packs a host film (PixelSample means, row-major from the top-left) for
 * the sinks, on the host. */
int pyre_film_pack(const pyre_pixel_sample* px, uint32_t width, uint32_t height, uint32_t what, void* out);
/* An 8-bit grey (channels 1) or RGB (channels 3) PNG of packed bytes. */
int pyre_write_png(const char* path, uint32_t width, uint32_t height, uint32_t channels, const uint8_t* px,
                   char* err, size_t err_len);
/* A single-part scanline OpenEXR of PYRE_PACK_EXR planes: R, G, B and the AOV
 * groups of aov_mask, FLOAT channels in save_exr's sorted order, stored
 * uncompressed (the reference writes the same channels RLE-compressed). */
int pyre_write_exr(const char* path, uint32_t width, uint32_t height, uint32_t aov_mask, const float* planes,
                   char* err, size_t err_len);

/* auto_frame_params (main.rs:337-351). */
void pyre_auto_frame(const float bmin[3], const float bmax[3], float vfov_deg, float aspect, float origin[3],
                     float target[3]);

#ifdef __cplusplus
}
#endif

#endif /* PYRE_GEN_H */
