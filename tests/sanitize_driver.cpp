// Test driver (CPU only): the host-side scene loaders and film sinks of
// include/pyre_gen.h, built with AddressSanitizer and UBSan by
// tests/test_host_sanitizers.py and run over valid, crafted and mutated
// scene files. Usage: sanitize_driver <out_dir> (<pbrt|gltf|hdr> <path>)...
// Prints one line per file; any sanitizer report aborts the run.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pyre_gen.h"

static double touch_preset(pyre_preset* p) {
  // read every array the preset hands out, so out-of-bounds contents surface
  pyre_scene_desc d{};
  double sum = 0.0;
  if (pyre_preset_desc(p, &d) != 0) return sum;
  for (uint32_t m = 0; m < d.n_meshes; ++m) {
    const pyre_mesh_desc& md = d.meshes[m];
    for (size_t k = 0; k < 3 * (size_t)md.n_vertices; ++k) sum += md.positions[k];
    for (size_t k = 0; k < 3 * (size_t)md.n_triangles; ++k) sum += md.indices[k];
    if (md.normals)
      for (size_t k = 0; k < 3 * (size_t)md.n_vertices; ++k) sum += md.normals[k];
    if (md.motion)
      for (int k = 0; k < 6; ++k) sum += md.motion[k];
  }
  for (uint32_t i = 0; i < d.n_materials; ++i) sum += d.materials[i].roughness;
  for (uint32_t i = 0; i < d.n_lights; ++i) sum += d.lights[i].emission[0];
  for (uint32_t i = 0; i < d.n_mesh_lights; ++i) {
    const pyre_mesh_light& ml = d.mesh_lights[i];
    for (size_t k = 0; k < 3 * (size_t)ml.n_vertices; ++k) sum += ml.positions[k];
    for (size_t k = 0; k < 3 * (size_t)ml.n_triangles; ++k) sum += ml.indices[k];
  }
  if (d.env)
    for (size_t k = 0; k < 3 * (size_t)d.env->width * d.env->height; ++k) sum += d.env->rgb[k];
  pyre_camera_desc cam{};
  pyre_preset_camera(p, 1.5f, 0.05f, &cam);
  sum += cam.origin[0];
  pyre_view v{};
  pyre_preset_view(p, &v);
  sum += v.width;
  return sum;
}

static void sinks(const std::string& dir) {
  const uint32_t w = 7, h = 5;
  std::vector<pyre_pixel_sample> px(w * h);
  const float odd[] = {0.0f, 0.5f, 1.0f, 2.0f, -1.0f, NAN, INFINITY, -INFINITY, 1e-30f, 3.4e38f};
  for (uint32_t i = 0; i < w * h; ++i) {
    float* f = &px[i].radiance[0];
    for (int k = 0; k < 10; ++k) f[k] = odd[(i * 3 + k) % 10];
  }
  std::vector<uint8_t> rgb(3 * w * h), grey(w * h);
  std::vector<float> planes(10 * w * h);
  char err[256];
  int rc = 0;
  for (uint32_t what : {PYRE_PACK_BEAUTY_PNG, PYRE_PACK_ALBEDO_PNG, PYRE_PACK_NORMAL_PNG}) {
    rc |= pyre_film_pack(px.data(), w, h, what, rgb.data());
    rc |= pyre_write_png((dir + "/s" + std::to_string(what) + ".png").c_str(), w, h, 3, rgb.data(), err, sizeof err);
  }
  rc |= pyre_film_pack(px.data(), w, h, PYRE_PACK_DEPTH_PNG, grey.data());
  rc |= pyre_write_png((dir + "/d.png").c_str(), w, h, 1, grey.data(), err, sizeof err);
  rc |= pyre_film_pack(px.data(), w, h, PYRE_PACK_EXR, planes.data());
  for (uint32_t mask = 0; mask < 8; ++mask)
    rc |= pyre_write_exr((dir + "/e" + std::to_string(mask) + ".exr").c_str(), w, h, mask, planes.data(), err,
                         sizeof err);
  std::printf("sinks rc=%d\n", rc);
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string dir = argv[1];
  double sum = 0.0;
  for (int i = 2; i + 1 < argc; i += 2) {
    const std::string kind = argv[i];
    const char* path = argv[i + 1];
    char err[512] = {0};
    int rc = -100;
    if (kind == "pbrt" || kind == "gltf") {
      pyre_preset* p = nullptr;
      rc = kind == "pbrt" ? pyre_preset_load_pbrt(path, &p, err, sizeof err)
                          : pyre_preset_load_gltf(path, &p, err, sizeof err);
      if (rc == 0 && p) {
        sum += touch_preset(p);
        pyre_preset_destroy(p);
      }
      if (kind == "pbrt") {  // again with area lights on any shape kept as emissive meshes
        pyre_preset* q = nullptr;
        char err2[512] = {0};
        const int rc2 = pyre_preset_load_pbrt_ex(path, PYRE_PBRT_MESH_EMITTERS, &q, err2, sizeof err2);
        if (rc2 != rc) {
          std::fprintf(stderr, "pyre_preset_load_pbrt_ex returned %d where pyre_preset_load_pbrt returned %d\n", rc2, rc);
          return 3;
        }
        if (rc2 == 0 && q) {
          sum += touch_preset(q);
          pyre_preset_destroy(q);
        }
      }
    } else if (kind == "hdr") {
      uint32_t w = 0, hh = 0;
      rc = pyre_load_hdr(path, &w, &hh, nullptr, 0, err, sizeof err);
      if (rc == 0) {
        std::vector<float> rgb(3 * (size_t)w * hh);
        rc = pyre_load_hdr(path, &w, &hh, rgb.data(), rgb.size(), err, sizeof err);
        for (float f : rgb) sum += f;
        // a buffer one float short is refused, not overrun
        if (!rgb.empty()) {
          std::vector<float> small(rgb.size() - 1);
          if (pyre_load_hdr(path, &w, &hh, small.data(), small.size(), err, sizeof err) == 0) rc = -99;
        }
      }
    }
    std::printf("%s %s rc=%d\n", kind.c_str(), path, rc);
  }
  sinks(dir);
  std::printf("done %g\n", std::isfinite(sum) ? 0.0 : 1.0);
  return 0;
}
