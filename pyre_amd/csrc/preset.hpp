// Host-side scene container behind the pyre_preset handle of pyre_gen.h:
// built-in presets (gen.cpp) and scenes read from files (scene_io.cpp).
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/pyre_gen.h"

// One primitive: world-space positions, optional per-vertex normals (empty =
// face normals), indices, the material id and an optional translation over
// the shutter.
struct PresetMesh {
  std::vector<float> pos, nrm;
  std::vector<uint32_t> idx;
  uint32_t material = 0;
  std::vector<float> motion;  // empty, or InstanceMotion {offset_t0, offset_t1}
};

// An emissive triangle mesh (pyre_mesh_light).
struct PresetMeshLight {
  std::vector<float> pos;
  std::vector<uint32_t> idx;
  float emission[3] = {0, 0, 0};
};

enum PresetKind : int {
  PRESET_CORNELL = 0,    // the CLI's fixed Cornell camera
  PRESET_INSTANCES = 1,  // auto-framed over the instance bounds
  PRESET_LOADED = 2,     // a scene file: its own camera, else auto-framed over the scene bounds
  PRESET_STUDIO = 3,     // the CLI's fixed studio camera
};

struct pyre_preset {
  std::vector<PresetMesh> meshes;
  std::vector<pyre_material> materials;
  std::vector<pyre_quad_light> lights;
  std::vector<PresetMeshLight> mesh_lights;
  std::vector<pyre_mesh_light> mldesc;
  std::vector<float> env_rgb;
  uint32_t env_w = 0, env_h = 0;
  float env_intensity = 1.0f;
  std::vector<pyre_mesh_desc> mdesc;
  pyre_env_desc edesc{};
  int kind = PRESET_INSTANCES;
  float frame_min[3] = {0, 0, 0}, frame_max[3] = {0, 0, 0};  // bounds the auto-frame camera uses
  pyre_view view{};                                          // what a scene file says about the render
};

// Points mdesc / edesc at the arrays (after every change to them).
void preset_finish(pyre_preset* p);
// frame_min / frame_max = union of the meshes' vertex bounds (Scene::bounds,
// scene.rs:146-152); empty bounds when there is no vertex.
void preset_scene_bounds(pyre_preset* p);
