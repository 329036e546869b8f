"""PBRT area lights on shapes that are not quads (io/pbrt/builder.rs:463-485,
513-516): dropped to plain primitives by default, as the reference does;
emissive mesh lights with PYRE_PBRT_MESH_EMITTERS."""
import numpy as np
import pytest

import pyre_amd

SCENE = """
LookAt 0 1 4  0 1 0  0 1 0
Camera "perspective" "float fov" [40]
WorldBegin
Material "diffuse" "rgb reflectance" [0.6 0.6 0.6]
Shape "trianglemesh" "integer indices" [0 1 2 0 2 3] "point3 P" [-2 0 -2 2 0 -2 2 0 2 -2 0 2]
AttributeBegin
  AreaLightSource "diffuse" "rgb L" [4 3 2]
  Translate 0.5 1.5 0
  Shape "sphere" "float radius" [0.25]
AttributeEnd
AttributeBegin
  AreaLightSource "diffuse" "rgb L" [1 2 3]
  Shape "trianglemesh" "integer indices" [0 1 2] "point3 P" [-1 2 -1 1 2 -1 0 2 1]
AttributeEnd
AttributeBegin
  AreaLightSource "diffuse" "rgb L" [5 5 5]
  Shape "trianglemesh" "integer indices" [0 1 2 0 2 3] "point3 P" [-0.5 2.9 -0.5 0.5 2.9 -0.5 0.5 2.9 0.5 -0.5 2.9 0.5]
AttributeEnd
"""


@pytest.fixture
def scene_file(tmp_path):
    p = tmp_path / "emitters.pbrt"
    p.write_text(SCENE)
    return str(p)


def test_reference_behaviour_by_default(scene_file):
    a = pyre_amd.Preset.from_pbrt(scene_file).arrays()
    # floor + the sphere + the emissive triangle as plain primitives; the quad is a light
    assert len(a["meshes"]) == 3
    assert len(a["lights"]) == 1 and a["mesh_lights"] == []


def test_mesh_emitters_flag(scene_file):
    a = pyre_amd.Preset.from_pbrt(scene_file).arrays()
    b = pyre_amd.Preset.from_pbrt(scene_file, mesh_emitters=True).arrays()
    assert len(b["meshes"]) == 1 and len(b["lights"]) == 1
    assert np.array_equal(b["lights"][0], a["lights"][0])
    assert len(b["mesh_lights"]) == 2
    sph, tri = b["mesh_lights"]
    # the same baked vertices as the default load's primitives, bit for bit
    assert np.array_equal(sph["positions"].view(np.uint32), a["meshes"][1]["positions"].view(np.uint32))
    assert np.array_equal(sph["indices"], a["meshes"][1]["indices"])
    assert np.array_equal(tri["positions"].view(np.uint32), a["meshes"][2]["positions"].view(np.uint32))
    assert sph["emission"] == pytest.approx((4, 3, 2)) and tri["emission"] == pytest.approx((1, 2, 3))


def test_unknown_flag_refused(scene_file):
    import ctypes as C
    import os
    from pyre_amd import _lib as L
    lib = L.load()
    h = C.c_void_p()
    assert lib.pyre_preset_load_pbrt_ex(os.fsencode(scene_file), 2, C.byref(h), None, 0) == L.PYRE_EINVAL
