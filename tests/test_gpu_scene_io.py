"""Scenes read from files, rendered on the GPU: the PBRT Cornell box against
the built-in preset (the reference's round trip for scenes/cornell.pbrt),
and PBRT / glTF scenes with a Radiance environment against the oracle's
PathIntegrator on the same loaded arrays (correlated samples)."""
import base64
import json

import numpy as np
import pytest

import pyre_amd
from tests.helpers import rel_mse
from tests.test_scene_io import _cornell_pbrt, _rgbe_header

pytestmark = pytest.mark.gpu


def _ocam(oracle, cam):
    return oracle.camera(cam.origin, cam.target, cam.up, cam.vfov_deg, cam.aspect, cam.aperture_radius,
                         cam.focus_distance if cam.focus_distance > 0 else None)


def _oracle_scene(oracle, pre):
    arr = pre.arrays()
    s = oracle.Scene()
    for m in arr["materials"]:
        s.add_material(oracle.material(m.kind, m.base_color, m.metallic, m.roughness, m.specular))
    for m in arr["meshes"]:
        s.add_mesh(m["positions"], m["indices"], m["normals"], m["material_id"])
    for l in arr["lights"]:
        s.add_light(l)
    if arr["env"] is not None:
        e = arr["env"]
        s.set_env(oracle.Hdri(e["width"], e["height"], e["rgb"], e["intensity"]))
    return s


def _sky_hdr(path, w=64, h=32):
    """A smooth sky with a bright patch, as flat RGBE scanlines."""
    y, x = np.mgrid[0:h, 0:w]
    px = np.zeros((h, w, 4), np.uint8)
    px[..., 0] = 60 + 100 * y // h
    px[..., 1] = 90 + 60 * x // w
    px[..., 2] = 200
    px[..., 3] = 128
    px[2:5, 40:44, :3] = 250
    px[2:5, 40:44, 3] = 134  # about 64x brighter
    path.write_bytes(_rgbe_header(w, h) + px.tobytes())


def _assert_env_scene_parity(g, o):
    """Pinhole scenes lit by an environment map: first-hit AOVs are the
    oracle's bits; radiance agrees to 1e-5 relative (the environment's
    direction mapping uses the device's acos/atan2/sin/cos, an ulp from
    libm's)."""
    same_aov = np.all(g[..., 3:].view(np.uint32) == o[..., 3:].view(np.uint32), axis=-1).mean()
    rel = (np.abs(g[..., :3] - o[..., :3]) / np.maximum(np.abs(o[..., :3]), 1e-6)).max(-1)
    e = rel_mse(g[..., :3], o[..., :3])
    print(f"  AOV bit-equal {same_aov:.4f} radiance within 1e-5 {np.mean(rel <= 1e-5):.4f} relMSE(gpu, cpu) {e:.2e}",
          flush=True)
    assert same_aov >= 0.99
    assert np.mean(rel <= 1e-5) >= 0.99
    assert e < 1e-8


def test_pbrt_cornell_renders_like_the_preset(ctx, tmp_path):
    print("[scene io] pbrt cornell vs preset, 64x64 16 spp", flush=True)
    f = tmp_path / "cornell.pbrt"
    f.write_text(_cornell_pbrt())
    imgs = []
    for pre in (pyre_amd.Preset.from_pbrt(str(f)), pyre_amd.Preset.cornell()):
        sc = pyre_amd.Scene.from_preset(ctx, pre)
        sc.build_bvh(bins=12, max_leaf=8)
        imgs.append(pyre_amd.Film(64, 64).render(sc, pre.camera(1.0), pyre_amd.PathIntegrator(8, 3), 16)
                    .reshape(64, 64, 10).copy())
        sc.close()
    a, b = imgs
    # the scenes differ only in the spheres' normals (normalised once more,
    # one ulp): most pixels are the same bits, the rest barely differ
    same = np.all(a == b, axis=-1).mean()
    e = rel_mse(a[..., :3], b[..., :3])
    print(f"  identical pixels {same:.4f} relMSE {e:.2e}", flush=True)
    assert same > 0.8
    assert e < 1e-3


def test_pbrt_scene_with_hdr_environment_against_oracle(ctx, oracle, tmp_path):
    print("[scene io] pbrt + hdr vs oracle, 48x32 16 spp", flush=True)
    _sky_hdr(tmp_path / "sky.hdr")
    (tmp_path / "s.pbrt").write_text("\n".join([
        "LookAt 0 1.2 4  0 0.3 0  0 1 0", 'Camera "perspective" "float fov" [40]',
        'Film "image" "integer xresolution" [48] "integer yresolution" [32]',
        "WorldBegin",
        'LightSource "infinite" "string mapname" ["sky.hdr"] "float scale" [1.5]',
        'AttributeBegin Material "matte" "rgb Kd" [0.55 0.55 0.55]',
        '  Shape "trianglemesh" "point P" [-4 0 -4 4 0 -4 4 0 4 -4 0 4] "normal N" [0 1 0 0 1 0 0 1 0 0 1 0]'
        ' "integer indices" [0 2 1 0 3 2]',
        "AttributeEnd",
        'AttributeBegin Material "plastic" "rgb Kd" [0.8 0.2 0.1] "float roughness" [0.3]',
        '  Translate -0.6 0.5 0 Scale 1 1.2 1 Shape "sphere" "float radius" [0.5]', "AttributeEnd",
        'AttributeBegin Material "metal" "float roughness" [0.2] Translate 0.7 0.4 0.3 Rotate 30 0 1 0',
        '  Shape "sphere" "float radius" [0.4]', "AttributeEnd",
        "WorldEnd"]))
    pre = pyre_amd.Preset.from_pbrt(str(tmp_path / "s.pbrt"))
    w, h = pre.view["film"]
    cam = pre.camera(w / h)
    sc = pyre_amd.Scene.from_preset(ctx, pre)
    sc.build_bvh(bins=12, max_leaf=8)
    g = pyre_amd.Film(w, h).render(sc, cam, pyre_amd.PathIntegrator(8, 3), 16).reshape(h, w, 10)
    sc.close()
    o = _oracle_scene(oracle, pre).render(_ocam(oracle, cam), w, h, 16).reshape(h, w, 10)
    _assert_env_scene_parity(g, o)
    assert o[..., :3].mean() > 0.05


def test_gltf_scene_against_oracle(ctx, oracle, tmp_path):
    print("[scene io] gltf + hdr vs oracle, 40x40 16 spp", flush=True)
    # a unit cube (24 vertices with face normals) under two nodes
    faces = []
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            n = np.zeros(3)
            n[axis] = sgn
            u, v = np.eye(3)[(axis + 1) % 3], np.eye(3)[(axis + 2) % 3]
            c = n * 0.5
            quad = [c - 0.5 * u - 0.5 * v, c + 0.5 * u - 0.5 * v, c + 0.5 * u + 0.5 * v, c - 0.5 * u + 0.5 * v]
            faces.append((quad, n))
    pos = np.array([p for q, _ in faces for p in q], np.float32)
    nrm = np.array([n for _, n in faces for _ in range(4)], np.float32)
    idx = np.array([k * 4 + i for k in range(6) for i in (0, 1, 2, 0, 2, 3)], np.uint32)
    blob = pos.tobytes() + nrm.tobytes() + idx.tobytes()
    doc = {
        "asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": [0, 2]}],
        "nodes": [{"translation": [0, 0.5, 0], "rotation": [0, 0.2588190451, 0, 0.9659258263], "children": [1]},
                  {"scale": [1, 0.6, 1], "mesh": 0},
                  {"scale": [6, 0.02, 6], "translation": [0, -0.01, 0], "mesh": 0}],
        "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "NORMAL": 1}, "indices": 2}]}],
        "buffers": [{"byteLength": len(blob),
                     "uri": "data:application/octet-stream;base64," + base64.b64encode(blob).decode()}],
        "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": pos.nbytes},
                        {"buffer": 0, "byteOffset": pos.nbytes, "byteLength": nrm.nbytes},
                        {"buffer": 0, "byteOffset": pos.nbytes + nrm.nbytes, "byteLength": idx.nbytes}],
        "accessors": [{"bufferView": 0, "componentType": 5126, "count": 24, "type": "VEC3"},
                      {"bufferView": 1, "componentType": 5126, "count": 24, "type": "VEC3"},
                      {"bufferView": 2, "componentType": 5125, "count": 36, "type": "SCALAR"}],
    }
    (tmp_path / "cube.gltf").write_text(json.dumps(doc))
    _sky_hdr(tmp_path / "sky.hdr")
    pre = pyre_amd.Preset.from_gltf(str(tmp_path / "cube.gltf"))
    pre.set_env_hdr(str(tmp_path / "sky.hdr"), 1.0)
    w = h = 40
    cam = pre.camera(1.0)  # framed over the scene bounds (main.rs:186-194)
    sc = pyre_amd.Scene.from_preset(ctx, pre)
    sc.build_bvh(bins=12, max_leaf=8)
    g = pyre_amd.Film(w, h).render(sc, cam, pyre_amd.PathIntegrator(8, 3), 16).reshape(h, w, 10)
    sc.close()
    o = _oracle_scene(oracle, pre).render(_ocam(oracle, cam), w, h, 16).reshape(h, w, 10)
    _assert_env_scene_parity(g, o)
