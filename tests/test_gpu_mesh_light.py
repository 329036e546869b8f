"""Emissive triangle meshes as lights (pyre_mesh_light) on the GPU against
the oracle's MeshLight: the general emitter the reference's PBRT builder
drops (io/pbrt/builder.rs:463-485,513-516), behind the Light trait
(light.rs:36-42) and the integrator's light loops (integrator.rs:113-167).

The device samples, intersects and weights the meshes with the oracle's f32
operations, so a scene without an environment renders to the oracle's bits,
as the Cornell box with its quad light does.
"""
import numpy as np
import pytest

import pyre_amd
from tests.helpers import rel_mse
from tests.test_oracle_mesh_light import octahedron

pytestmark = pytest.mark.gpu


def _quad_tris(l):
    p0, eu, ev = l[0:3], l[3:6], l[6:9]
    return np.array([p0, p0 + eu, p0 + eu + ev, p0 + ev], np.float32), np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


def _scenes(ctx, oracle, keep_quad, extra_emitter, env=None):
    """The Cornell preset with its light as a two-triangle mesh light (or kept
    as a quad) and optionally an emissive octahedron inside the box."""
    arr = pyre_amd.Preset.cornell().arrays()
    mls = []
    if not keep_quad:
        for l in arr["lights"]:
            v, f = _quad_tris(l)
            mls.append(dict(positions=v, indices=f, emission=tuple(l[9:12])))
    if extra_emitter:
        v, f = octahedron((0.25, -0.55, 0.2), 0.12)
        mls.append(dict(positions=v, indices=f, emission=(4.0, 2.5, 1.0)))
    quads = arr["lights"] if keep_quad else []
    g = pyre_amd.Scene.from_arrays(ctx, arr["meshes"], arr["materials"], quads, env=env, mesh_lights=mls)
    g.build_bvh(bins=12, max_leaf=8)
    o = oracle.Scene()
    for m in arr["materials"]:
        o.add_material(oracle.material(m.kind, m.base_color, m.metallic, m.roughness, m.specular))
    for m in arr["meshes"]:
        o.add_mesh(m["positions"], m["indices"], m["normals"], m["material_id"])
    for l in quads:
        o.add_light(l)
    for m in mls:
        o.add_mesh_light(m["positions"], m["indices"], m["emission"])
    return g, o


@pytest.mark.parametrize("keep_quad,extra", [(False, False), (False, True), (True, True)])
def test_mesh_lights_render_to_the_oracles_bits(ctx, oracle, keep_quad, extra):
    print(f"[mesh light] quad={keep_quad} octahedron={extra}", flush=True)
    g_sc, o_sc = _scenes(ctx, oracle, keep_quad, extra)
    cam = pyre_amd.Preset.cornell().camera(1.0)
    ocam = oracle.camera(cam.origin, cam.target, cam.up, cam.vfov_deg, cam.aspect)
    w = h = 40
    spp = 32
    g = pyre_amd.Film(w, h).render(g_sc, cam, pyre_amd.PathIntegrator(8, 3), spp).reshape(h, w, 10)
    o = o_sc.render(ocam, w, h, spp).reshape(h, w, 10)
    same = np.all(g.view(np.uint32) == o.view(np.uint32), axis=-1)
    e = rel_mse(g[..., :3], o[..., :3])
    print(f"  bit-equal pixels {same.mean():.4f}, relMSE(gpu, cpu)={e:.3e}, mean {o[..., :3].mean():.4f}")
    assert same.mean() >= 0.99
    assert e < 1e-9
    assert o[..., :3].mean() > 0.05
    if extra:
        # the octahedron is seen directly: its pixels end on a light (albedo AOV = 1)
        assert np.any(np.all(o[..., 3:6] == 1.0, -1))


def test_mesh_light_with_environment(ctx, oracle):
    """Slot order quads, meshes, environment: the octahedron next to a sky."""
    rng = np.random.default_rng(7)
    rgb = (0.2 + rng.random((16, 32, 3))).astype(np.float32)
    env = dict(width=32, height=16, rgb=rgb, intensity=0.5)
    g_sc, o_sc = _scenes(ctx, oracle, True, True, env=env)
    o_sc.set_env(oracle.Hdri(32, 16, rgb, 0.5))
    cam = pyre_amd.Preset.cornell().camera(1.0)
    ocam = oracle.camera(cam.origin, cam.target, cam.up, cam.vfov_deg, cam.aspect)
    g = pyre_amd.Film(32, 32).render(g_sc, cam, pyre_amd.PathIntegrator(6, 3), 16).reshape(32, 32, 10)
    o = o_sc.render(ocam, 32, 32, 16, max_depth=6).reshape(32, 32, 10)
    # the environment's direction mapping uses the device math library: AOVs
    # bit-equal, radiance within 1e-5 (as for every scene with an HDRI)
    assert np.array_equal(g[..., 3:].view(np.uint32), o[..., 3:].view(np.uint32))
    close = np.abs(g[..., :3] - o[..., :3]) <= 1e-5 * np.maximum(1.0, np.abs(o[..., :3]))
    print(f"[mesh light + env] pixels within 1e-5: {close.all(-1).mean():.4f}")
    assert close.all(-1).mean() >= 0.99
    assert rel_mse(g[..., :3], o[..., :3]) < 1e-6


def test_mesh_light_refusals(ctx):
    arr = pyre_amd.Preset.cornell().arrays()
    v, f = octahedron((0, 0, 0), 0.1)
    bad = [dict(positions=v, indices=np.array([[0, 1, 9]], np.uint32), emission=(1, 1, 1)),  # index out of range
           dict(positions=np.zeros((3, 3), np.float32), indices=np.array([[0, 1, 2]], np.uint32),
                emission=(1, 1, 1)),  # no area
           dict(positions=v, indices=np.zeros((0, 3), np.uint32), emission=(1, 1, 1))]  # no triangles
    for m in bad:
        with pytest.raises(pyre_amd.PyreError):
            pyre_amd.Scene.from_arrays(ctx, arr["meshes"], arr["materials"], arr["lights"], mesh_lights=[m])


def test_pbrt_mesh_emitters_render_to_the_oracles_bits(ctx, oracle, tmp_path):
    """A PBRT file whose emissive sphere (48 x 24 uv sphere, 2,208 triangles)
    and emissive triangle load as mesh lights (PYRE_PBRT_MESH_EMITTERS)."""
    from tests.test_scene_io_emitters import SCENE
    f = tmp_path / "emitters.pbrt"
    f.write_text(SCENE)
    pre = pyre_amd.Preset.from_pbrt(str(f), mesh_emitters=True)
    arr = pre.arrays()
    assert len(arr["mesh_lights"]) == 2
    g_sc = pyre_amd.Scene.from_preset(ctx, pre)
    g_sc.build_bvh(bins=12, max_leaf=8)
    o_sc = oracle.Scene()
    for m in arr["materials"]:
        o_sc.add_material(oracle.material(m.kind, m.base_color, m.metallic, m.roughness, m.specular))
    for m in arr["meshes"]:
        o_sc.add_mesh(m["positions"], m["indices"], m["normals"], m["material_id"])
    for l in arr["lights"]:
        o_sc.add_light(l)
    for m in arr["mesh_lights"]:
        o_sc.add_mesh_light(m["positions"], m["indices"], m["emission"])
    cam = pre.camera(1.0)
    ocam = oracle.camera(cam.origin, cam.target, cam.up, cam.vfov_deg, cam.aspect)
    print(f"[pbrt mesh emitters] {sum(len(m['indices']) for m in arr['mesh_lights'])} emitter triangles", flush=True)
    g = pyre_amd.Film(32, 32).render(g_sc, cam, pyre_amd.PathIntegrator(5, 3), 8).reshape(32, 32, 10)
    o = o_sc.render(ocam, 32, 32, 8, max_depth=5).reshape(32, 32, 10)
    same = np.all(g.view(np.uint32) == o.view(np.uint32), axis=-1)
    print(f"  bit-equal pixels {same.mean():.4f}, mean radiance {o[..., :3].mean():.4f}")
    assert same.mean() >= 0.99
    assert rel_mse(g[..., :3], o[..., :3]) < 1e-9
    assert o[..., :3].mean() > 0.01


@pytest.mark.parametrize("pool", [3000, 1 << 20])
def test_mesh_light_jobs_in_a_batch(ctx, oracle, pool):
    """Jobs of scenes with 2 and 1 light slots share one path pool; each job
    equals its own pyre_render (compared with the oracle above)."""
    g_sc, _ = _scenes(ctx, oracle, True, True)
    pre = pyre_amd.Preset.cornell()
    c_sc = pyre_amd.Scene.from_preset(ctx, pre)
    c_sc.build_bvh(bins=12, max_leaf=8)
    cam = pre.camera(1.0)
    ip = pyre_amd.PathIntegrator(8, 3)
    jobs = [pyre_amd.Job(g_sc, cam, ip, 24, 24, 8), pyre_amd.Job(c_sc, cam, ip, 16, 16, 4),
            pyre_amd.Job(g_sc, cam, ip, 20, 20, 4, 3)]
    print(f"[mesh light batch] pool={pool}", flush=True)
    outs, accs, st = pyre_amd.render_batch(ctx, jobs, pool_paths=pool, accum64=True)
    for k, j in enumerate(jobs):
        f = pyre_amd.Film(j.width, j.height)
        ref = f.render(j.scene, j.camera, j.integrator, j.spp, spp_offset=j.spp_offset, accum64=True)
        assert np.array_equal(np.ascontiguousarray(outs[k], np.float32).view(np.uint32),
                              np.ascontiguousarray(ref, np.float32).view(np.uint32)), f"job {k}"
        assert np.array_equal(accs[k], f.accum64), f"job {k}: FP64 sums"


def test_mesh_lights_in_cache_renders(ctx, oracle):
    """A radiance-cache render of a scene with a quad and an emissive mesh:
    its plain frames (training paths riding in the same passes) give the
    plain session's FP64 sums bit for bit, and a cache frame after them
    queries the cache and stays finite."""
    w, h = 32, 32
    g_sc, _ = _scenes(ctx, oracle, True, True)
    cam = pyre_amd.Preset.cornell().camera(w / h)
    lo, hi = g_sc.bounds()
    cache = pyre_amd.RadianceCache(ctx, lo, hi, seed=5)
    ip = pyre_amd.PathIntegrator(8, 3)
    kw = dict(spp_per_frame=2, train_paths=1 << 13, train_batch=1 << 12, train_steps=2)
    print("[mesh light cache render]", flush=True)
    s = pyre_amd.Session(g_sc, cam, ip, w, h, nrc_train_paths=1 << 13)
    losses = s.render_nrc(cache, frames=2, plain_frames=2, **kw)
    _, r64, spp = s.read(mean=False, radiance64=True)
    assert spp == 4 and np.all(np.isfinite(losses))
    assert s.stats()["training_samples"] > 0
    p = pyre_amd.Session(g_sc, cam, ip, w, h)
    p.render(4)
    _, p64, _ = p.read(mean=False, radiance64=True)
    p.close()
    assert np.array_equal(r64.view(np.uint64), p64.view(np.uint64))
    s.render_nrc(cache, frames=1, **kw)
    assert s.stats()["cache_queries"] > 0
    m, _, _ = s.read()
    assert np.all(np.isfinite(m))
    s.close()
    cache.close()
