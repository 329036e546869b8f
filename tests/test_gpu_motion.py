"""Motion blur: meshes translating over the shutter (InstanceMotion,
crates/pyre/src/geometry/mesh.rs:43-148).

A moving mesh is hit by the ray shifted by -lerp(offset_t0, offset_t1,
clamp(time, 0, 1)); its hit position is Ray::at on the shifted ray, shifted
back. The oracle restates that per primitive (its own BVH per mesh, like the
reference); the device keeps one scene-global tree whose boxes for moving
triangles are swept over the shutter and shifts the origin at the triangle
test. Hits must match the oracle bit for bit (ids, t, u, v), at times inside
and outside [0, 1]; renders of the motion-blurred Cornell box are correlated
with the oracle's like the static ones, and so is the CLI's studio preset
with its moving sphere and gradient sky; batches with moving and static
scenes equal single renders.
"""
import numpy as np
import pytest

import pyre_amd
from tests.helpers import edge_distance, near_edge_hits, rel_mse

pytestmark = pytest.mark.gpu
F32 = np.float32

MOTIONS = [None, ((0.0, 0.0, 0.0), (0.3, -0.2, 0.1)), ((-0.2, 0.1, 0.05), (0.1, 0.1, -0.3))]


def _soups():
    out = []
    for k, (seed, n) in enumerate([(61, 20000), (62, 6000), (63, 6000)]):
        tris = pyre_amd.random_triangles(seed, n, 0.01)
        pos = tris.reshape(-1, 3)
        idx = np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)
        out.append((tris, pos, idx))
    return out


def _scenes(ctx, oracle, bins=32, leaf=4):
    soups = _soups()
    meshes = []
    o_sc = oracle.Scene()
    o_sc.add_material(oracle.material(0, (0.5, 0.5, 0.5)))
    for (tris, pos, idx), mo in zip(soups, MOTIONS):
        m = dict(positions=pos, indices=idx, material_id=0)
        if mo is not None:
            m["motion"] = np.array(mo, np.float32)
        meshes.append(m)
        prim = o_sc.add_mesh(pos, idx, None, 0, bins, leaf)
        if mo is not None:
            o_sc.set_motion(prim, mo[0], mo[1])
    sc = pyre_amd.Scene.from_arrays(ctx, meshes, [pyre_amd.Material(pyre_amd.MAT_LAMBERTIAN, (0.5, 0.5, 0.5))])
    sc.build_bvh(bins=bins, max_leaf=leaf)
    return soups, sc, o_sc


def _timed_rays(seed, n, n_closest):
    rays = pyre_amd.random_rays(seed, n, n_closest)
    rays[:, 8] = np.random.default_rng(seed).uniform(-0.25, 1.25, n).astype(np.float32)
    return rays


def test_moving_meshes_closest_hit_bit_exact(ctx, oracle):
    print("[motion closest] 3 meshes (2 moving), 40000 rays", flush=True)
    soups, sc, o_sc = _scenes(ctx, oracle)
    rays = _timed_rays(71, 40000, 40000)
    g = sc.intersect(rays)
    o = o_sc.intersect(rays)
    o_hit = o["kind"] == 1
    hi = np.where(o_hit)[0]
    ed = np.empty(len(hi))
    for p in range(3):
        sel = o["id"][hi] == p
        t3 = soups[p][0].reshape(-1, 3, 3)
        tri = o["tri"][hi][sel]
        ed[sel] = edge_distance(t3[tri, 0], t3[tri, 1], t3[tri, 2], o["u"][hi][sel], o["v"][hi][sel])
    keep = np.ones(len(rays), bool)
    keep[hi[ed < 1e-6]] = False
    g_hit = g["inst"] != pyre_amd.MISS
    both = keep & o_hit
    moving = both & (o["id"] > 0)
    print(f"  hits {o_hit.sum()} (on moving meshes {moving.sum()}), excluded {(~keep).sum()}", flush=True)
    assert moving.sum() > 500
    assert (~keep).sum() < len(rays) // 1000 + 5
    assert np.array_equal(g_hit[keep], o_hit[keep])
    assert np.array_equal(g["inst"][both], o["id"][both])
    assert np.array_equal(g["tri"][both], o["tri"][both])
    for k in ("t", "u", "v"):
        assert np.array_equal(g[k][both].view(np.uint32), o[k][both].view(np.uint32)), k
    # the same rays at another time hit differently
    r2 = rays.copy()
    r2[:, 8] = 1.0 - r2[:, 8]
    assert (sc.intersect(r2)["t"] != g["t"]).mean() > 0.01


def test_moving_meshes_any_hit_and_bounds(ctx, oracle):
    print("[motion any-hit] 40000 rays", flush=True)
    soups, sc, o_sc = _scenes(ctx, oracle)
    rays = _timed_rays(72, 40000, 0)
    g = sc.occluded(rays).astype(bool)
    o = o_sc.any_hit(rays).astype(bool)
    c = sc.intersect(rays)
    assert np.array_equal(g, c["inst"] != pyre_amd.MISS)
    excl = near_edge_hits(o_sc, rays, [sp[0].reshape(-1, 3, 3) for sp in soups])
    print(f"  excluded near an edge {excl.sum()}", flush=True)
    assert ((g != o) & ~excl).sum() == 0
    assert excl.sum() < len(rays) // 1000 + 5
    assert 0.05 < g.mean() < 0.95
    # Scene::bounds with each moving mesh's box at both ends of the shutter
    lo = np.full(3, np.inf, np.float32)
    hi = np.full(3, -np.inf, np.float32)
    for (tris, pos, idx), mo in zip(soups, MOTIONS):
        b0, b1 = pos.min(0), pos.max(0)
        offs = [np.zeros(3, np.float32)] if mo is None else [np.float32(mo[0]), np.float32(mo[1])]
        for off in offs:
            lo = np.minimum(lo, b0 + np.asarray(off, np.float32))
            hi = np.maximum(hi, b1 + np.asarray(off, np.float32))
    blo, bhi = sc.bounds()
    assert np.array_equal(np.asarray(blo, np.float32), lo) and np.array_equal(np.asarray(bhi, np.float32), hi)


def _ocam(oracle, cam):
    return oracle.camera(cam.origin, cam.target, cam.up, cam.vfov_deg, cam.aspect, cam.aperture_radius,
                         cam.focus_distance if cam.focus_distance > 0 else None)


def test_cornell_motion_blur_correlated_with_oracle(ctx, oracle):
    print("[motion cornell] 48x48 32 spp", flush=True)
    w = h = 48
    pre = pyre_amd.Preset.cornell(motion_blur=True)
    assert pre.arrays()["meshes"][5]["motion"].tolist() == [[0, 0, 0], [0, 0.25, 0]]
    sc = pyre_amd.Scene.from_preset(ctx, pre)
    sc.build_bvh(bins=12, max_leaf=8)
    cam = pre.camera(1.0)
    ip = pyre_amd.PathIntegrator(8, 3)
    g = pyre_amd.Film(w, h).render(sc, cam, ip, 32).reshape(h, w, 10)
    o = oracle.Scene.cornell(motion_blur=True).render(_ocam(oracle, cam), w, h, 32).reshape(h, w, 10)
    same = np.all(g.view(np.uint32) == o.view(np.uint32), axis=-1).mean()
    e = rel_mse(g[..., :3], o[..., :3])
    # the blur is there: the static box renders differently around the sphere
    st = pyre_amd.Scene.from_preset(ctx, pyre_amd.Preset.cornell())
    st.build_bvh(bins=12, max_leaf=8)
    s = pyre_amd.Film(w, h).render(st, cam, ip, 32).reshape(h, w, 10)
    moved = np.any(s[..., 3:6] != g[..., 3:6], axis=-1).mean()
    print(f"  bit-equal pixels {same:.4f} relMSE(gpu, cpu) {e:.2e}; pixels changed by the motion {moved:.3f}",
          flush=True)
    assert same >= 0.99
    assert e < 1e-9
    assert moved > 0.01
    sc.close()
    st.close()


def test_batch_with_moving_and_static_scenes(ctx):
    print("[motion batch]", flush=True)
    mb = pyre_amd.Preset.cornell(motion_blur=True)
    s0 = pyre_amd.Scene.from_preset(ctx, mb)
    s0.build_bvh()
    cb = pyre_amd.Preset.cornell()
    s1 = pyre_amd.Scene.from_preset(ctx, cb)
    s1.build_bvh()
    J = pyre_amd.Job
    jobs = [J(s0, mb.camera(1.0), pyre_amd.PathIntegrator(8, 3), 24, 24, 6),
            J(s1, cb.camera(1.0), pyre_amd.PathIntegrator(8, 3), 20, 16, 5, spp_offset=3),
            J(s0, mb.camera(1.5), pyre_amd.PathIntegrator(4, 2), 30, 20, 3)]
    outs, accs, _ = pyre_amd.render_batch(ctx, jobs, pool_paths=2000, accum64=True)
    for k, j in enumerate(jobs):
        f = pyre_amd.Film(j.width, j.height)
        ref = f.render(j.scene, j.camera, j.integrator, j.spp, spp_offset=j.spp_offset, accum64=True)
        assert np.array_equal(outs[k].view(np.uint32), ref.view(np.uint32)), k
        assert np.array_equal(accs[k], f.accum64), k


def test_studio_preset_with_motion_against_oracle(ctx, oracle):
    """The CLI's studio preset (main.rs:502-562) with its default gradient sky
    (light.rs:252-272) and the moving gold sphere, against the oracle built
    from the same meshes with the oracle's own gradient environment."""
    print("[motion studio] 48x32 16 spp", flush=True)
    pre = pyre_amd.Preset.studio(motion_blur=True, env_intensity=1.0)
    arr = pre.arrays()
    assert [m["motion"] is not None for m in arr["meshes"]] == [False, True, False]
    assert (arr["env"]["width"], arr["env"]["height"]) == (128, 64)
    w, h = 48, 32
    cam = pre.camera(w / h)
    assert cam.origin == (F32(1.6), F32(0.5), F32(2.6)) and cam.vfov_deg == 38.0
    sc = pyre_amd.Scene.from_preset(ctx, pre)
    sc.build_bvh(bins=12, max_leaf=8)
    g = pyre_amd.Film(w, h).render(sc, cam, pyre_amd.PathIntegrator(8, 3), 16).reshape(h, w, 10)
    sc.close()
    o_sc = oracle.Scene()
    for m in arr["materials"]:
        o_sc.add_material(oracle.material(m.kind, m.base_color, m.metallic, m.roughness, m.specular))
    for m in arr["meshes"]:
        prim = o_sc.add_mesh(m["positions"], m["indices"], m["normals"], m["material_id"])
        if m["motion"] is not None:
            o_sc.set_motion(prim, m["motion"][0], m["motion"][1])
    o_sc.set_env(oracle.Hdri(gradient=((0.20, 0.35, 0.65), (0.95, 0.85, 0.70)), intensity=1.0))
    o = o_sc.render(_ocam(oracle, cam), w, h, 16).reshape(h, w, 10)
    same_aov = np.all(g[..., 3:].view(np.uint32) == o[..., 3:].view(np.uint32), axis=-1).mean()
    rel = (np.abs(g[..., :3] - o[..., :3]) / np.maximum(np.abs(o[..., :3]), 1e-6)).max(-1)
    e = rel_mse(g[..., :3], o[..., :3])
    print(f"  AOV bit-equal {same_aov:.4f} radiance within 1e-5 {np.mean(rel <= 1e-5):.4f} relMSE(gpu, cpu) {e:.2e}",
          flush=True)
    # the gradient sky's direction mapping (acos/atan2) is the device math
    # library's: radiance to 1e-5, first-hit AOVs (pinhole camera) bit-exact
    assert same_aov >= 0.99
    assert np.mean(rel <= 1e-5) >= 0.99
    assert e < 1e-8
    assert o[..., :3].mean() > 0.05
