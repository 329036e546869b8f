"""GPU wavefront integrator vs the oracle's PathIntegrator (integrator.rs).

Each GPU path consumes the same per-sample stream (pixel_seed, xoshiro256++)
as the reference, with the same f32 operations, and the BSDF samplers' sin /
cos reproduce the host libm's. Scenes without an environment or thin lens
(the Cornell box) therefore render to the oracle's bits. The environment's
direction mapping (acos, atan2, sin, cos of arbitrary angles) and the
thin-lens disk use the device math library, which can differ from libm by an
ulp: there the images agree to 1e-5 relative, far closer than two independent
renders would (correlated-sampling check).
"""
import numpy as np
import pytest

import pyre_amd
from tests.helpers import rel_mse

pytestmark = pytest.mark.gpu


def _cornell(ctx):
    pre = pyre_amd.Preset.cornell()
    sc = pyre_amd.Scene.from_preset(ctx, pre)
    sc.build_bvh(bins=12, max_leaf=8)
    return pre, sc


def _ocam(oracle, cam):
    return oracle.camera(cam.origin, cam.target, cam.up, cam.vfov_deg, cam.aspect, cam.aperture_radius,
                         cam.focus_distance if cam.focus_distance > 0 else None)


@pytest.mark.parametrize("w,h,spp", [(64, 48, 16), (40, 40, 64)])
def test_cornell_correlated_with_oracle(ctx, oracle, w, h, spp):
    print(f"[render cornell] {w}x{h} spp={spp}", flush=True)
    pre, sc = _cornell(ctx)
    cam = pre.camera(w / h)
    film = pyre_amd.Film(w, h)
    g = film.render(sc, cam, pyre_amd.PathIntegrator(8, 3), spp).reshape(h, w, 10)
    o = oracle.Scene.cornell().render(_ocam(oracle, cam), w, h, spp).reshape(h, w, 10)
    # every PixelSample field of every pixel bit-equal to the oracle's
    same = np.all(g.view(np.uint32) == o.view(np.uint32), axis=-1)
    e = rel_mse(g[..., :3], o[..., :3])
    print(f"  bit-equal pixels {same.mean():.4f}, relMSE(gpu, cpu)={e:.3e}")
    assert same.mean() >= 0.99
    assert e < 1e-9
    # and the image is not trivially dark
    assert o[..., :3].mean() > 0.05


def test_render_deterministic_and_resumable(ctx):
    pre, sc = _cornell(ctx)
    cam = pre.camera(1.0)
    ip = pyre_amd.PathIntegrator(8, 3)
    f1 = pyre_amd.Film(48, 48)
    a = f1.render(sc, cam, ip, 16, accum64=True).copy()
    f2 = pyre_amd.Film(48, 48)
    b = f2.render(sc, cam, ip, 16).copy()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    f3 = pyre_amd.Film(48, 48)
    f3.render(sc, cam, ip, 8, spp_offset=0, accum64=True)
    f3.render(sc, cam, ip, 8, spp_offset=8, accum64=True)
    assert np.array_equal(f3.accum64, f1.accum64)
    assert np.allclose(f1.accum64 / 16, a[:, :3], rtol=1e-5, atol=1e-6)


def test_production_scene_small(ctx, oracle):
    """Instanced icospheres + sun/sky HDRI + thin lens, small: GPU vs oracle."""
    print("[render production small]", flush=True)
    pre = pyre_amd.Preset.production(seed=3, n_instances=6, level=2, env_w=128, env_h=64)
    sc = pyre_amd.Scene.from_preset(ctx, pre)
    sc.build_bvh(bins=12, max_leaf=8)
    w, h, spp = 48, 27, 16
    cam = pre.camera(16 / 9, aperture=0.05)
    g = pyre_amd.Film(w, h).render(sc, cam, pyre_amd.PathIntegrator(12, 3), spp).reshape(h, w, 10)
    arr = pre.arrays()
    os_ = oracle.Scene()
    for m in arr["materials"]:
        os_.add_material(oracle.material(m.kind, m.base_color, m.metallic, m.roughness, m.specular))
    for m in arr["meshes"]:
        os_.add_mesh(m["positions"], m["indices"], m["normals"], m["material_id"])
    e = arr["env"]
    os_.set_env(oracle.Hdri(e["width"], e["height"], e["rgb"], e["intensity"]))
    o = os_.render(_ocam(oracle, cam), w, h, spp).reshape(h, w, 10)
    same_aov = np.all(g[..., 3:].view(np.uint32) == o[..., 3:].view(np.uint32), axis=-1).mean()
    rel = (np.abs(g[..., :3] - o[..., :3]) / np.maximum(np.abs(o[..., :3]), 1e-6)).max(-1)
    e2 = rel_mse(g[..., :3], o[..., :3])
    print(f"  AOV bit-equal={same_aov:.4f} radiance within 1e-5: {np.mean(rel <= 1e-5):.4f} "
          f"relMSE(gpu,cpu)={e2:.3e}")
    # first-hit AOVs: thin-lens camera rays are within an ulp of the oracle's,
    # so a few pixels see another triangle or a depth an ulp away
    assert same_aov > 0.95
    assert np.mean(rel <= 1e-5) >= 0.99
    assert e2 < 1e-8


def test_many_small_meshes_first_hit_aovs(ctx, oracle):
    """2,500 meshes of 0..9 triangles (every 97th empty), each with its own
    material, so mesh boundaries fall inside the blocks of 4096 triangles the
    shade stage's mesh lookup table brackets: the first hit's albedo (its
    mesh's material), normal and depth AOVs equal the oracle's bit for bit
    (pinhole camera, no lights: the radiance is zero on both sides)."""
    print("[render many meshes] 2500 meshes, 128x128, 1 spp", flush=True)
    rng = np.random.default_rng(12)
    n_mesh = 2500
    counts = rng.integers(0, 10, n_mesh)
    counts[::97] = 0
    meshes, mats = [], []
    os_ = oracle.Scene()
    for m in range(n_mesh):
        base = (((m % 13) + 1) / 14.0, ((m % 7) + 1) / 8.0, ((m % 5) + 1) / 6.0)
        mats.append(pyre_amd.Material(pyre_amd.MAT_LAMBERTIAN, base))
        os_.add_material(oracle.material(0, base))
        # the mesh's triangles side by side in cell m of a 50 x 50 grid on z = 0
        cx, cy = (m % 50) - 24.5, (m // 50) - 24.5
        pos = np.array([[(cx - 0.45 + 0.1 * k, cy - 0.4, 0.0), (cx - 0.35 + 0.1 * k, cy - 0.4, 0.0),
                         (cx - 0.45 + 0.1 * k, cy + 0.4, 0.0)] for k in range(int(counts[m]))],
                       np.float32).reshape(-1, 3)
        idx = np.arange(pos.shape[0], dtype=np.uint32).reshape(-1, 3)
        meshes.append(dict(positions=pos, indices=idx, material_id=m))
        if counts[m]:
            os_.add_mesh(pos, idx, None, m)
    assert int(counts.sum()) > 2 * 4096
    sc = pyre_amd.Scene.from_arrays(ctx, meshes, mats)
    sc.build_bvh(bins=12, max_leaf=8)
    w = h = 128
    cam = pyre_amd.Camera((0.0, 0.0, 60.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, 1.0)
    g = pyre_amd.Film(w, h).render(sc, cam, pyre_amd.PathIntegrator(8, 3), 1).reshape(h, w, 10)
    o = os_.render(_ocam(oracle, cam), w, h, 1).reshape(h, w, 10)
    hit = o[..., 9] > 0
    same = np.all(g[..., 3:].view(np.uint32) == o[..., 3:].view(np.uint32), axis=-1)
    n_alb = len(np.unique(o[hit][:, 3:6], axis=0))
    print(f"  pixels on a triangle {hit.mean():.3f}, materials seen {n_alb}, AOV bit-equal {same.mean():.4f}")
    assert hit.mean() > 0.15 and n_alb > 100
    assert same.all()


def test_ray_counts_match_reference_queries(ctx, oracle):
    """The device's extension and shadow ray counts are the reference's
    Scene::intersect and Scene::occluded calls for the same samples (paths of
    the same streams take the same bounces, so the counts are equal). These
    counts are the units of the traversal byte model."""
    w, h, spp = 48, 48, 8
    pre, sc = _cornell(ctx)
    cam = pre.camera(1.0)
    pyre_amd.Film(w, h).render(sc, cam, pyre_amd.PathIntegrator(8, 3), spp)
    ext, sh = sc.render_stats()
    st = oracle.Scene.cornell().render_stats(_ocam(oracle, cam), w, h, spp=spp, stride=1, max_depth=8, min_rr=3)
    print(f"  gpu ext={ext} shadow={sh}; oracle {st}", flush=True)
    assert ext == st["extension_queries"]
    assert sh == st["shadow_queries"]
    assert st["n_node_extension"] > 1 and st["n_tri_extension"] > 0


@pytest.mark.parametrize("switch", [dict(shade_waves=1), dict(stack_ring=4, trace_blocks_per_cu=4),
                                    dict(stack_ring=16), dict(pool_paths=5000), dict(eager=1),
                                    dict(memory_fraction=1e-6)])
def test_render_tuning_switches_same_bits(ctx, switch):
    """Tuning switches (the shade kernel built without an occupancy target,
    stack rings, traversal grid, pool size, eager launches, a pool bounded by
    a tiny share of device memory) change speed only: the film is
    bit-identical to the default render."""
    pre, sc = _cornell(ctx)
    cam = pre.camera(1.0)
    ip = pyre_amd.PathIntegrator(8, 3)
    ref = pyre_amd.Film(32, 32).render(sc, cam, ip, 6).copy()
    with ctx.tuning(**switch):
        got = pyre_amd.Film(32, 32).render(sc, cam, ip, 6)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_split_continuations_bound_the_cache(ctx):
    """pyre_render_split: paths are the plain renderer's up to the vertex a
    cache render would query the cache at, then continue on a stream of
    their own. With no such vertex (max depth 1) the render is the plain one
    bit for bit; a continuation is deterministic; the mean over K
    continuations is unbiased (its mean radiance is the reference's) and has
    the variance of the prefix alone plus 1/K of the suffix's, below the
    plain render's at the same spp."""
    pre, sc = _cornell(ctx)
    w = h = 32
    cam = pre.camera(1.0)
    one = pyre_amd.PathIntegrator(1, 3)
    a = pyre_amd.Film(w, h)
    a.render_split(sc, cam, one, 4, continuation=3)
    b = pyre_amd.Film(w, h)
    b.render(sc, cam, one, 4, accum64=True)
    assert np.array_equal(a.accum64.view(np.uint64), b.accum64.view(np.uint64))
    ip = pyre_amd.PathIntegrator(8, 3)
    c1, c2 = pyre_amd.Film(w, h), pyre_amd.Film(w, h)
    c1.render_split(sc, cam, ip, 4, continuation=5)
    c2.render_split(sc, cam, ip, 4, continuation=5)
    assert np.array_equal(c1.accum64.view(np.uint64), c2.accum64.view(np.uint64))
    K = 64
    sp = pyre_amd.Film(w, h)
    for k in range(K):
        sp.render_split(sc, cam, ip, 4, continuation=k)
    split_mean = sp.accum64 / (4 * K)
    ref = pyre_amd.Film(w, h)
    ref.render(sc, cam, ip, 4096, spp_offset=1 << 20, accum64=True)
    ref = ref.accum64 / 4096
    plain = pyre_amd.Film(w, h)
    plain.render(sc, cam, ip, 4, accum64=True)
    plain = plain.accum64 / 4
    e_split, e_plain = rel_mse(split_mean, ref), rel_mse(plain, ref)
    print(f"  relMSE at 4 spp: plain {e_plain:.4f}, {K} continuations {e_split:.4f}; "
          f"mean {split_mean.mean():.4f} vs reference {ref.mean():.4f}", flush=True)
    assert abs(split_mean.mean() - ref.mean()) < 0.01 * ref.mean()
    assert e_split < e_plain


def test_split_continuations_unbiased_with_environment(ctx):
    """The same on the small production scene (sky/sun HDRI, thin lens,
    Disney materials): the mean over continuations agrees with a 4096-spp
    reference image to 1 % in mean radiance."""
    pre = pyre_amd.Preset.production(seed=3, n_instances=6, level=2, env_w=128, env_h=64)
    sc = pyre_amd.Scene.from_preset(ctx, pre)
    sc.build_bvh(bins=12, max_leaf=8)
    w, h = 48, 27
    cam = pre.camera(16 / 9, aperture=0.05)
    ip = pyre_amd.PathIntegrator(12, 3)
    K = 32
    sp = pyre_amd.Film(w, h)
    for k in range(K):
        sp.render_split(sc, cam, ip, 16, continuation=k)
    ref = pyre_amd.Film(w, h)
    ref.render(sc, cam, ip, 4096, spp_offset=1 << 20, accum64=True)
    m, r = sp.accum64.mean() / (16 * K), ref.accum64.mean() / 4096
    print(f"  mean radiance: {K} continuations {m:.5f}, reference {r:.5f}", flush=True)
    assert abs(m - r) < 0.01 * r
    sc.close()
