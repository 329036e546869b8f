"""The traversal benchmark workload at its full size: 2^20 random triangles
(binned SAH, 32 bins, leaf <= 4) and 2^24 rays, 2^23 closest-hit and 2^23
any-hit, on device arrays through pyre_trace_closest_dev / pyre_trace_any_dev.

The oracle cannot trace 2^24 rays in seconds, so the full set is checked by
properties that do not depend on size, and a slice of it against the oracle:
  * any-hit on the closest-hit rays is set exactly where a closest hit exists
    (scene.rs:102-140: occluded is intersect's existence question);
  * no triangle lies strictly before the closest hit: any-hit on the same
    rays with t_max lowered to the hit's t finds nothing;
  * every hit t lies in (t_min, t_max);
  * the first 2^13 closest-hit rays match the oracle bit for bit (id, t), rays
    whose oracle hit lies within 1e-6 of a triangle edge excluded and counted,
    and the first 2^13 any-hit rays agree with the oracle's occluded.
"""
import numpy as np
import pytest

import pyre_amd
from tests.helpers import edge_distance, near_edge_hits

pytestmark = pytest.mark.gpu

N_TRIS, N_RAYS, N_CLOSEST = 1 << 20, 1 << 24, 1 << 23
N_SPOT = 1 << 13


def _ray36(r32):
    """(o, t_min, d, t_max) device records -> (o, d, t_min, t_max, time) host rays."""
    out = np.zeros((len(r32), 9), np.float32)
    out[:, 0:3] = r32[:, 0:3]
    out[:, 3:6] = r32[:, 4:7]
    out[:, 6] = r32[:, 3]
    out[:, 7] = r32[:, 7]
    return out


def test_benchmark_workload_full_size(ctx, oracle):
    import torch
    print(f"[full size] tris={N_TRIS} rays={N_RAYS}", flush=True)
    tris = pyre_amd.random_triangles(1, N_TRIS, 0.01)
    rays = pyre_amd.random_rays(2, N_RAYS, N_CLOSEST, "ray32")
    pos = tris.reshape(-1, 3)
    idx = np.arange(3 * N_TRIS, dtype=np.uint32).reshape(-1, 3)
    sc = pyre_amd.Scene.from_arrays(ctx, [dict(positions=pos, indices=idx, material_id=0)],
                                    [pyre_amd.Material(pyre_amd.MAT_LAMBERTIAN, (0.5, 0.5, 0.5))])
    sc.build_bvh(bins=32, max_leaf=4)

    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        d = torch.from_numpy(rays).to("cuda")
        hits = torch.empty((N_CLOSEST, 2), dtype=torch.int32, device="cuda")
        occ_c = torch.empty(N_CLOSEST, dtype=torch.int32, device="cuda")
        occ_a = torch.empty(N_RAYS - N_CLOSEST, dtype=torch.int32, device="cuda")
        occ_t = torch.empty(N_CLOSEST, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    closest, shadow = d[:N_CLOSEST], d[N_CLOSEST:]
    sc.trace_closest_dev(closest.data_ptr(), N_CLOSEST, hits.data_ptr(), s.cuda_stream)
    sc.trace_any_dev(closest.data_ptr(), N_CLOSEST, occ_c.data_ptr(), s.cuda_stream)
    sc.trace_any_dev(shadow.data_ptr(), N_RAYS - N_CLOSEST, occ_a.data_ptr(), s.cuda_stream)
    s.synchronize()

    h = hits.cpu().numpy()
    prim = h[:, 0].view(np.uint32)
    t = h[:, 1].view(np.float32)
    hit = prim != pyre_amd.MISS
    print(f"  closest hits {hit.mean():.3f}, any-hit rate {occ_a.float().mean().item():.3f}", flush=True)
    assert 0.05 < hit.mean() < 0.95
    assert np.array_equal(occ_c.cpu().numpy().astype(bool), hit)
    r = rays[:N_CLOSEST]
    assert np.all(prim[hit] < N_TRIS)
    assert np.all((t[hit] > r[hit, 3]) & (t[hit] < r[hit, 7]))

    # nothing strictly before the closest hit
    with torch.cuda.stream(s):
        lowered = closest.clone()
        th = torch.from_numpy(np.where(hit, t, r[:, 7]).astype(np.float32)).to("cuda")
        lowered[:, 7] = th
    s.synchronize()
    sc.trace_any_dev(lowered.data_ptr(), N_CLOSEST, occ_t.data_ptr(), s.cuda_stream)
    s.synchronize()
    assert not occ_t.cpu().numpy().any()

    # a slice against the oracle
    o_sc = oracle.Scene()
    o_sc.add_material(oracle.material(0, (0.5, 0.5, 0.5)))
    o_sc.add_mesh(pos, idx, None, 0, 32, 4)
    o = o_sc.intersect(_ray36(rays[:N_SPOT]))
    o_hit = o["kind"] == 1
    t3 = tris.reshape(-1, 3, 3)
    hi = np.where(o_hit)[0]
    ed = edge_distance(t3[o["tri"][hi], 0], t3[o["tri"][hi], 1], t3[o["tri"][hi], 2], o["u"][hi], o["v"][hi])
    keep = np.ones(N_SPOT, bool)
    keep[hi[ed < 1e-6]] = False
    print(f"  spot check: {N_SPOT} rays, {o_hit.sum()} hits, {(~keep).sum()} excluded near an edge", flush=True)
    assert (~keep).sum() < N_SPOT // 1000 + 5
    assert np.array_equal(hit[:N_SPOT][keep], o_hit[keep])
    both = keep & o_hit
    assert np.array_equal(prim[:N_SPOT][both], o["tri"][both].astype(np.uint32))
    assert np.array_equal(t[:N_SPOT][both].view(np.uint32), o["t"][both].view(np.uint32))
    ar = _ray36(rays[N_CLOSEST:N_CLOSEST + N_SPOT])
    oa = o_sc.any_hit(ar)
    ga = occ_a[:N_SPOT].cpu().numpy().astype(bool)
    excl = near_edge_hits(o_sc, ar, [t3])
    print(f"  any-hit spot check: {N_SPOT} rays, {excl.sum()} excluded near an edge", flush=True)
    assert ((ga != oa.astype(bool)) & ~excl).sum() == 0
    assert excl.sum() < N_SPOT // 1000 + 5
    sc.close()


def test_production_render_full_size(ctx):
    """The production configuration at full size (64 level-7 icosphere
    instances = 20,971,522 triangles with the ground, 2048x1024 sky/sun HDRI,
    1920x1080, depth 12, thin lens 0.05). Every sample is its own stream
    (pixel_seed(x, y, s), sampler.rs:41-49) summed in sample order, so 8 spp
    in one pass, 4 + 4 spp in two session passes and two one-shot renders of
    samples [0, 4) and [4, 8) give the same FP64 sums bit for bit; the
    radiance is finite and non-negative everywhere."""
    print("[full size] production 1920x1080, 8 spp", flush=True)
    pre = pyre_amd.Preset.production(seed=3, n_instances=64, level=7, env_w=2048, env_h=1024)
    sc = pyre_amd.Scene.from_preset(ctx, pre)
    info = sc.build_bvh()
    assert info["n_nodes"] > 20_000_000
    w, h = 1920, 1080
    cam = pre.camera(w / h, aperture=0.05)
    ip = pyre_amd.PathIntegrator(12, 3)
    f = pyre_amd.Film(w, h)
    img = f.render(sc, cam, ip, 8, accum64=True)
    one = f.accum64.copy()
    assert np.all(np.isfinite(img)) and img[:, :3].min() >= 0.0
    sess = pyre_amd.Session(sc, cam, ip, w, h)
    sess.render(4)
    sess.render(4)
    _, two, spp = sess.read(mean=False, radiance64=True)
    sess.close()
    assert spp == 8
    assert np.array_equal(two.reshape(-1).view(np.uint64), one.reshape(-1).view(np.uint64))
    g = pyre_amd.Film(w, h)
    g.render(sc, cam, ip, 4, spp_offset=0, accum64=True)
    g.render(sc, cam, ip, 4, spp_offset=4, accum64=True)
    assert np.array_equal(g.accum64.reshape(-1).view(np.uint64), one.reshape(-1).view(np.uint64))
    print(f"  mean radiance {img[:, :3].mean():.4f}", flush=True)
    # the configuration's 64 spp: one pass of the default 2^27-path pool
    # (64 x 1920 x 1080 = 132,710,400 paths) against eight one-shot renders of
    # 8 spp, FP64 sums bit for bit, and the same rays traced
    print("[full size] production 1920x1080, 64 spp in one pass vs 8 x 8 spp", flush=True)
    f64 = pyre_amd.Film(w, h)
    m64 = f64.render(sc, cam, ip, 64, accum64=True)
    ext64, sh64 = sc.render_stats()
    f8 = pyre_amd.Film(w, h)
    ext8 = sh8 = 0
    for k in range(8):
        f8.render(sc, cam, ip, 8, spp_offset=8 * k, accum64=True)
        e, s = sc.render_stats()
        ext8, sh8 = ext8 + e, sh8 + s
    assert np.array_equal(f64.accum64.reshape(-1).view(np.uint64), f8.accum64.reshape(-1).view(np.uint64))
    assert (ext64, sh64) == (ext8, sh8) and ext64 >= 64 * w * h
    assert np.all(np.isfinite(m64)) and m64[:, :3].min() >= 0.0
    sc.close()


def test_radiance_cache_on_the_production_view(ctx):
    """The radiance-cache configuration at full size: the production scene
    at 1920x1080, depth 12, thin lens 0.05, paths ending in the cache after
    the first diffuse bounce, the cache trained online on 2^16 samples per
    frame (4 Adam steps of 2^14 from 2^18 training paths). Properties: the
    image is finite and non-negative, paths do end in the cache and training
    samples are produced, and a trained cache renders closer to a 256-spp
    FP64 reference than an untrained one at equal spp."""
    from tests.helpers import rel_mse
    print("[full size] radiance cache, production 1920x1080", flush=True)
    pre = pyre_amd.Preset.production(seed=3, n_instances=64, level=7, env_w=2048, env_h=1024)
    sc = pyre_amd.Scene.from_preset(ctx, pre)
    sc.build_bvh()
    w, h = 1920, 1080
    cam = pre.camera(w / h, aperture=0.05)
    ip = pyre_amd.PathIntegrator(12, 3)
    ref = pyre_amd.Film(w, h)
    ref.render(sc, cam, ip, 256, spp_offset=1 << 20, accum64=True)
    ref = ref.accum64 / 256
    lo, hi = sc.bounds()
    errs = {}
    for kind, steps in (("trained", 4), ("untrained", 0)):
        cache = pyre_amd.RadianceCache(ctx, lo, hi, seed=5)
        sess = pyre_amd.Session(sc, cam, ip, w, h, nrc_train_paths=1 << 18)
        kw = dict(train_paths=1 << 18, train_batch=1 << 14, train_steps=steps)
        sess.render_nrc(cache, frames=32, spp_per_frame=0, warmup_frames=32, **kw)  # training only
        losses = sess.render_nrc(cache, frames=4, spp_per_frame=4, **kw)
        mean, r64, spp = sess.read(mean=True, radiance64=True)
        st = sess.stats()
        assert spp == 16
        assert np.all(np.isfinite(mean)) and mean[:, :3].min() >= 0.0
        assert st["cache_queries"] > 0 and st["training_samples"] > 0
        if steps:
            assert np.all(np.isfinite(losses)) and cache.stats()[0] > 100
        errs[kind] = rel_mse(r64 / spp, ref)
        sess.close()
        cache.close()
    print(f"  relMSE at 16 spp vs 256 spp: {errs}", flush=True)
    assert errs["trained"] < errs["untrained"]
    sc.close()
