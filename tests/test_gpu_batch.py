"""Batched renders (pyre_render_batch): jobs of different scenes, cameras,
resolutions and spp share one path pool. Every job's PixelSample means and
FP64 sums must equal pyre_render of that job alone, bit for bit: each path
uses its own per-sample stream (pixel_seed(x, y, s), sampler.rs:41-49) and
each pixel sums its samples in order (main.rs:275-285), whatever pass or
pool position the path gets.
"""
import numpy as np
import pytest

import pyre_amd

pytestmark = pytest.mark.gpu


def _scenes(ctx):
    cb = pyre_amd.Preset.cornell()
    s0 = pyre_amd.Scene.from_preset(ctx, cb)
    s0.build_bvh()
    p1 = pyre_amd.Preset.instances(41, 5, 2, 64, 32, 0.5)
    s1 = pyre_amd.Scene.from_preset(ctx, p1)
    s1.build_bvh(bins=16, max_leaf=4)
    p2 = pyre_amd.Preset.instances(42, 3, 3, 128, 64, 0.5)
    s2 = pyre_amd.Scene.from_preset(ctx, p2)
    s2.build_bvh()
    return [(cb, s0), (p1, s1), (p2, s2)]


def _jobs(ctx):
    sc = _scenes(ctx)
    (cb, s0), (p1, s1), (p2, s2) = sc
    J = pyre_amd.Job
    jobs = [
        J(s0, cb.camera(40 / 30), pyre_amd.PathIntegrator(8, 3), 40, 30, 6),
        J(s1, p1.camera(1.0, aperture=0.05), pyre_amd.PathIntegrator(12, 3), 24, 24, 9, spp_offset=5),
        J(s2, p2.camera(16 / 9), pyre_amd.PathIntegrator(5, 2), 32, 18, 4),
        J(s0, cb.camera(1.0), pyre_amd.PathIntegrator(3, 1), 16, 16, 13),
        J(s1, p1.camera(2.0), pyre_amd.PathIntegrator(12, 3), 50, 25, 2),
    ]
    return sc, jobs


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("pool", [3000, 1 << 21])
def test_batch_equals_single_renders(ctx, pool):
    """pool 3000 paths: many passes, several jobs per pass, graph replays;
    pool 2^21: one pass holds everything."""
    print(f"[batch] pool={pool}", flush=True)
    sc, jobs = _jobs(ctx)
    outs, accs, stats = pyre_amd.render_batch(ctx, jobs, pool_paths=pool, accum64=True)
    print(f"  stats {stats}", flush=True)
    if pool == 3000:
        assert stats["passes"] > 5
    else:
        assert stats["passes"] == 1
    assert stats["extension_rays"] > 0 and stats["shadow_rays"] > 0
    for k, j in enumerate(jobs):
        f = pyre_amd.Film(j.width, j.height)
        ref = f.render(j.scene, j.camera, j.integrator, j.spp, spp_offset=j.spp_offset, accum64=True)
        assert np.array_equal(_bits(outs[k]), _bits(ref)), f"job {k}: means differ from pyre_render"
        assert np.array_equal(accs[k], f.accum64), f"job {k}: FP64 sums differ"


def test_batch_autotune_same_result(ctx):
    sc, jobs = _jobs(ctx)
    base, _, _ = pyre_amd.render_batch(ctx, jobs, pool_paths=5000)
    tuned, _, st = pyre_amd.render_batch(ctx, jobs, autotune=True)
    print(f"[batch autotune] chose pool={st['pool_paths']} bpc={st['trace_blocks_per_cu']} ring={st['stack_ring']}")
    total = sum(j.width * j.height * j.spp for j in jobs)
    big = max(j.width * j.height for j in jobs)
    assert st["pool_paths"] in [max(min(p, total), big) for p in (1 << 22, 1 << 24, 1 << 26)]
    assert st["stack_ring"] in (4, 8, 16) and st["trace_blocks_per_cu"] in (4, 8)
    for a, b in zip(base, tuned):
        assert np.array_equal(_bits(a), _bits(b))


def test_batch_rejects_bad_jobs(ctx):
    sc, jobs = _jobs(ctx)
    unbuilt = pyre_amd.Scene.from_preset(ctx, pyre_amd.Preset.cornell())
    bad = [pyre_amd.Job(unbuilt, jobs[0].camera, jobs[0].integrator, 8, 8, 1)]
    with pytest.raises(pyre_amd.PyreError):
        pyre_amd.render_batch(ctx, bad)
    zero = [pyre_amd.Job(jobs[0].scene, jobs[0].camera, jobs[0].integrator, 8, 8, 0)]
    with pytest.raises(pyre_amd.PyreError):
        pyre_amd.render_batch(ctx, zero)
    outs, _, st = pyre_amd.render_batch(ctx, [])
    assert outs == [] and st["passes"] == 0


def test_batch_of_the_96_job_workload(ctx):
    """The batched-jobs workload itself (bench.py's job list, seed 4: 96
    jobs of 128^2..1024^2 pixels, 4..256 spp, 10k..2M-triangle scenes with
    2048x1024 sky/sun maps) in one pyre_render_batch: every output finite
    and non-negative, and the four smallest jobs' FP64 sums bit-equal to
    their own pyre_render."""
    import bench
    print("[batch] the 96-job workload", flush=True)
    jobs, scenes = bench.batch_jobs(ctx)
    assert len(jobs) == 96
    outs, accs, st = pyre_amd.render_batch(ctx, jobs, accum64=True)
    print(f"  {st['passes']} passes, gpu {st['gpu_ms']:.1f} ms, pool {st['pool_paths']}", flush=True)
    for k, (o, a) in enumerate(zip(outs, accs)):
        assert np.all(np.isfinite(o)) and np.all(np.isfinite(a)), f"job {k}"
        assert o[:, :3].min() >= 0.0 and a.min() >= 0.0, f"job {k}"
    small = sorted(range(len(jobs)), key=lambda k: jobs[k].width * jobs[k].height * jobs[k].spp)[:4]
    for k in small:
        j = jobs[k]
        f = pyre_amd.Film(j.width, j.height)
        f.render(j.scene, j.camera, j.integrator, j.spp, spp_offset=j.spp_offset, accum64=True)
        assert np.array_equal(accs[k].view(np.uint64), f.accum64.view(np.uint64)), f"job {k}"
    for _, sc, _ in scenes:
        sc.close()


def test_mixed_batch_against_oracle(ctx, oracle):
    """A batch of different scenes -- the Cornell box, instanced icospheres
    under a sky/sun map with a thin lens, and the studio preset with its
    moving sphere and gradient sky -- every job against the oracle's render
    of the same samples (correlated samples; Cornell to the bit)."""
    from tests.helpers import assert_render_parity, oracle_camera, oracle_scene
    cb = pyre_amd.Preset.cornell()
    pi = pyre_amd.Preset.instances(43, 4, 2, 128, 64, 0.5)
    st_ = pyre_amd.Preset.studio(motion_blur=True, env_intensity=1.0)
    scs = []
    for pre in (cb, pi, st_):
        sc = pyre_amd.Scene.from_preset(ctx, pre)
        sc.build_bvh()
        scs.append(sc)
    J = pyre_amd.Job
    jobs = [J(scs[0], cb.camera(1.0), pyre_amd.PathIntegrator(8, 3), 24, 24, 6, spp_offset=2),
            J(scs[1], pi.camera(16 / 9, aperture=0.05), pyre_amd.PathIntegrator(12, 3), 32, 18, 5),
            J(scs[2], st_.camera(1.5), pyre_amd.PathIntegrator(8, 3), 24, 16, 4, spp_offset=7)]
    outs, _, st = pyre_amd.render_batch(ctx, jobs, pool_paths=1500)
    assert st["passes"] > 1
    for k, (pre, j, exact) in enumerate(zip((cb, pi, st_), jobs, (True, False, False))):
        o = oracle_scene(oracle, pre)
        if k == 2:  # the studio preset's sky is HdriEnvironmentLight::gradient (light.rs:252-272)
            o.set_env(oracle.Hdri(gradient=((0.20, 0.35, 0.65), (0.95, 0.85, 0.70)), intensity=1.0))
        ref = o.render(oracle_camera(oracle, j.camera), j.width, j.height, j.spp, spp_offset=j.spp_offset,
                       max_depth=j.integrator.max_depth, min_rr=j.integrator.min_rr_depth)
        fr = assert_render_parity(outs[k].reshape(j.height, j.width, 10), ref.reshape(j.height, j.width, 10), exact)
        print(f"  job {k}: {fr}", flush=True)
    for sc in scs:
        sc.close()
