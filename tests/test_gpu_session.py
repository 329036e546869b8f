"""Progressive sessions (pyre_session_*) and the radiance-cache render mode.

A session renders samples s = 0, 1, 2, ... per pixel with the streams
pixel_seed(x, y, s) (sampler.rs:41-49), accumulating in sample order as
main.rs:275-285 does, so any split of N spp into calls and passes -- eager or
replayed from the captured hipGraph -- equals pyre_render(spp = N) bit for bit.
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


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_session_graph_replays_match_eager_render(ctx):
    """Two samples per pass (a pool of 2 * pixels paths), so every call
    replays the captured pass graph; each replay is checked against an eager
    one-shot render of the same sample count, with reads (device work and a
    synchronisation) between the replays."""
    pre, sc = _cornell(ctx)
    w, h = 40, 32
    cam = pre.camera(w / h)
    ip = pyre_amd.PathIntegrator(8, 3)
    with ctx.tuning(pool_paths=2 * w * h):
        sess = pyre_amd.Session(sc, cam, ip, w, h)
        total = 0
        for k in range(4):
            print(f"[session graph] replay {k}", flush=True)
            sess.render(2)
            total += 2
            m, r64, spp = sess.read(radiance64=True)
            assert spp == total
            with ctx.tuning(eager=1, pool_paths=1 << 21):
                film = pyre_amd.Film(w, h)
                ref = film.render(sc, cam, ip, total, accum64=True)
            assert np.array_equal(_bits(m), _bits(ref)), f"replay {k}: graph pass differs from eager render"
            assert np.array_equal(r64, film.accum64)
        sess.close()


def test_session_resume_and_reset(ctx):
    pre, sc = _cornell(ctx)
    w, h = 32, 32
    cam = pre.camera(1.0)
    ip = pyre_amd.PathIntegrator(8, 3)
    sess = pyre_amd.Session(sc, cam, ip, w, h)
    # the pool grows from 3 to 5 samples per pass (the pass graph is captured
    # again on the new buffers), then the graph replays a pass of 2 samples
    sess.render(3)
    sess.render(5)
    sess.render(2)
    m, _, spp = sess.read()
    assert spp == 10
    with ctx.tuning(eager=1):
        ref = pyre_amd.Film(w, h).render(sc, cam, ip, 10)
    assert np.array_equal(_bits(m), _bits(ref))
    # reset clears the sums; the sample index keeps advancing (samples 10..13)
    sess.reset()
    sess.render(4)
    m2, _, spp2 = sess.read()
    assert spp2 == 4
    with ctx.tuning(eager=1):
        ref2 = pyre_amd.Film(w, h).render(sc, cam, ip, 4, spp_offset=10)
    assert np.array_equal(_bits(m2), _bits(ref2))
    ext, shadow = sc.render_stats()
    assert ext > 0 and shadow > 0
    sess.close()


def _nrc_setup(ctx, w, h):
    pre, sc = _cornell(ctx)
    cam = pre.camera(w / h)
    lo, hi = sc.bounds()
    cache = pyre_amd.RadianceCache(ctx, lo, hi, seed=5)
    return pre, sc, cam, cache


def test_nrc_render_trains_and_terminates(ctx):
    w, h = 48, 48
    pre, sc, cam, cache = _nrc_setup(ctx, w, h)
    ip = pyre_amd.PathIntegrator(8, 3)
    sess = pyre_amd.Session(sc, cam, ip, w, h, nrc_train_paths=1 << 13)
    losses = sess.render_nrc(cache, frames=40, spp_per_frame=1, train_paths=1 << 13, train_batch=1 << 12,
                             train_steps=4, warmup_frames=20)
    print(f"[nrc render] losses first {losses[:5]} last {losses[-5:]}", flush=True)
    # single-path targets are noisy (relative L2 stays near their variance), so
    # the loss is only checked to be finite; the gain test checks the cache
    assert np.all(np.isfinite(losses)) and np.all(losses >= 0)
    steps, last = cache.stats()
    assert steps == 40 * 4 and np.isfinite(last)
    m, _, spp = sess.read()
    assert spp == 20
    img = m.reshape(h, w, 10)[..., :3]
    assert np.all(np.isfinite(img)) and img.min() >= 0
    # cache paths stop early: fewer extension rays than a plain render of the same spp
    ext_nrc, _ = sc.render_stats()
    pyre_amd.Film(w, h).render(sc, cam, ip, 20)
    ext_plain, _ = sc.render_stats()
    print(f"  extension rays nrc={ext_nrc} plain={ext_plain}")
    sess.close()


def test_nrc_plain_frames_are_the_plain_renderers_samples(ctx):
    """Plain frames of a cache render trace their image samples to the end:
    with the frames' training paths riding in the same passes and the cache
    taking its Adam steps, the film's FP64 sums equal a plain session's of
    the same samples bit for bit and no path queries the cache; a cache
    frame after them stops paths in the cache."""
    w, h = 40, 32
    pre, sc, cam, cache = _nrc_setup(ctx, w, h)
    ip = pyre_amd.PathIntegrator(8, 3)
    kw = dict(spp_per_frame=2, train_paths=1 << 13, train_batch=1 << 12, train_steps=2)
    s = pyre_amd.Session(sc, cam, ip, w, h, nrc_train_paths=1 << 13)
    losses = s.render_nrc(cache, frames=3, plain_frames=3, **kw)
    _, r64, spp = s.read(mean=False, radiance64=True)
    st = s.stats()
    assert spp == 6
    assert st["cache_queries"] == 0 and st["training_samples"] > 0
    assert cache.stats()[0] == 3 * 2 and np.all(np.isfinite(losses))
    p = pyre_amd.Session(sc, cam, ip, w, h)
    p.render(6)
    _, p64, pspp = p.read(mean=False, radiance64=True)
    p.close()
    assert pspp == 6
    assert np.array_equal(r64.view(np.uint64), p64.view(np.uint64))
    s.render_nrc(cache, frames=1, **kw)
    assert s.stats()["cache_queries"] > 0
    s.close()
    cache.close()


def test_nrc_gain_over_plain_at_equal_spp(ctx):
    """What the trained cache gains: at 4 spp, relMSE against a 2048-spp
    plain reference is lower with paths ending in the trained cache than
    with the plain path tracer, and lower than with an untrained cache."""
    w, h = 48, 48
    pre, sc, cam, cache = _nrc_setup(ctx, w, h)
    ip = pyre_amd.PathIntegrator(8, 3)
    ref = pyre_amd.Film(w, h).render(sc, cam, ip, 2048).reshape(h, w, 10)[..., :3]
    plain = pyre_amd.Film(w, h).render(sc, cam, ip, 4, spp_offset=4096).reshape(h, w, 10)[..., :3]
    e_plain = rel_mse(plain, ref)

    def nrc_image(train_frames):
        c = pyre_amd.RadianceCache(ctx, *sc.bounds(), seed=5)
        s = pyre_amd.Session(sc, cam, ip, w, h, nrc_train_paths=1 << 13)
        if train_frames:
            s.render_nrc(c, frames=train_frames, spp_per_frame=1, train_paths=1 << 13, train_batch=1 << 12,
                         train_steps=4, warmup_frames=train_frames)
        s.reset()
        # the measured frames do not train: the cache is frozen (or never trained)
        s.render_nrc(c, frames=4, spp_per_frame=1, train_paths=1 << 13, train_batch=1 << 12, train_steps=0)
        m, _, spp = s.read()
        assert spp == 4
        s.close()
        c.close()
        return m.reshape(h, w, 10)[..., :3]

    e_trained = rel_mse(nrc_image(120), ref)
    e_untrained = rel_mse(nrc_image(0), ref)
    print(f"[nrc gain] relMSE plain={e_plain:.4f} trained={e_trained:.4f} untrained={e_untrained:.4f}", flush=True)
    assert e_trained < e_plain
    assert e_trained < e_untrained


def test_viewer_running_mean_equals_welford_of_samples(ctx):
    """viewer.rs:165-177 keeps d += (s - d) / n per pass in f32. The session's
    running mean must equal that update applied on the host to the per-sample
    radiance (one-sample renders return the sample itself, mean x * (1/1)),
    also when several samples share one pass."""
    pre, sc = _cornell(ctx)
    w, h = 24, 20
    cam = pre.camera(w / h)
    ip = pyre_amd.PathIntegrator(8, 3)
    with ctx.tuning(pool_paths=3 * w * h):
        sess = pyre_amd.Session(sc, cam, ip, w, h)
        sess.render(2)
        sess.render(5)
        got = sess.read_viewer()
        sess.close()
    d = np.zeros((w * h, 3), np.float32)
    for s in range(7):
        L = pyre_amd.Film(w, h).render(sc, cam, ip, 1, spp_offset=s)[:, :3].astype(np.float32)
        n = np.float32(s + 1)
        d = (d + (L - d) / n).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), d.view(np.uint32))
