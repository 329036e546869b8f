"""Film sinks packed on the device (pyre_session_pack) against the host
packing (pyre_film_pack) of the same session's means (pyre_session_read):
the linear conversions and the EXR planes bit for bit; the gamma-encoded
PNG bytes within one unit where the device's and the host's powf round the
last ulp differently next to a truncation boundary."""
import numpy as np
import pytest

import pyre_amd
from pyre_amd import _lib as L

pytestmark = pytest.mark.gpu


def test_session_pack_matches_host_pack(ctx, tmp_path):
    print("[film io] cornell 96x72, 8 spp", flush=True)
    pre = pyre_amd.Preset.cornell()
    sc = pyre_amd.Scene.from_preset(ctx, pre)
    sc.build_bvh()
    w, h = 96, 72
    sess = pyre_amd.Session(sc, pre.camera(w / h), pyre_amd.PathIntegrator(8, 3), w, h)
    sess.render(8)
    mean, _, spp = sess.read()
    assert spp == 8
    host = pyre_amd.Film(w, h)
    host.samples[:] = mean
    for what in (L.PACK_NORMAL_PNG, L.PACK_DEPTH_PNG, L.PACK_EXR):
        a, b = sess.pack(what), host.pack(what)
        assert a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8)), what
    for what in (L.PACK_BEAUTY_PNG, L.PACK_ALBEDO_PNG):
        d = np.abs(sess.pack(what).astype(int) - host.pack(what).astype(int))
        print(f"  what={what}: bytes differing {(d > 0).sum()} of {d.size}", flush=True)
        assert d.max() <= 1 and (d > 0).mean() < 1e-3
    assert sess.pack(L.PACK_BEAUTY_PNG).mean() > 10
    sess.save_png(tmp_path / "c.png")
    sess.save_aov_pngs(tmp_path / "c.png")
    sess.save_exr(tmp_path / "c.exr")
    assert (tmp_path / "c.depth.png").exists() and (tmp_path / "c.exr").stat().st_size > 10 * 4 * w * h
    sess.close()
    sc.close()
