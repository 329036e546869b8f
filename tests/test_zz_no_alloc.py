"""pyre_trace_closest_dev_noalloc (no counterpart in the reference; for
callers that cannot allocate): the closest-hit kernel on the library's fixed
device scratch array must give pyre_trace_closest_dev's hits bit for bit, on
two calls in a row, the second with a smaller input (fewer rays, and a
smaller scene), so the second call runs on the array the first one left
behind. Run alone in a fresh process and after the other tests."""
import numpy as np
import pytest

import pyre_amd

pytestmark = pytest.mark.gpu


def _scene(ctx, seed, n):
    tris = pyre_amd.random_triangles(seed, n, 0.01)
    sc = pyre_amd.Scene.from_arrays(
        ctx, [dict(positions=tris.reshape(-1, 3), indices=np.arange(3 * n, dtype=np.uint32).reshape(-1, 3),
                   material_id=0)], [pyre_amd.Material(pyre_amd.MAT_LAMBERTIAN, (0.5, 0.5, 0.5))])
    sc.build_bvh(bins=32, max_leaf=4)
    return sc


def test_noalloc_closest_hit_equals_trace_closest_dev(ctx):
    import torch
    s = torch.cuda.Stream()
    cases = [(_scene(ctx, 91, 1 << 18), 1 << 21), (_scene(ctx, 92, 1 << 14), 1 << 16)]
    for k, (sc, n) in enumerate(cases):
        print(f"[no-alloc] call {k}: {n} rays, depth {sc.bvh_info()['depth']}", flush=True)
        rays = pyre_amd.random_rays(93 + k, n, n // 2, "ray32")
        with torch.cuda.stream(s):
            d = torch.from_numpy(rays).to("cuda")
            ref = torch.empty((n, 2), dtype=torch.int32, device="cuda")
            got = torch.full((n, 2), 7, dtype=torch.int32, device="cuda")
        s.synchronize()
        sc.trace_closest_dev(d.data_ptr(), n, ref.data_ptr(), s.cuda_stream)
        sc.trace_closest_dev_noalloc(d.data_ptr(), n, got.data_ptr(), s.cuda_stream)
        s.synchronize()
        r, g = ref.cpu().numpy(), got.cpu().numpy()
        assert np.array_equal(g, r), f"call {k}: {(g != r).any(axis=1).sum()} hits differ"
        assert (r[:, 0] != -1).mean() > 0.01
    for sc, _ in cases:
        sc.close()
