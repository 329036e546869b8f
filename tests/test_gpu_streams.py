"""Trace entry points on several HIP streams.

The trace calls of a context share one traversal scratch (the persistent
kernels' work counter and stack spill columns), sized by pyre_bvh_build.
Calls enqueued on different streams are ordered on the device in call order,
so interleaving streams (and scenes) gives the same hits as one stream.
"""
import numpy as np
import pytest

import pyre_amd

pytestmark = pytest.mark.gpu


def _scene(ctx, seed, n):
    tris = pyre_amd.random_triangles(seed, n, 0.01)
    pos = tris.reshape(-1, 3)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)
    sc = pyre_amd.Scene.from_arrays(ctx, [dict(positions=pos, indices=idx, material_id=0)],
                                    [pyre_amd.Material(pyre_amd.MAT_LAMBERTIAN, (0.5, 0.5, 0.5))])
    sc.build_bvh(bins=32, max_leaf=4)
    return sc


def test_trace_calls_on_two_streams_and_two_scenes(ctx):
    import torch
    print("[streams] 2 scenes, 2 streams, 2^18 rays each", flush=True)
    n = 1 << 18
    scenes = [_scene(ctx, 81, 1 << 16), _scene(ctx, 82, 1 << 17)]
    rays = pyre_amd.random_rays(83, n, n // 2, "ray32")
    d_rays = torch.from_numpy(rays).to("cuda")
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]

    def run(assign):
        """assign[k] = (scene, stream) of call k; each call traces all rays."""
        outs = []
        for sc_i, st_i in assign:
            with torch.cuda.stream(streams[st_i]):
                hits = torch.empty((n, 2), dtype=torch.int32, device="cuda")
                occ = torch.empty(n, dtype=torch.int32, device="cuda")
            torch.cuda.synchronize()
            scenes[sc_i].trace_closest_dev(d_rays.data_ptr(), n, hits.data_ptr(), streams[st_i].cuda_stream)
            scenes[sc_i].trace_any_dev(d_rays.data_ptr(), n, occ.data_ptr(), streams[st_i].cuda_stream)
            outs.append((hits, occ))
        torch.cuda.synchronize()
        return [(h.cpu().numpy(), o.cpu().numpy()) for h, o in outs]

    ref = run([(0, 0), (1, 0)])  # one stream
    for k in range(3):
        got = run([(0, 0), (1, 1), (0, 1), (1, 0)])  # interleaved streams, no waits between calls
        for (gh, go), (rh, ro) in zip(got, [ref[0], ref[1], ref[0], ref[1]]):
            assert np.array_equal(gh, rh), f"round {k}: closest hits differ"
            assert np.array_equal(go, ro), f"round {k}: any-hit differs"
    assert (ref[0][0][:, 0] != -1).mean() > 0.01
    for sc in scenes:
        sc.close()
