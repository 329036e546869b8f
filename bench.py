#!/usr/bin/env python
"""Headline benchmark: closest-hit traversal throughput on one MI355X.

Workload (the raw traversal microbench): 2^20 random triangles (centroids
U[-1,1]^3, vertices offset by N(0, 0.01^2)), a binned-SAH BVH (32 bins,
leaf <= 4) built on the GPU, and 2^24 rays (origins U[-1.5,1.5]^3, directions
uniform on the sphere): the first half closest-hit with t_max = inf, the
second half any-hit shadow rays with t_max ~ U[0.1, 3]. Seeds 1 (triangles)
and 2 (rays).

A step is one closest-hit pass over the 2^23 closest-hit rays, inputs
resident in HBM; value = closest-hit Mrays/s over K timed steps. The any-hit
pass is timed the same way in its own block and reported beside it.

The metric's second half follows on the same line: seconds to relMSE 0.01
at 1920x1080 on the production scene (64 level-7 icospheres, sky/sun HDRI,
depth 12, thin lens 0.05), the plain path tracer vs the radiance cache
trained online on 2^16 samples per frame, against a 16384-spp FP64
reference rendered on the GPU in the same run: from a cold start
(nrc_over_plain: the cache renderer's first frame is plain while its cache
trains), with a pre-trained cache (nrc_warm_cache_over_plain), and the
bound of a perfect cache (nrc_best_case_over_plain). --no-nrc skips it.

Roofline: algorithmic bytes per closest-hit ray =
32 (ray) + 8 (hit) + 32 * N_node + 36 * N_tri, with N_node / N_tri the node
fetches and triangle tests of the reference traversal (geometry/bvh.rs:86-143)
on the reference build (32 bins, leaf 4) of the same triangles, measured by
the oracle on the first 2^20 closest-hit rays.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "Mrays/s closest-hit; seconds to relMSE 0.01 at 1080p, plain vs NRC"
N_TRIS = 1 << 20
N_RAYS = 1 << 24
N_CLOSEST = 1 << 23
BINS, MAX_LEAF = 32, 4
# Reference traversal statistics on this workload (oracle, 2^20 closest rays).
N_NODE_CLOSEST = 61.5237
N_TRI_CLOSEST = 2.5797
N_NODE_ANY = 57.9721
N_TRI_ANY = 2.4079
BYTES_CLOSEST = 32 + 8 + 32 * N_NODE_CLOSEST + 36 * N_TRI_CLOSEST
BYTES_ANY = 32 + 4 + 32 * N_NODE_ANY + 36 * N_TRI_ANY
HBM_PEAK_GBS = 8000.0
CPU_SAMPLE = 1 << 22
PROFILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")
TRAFFIC_PROFILE = os.path.join(PROFILES, "k_trace4_traversal_2e20tris.md")
PRODUCTION_PROFILE = os.path.join(PROFILES, "render_production.md")


def measured_traffic(key, path=TRAFFIC_PROFILE):
    """A figure of a committed rocprofv3 summary (a "key: value" line), e.g.
    the traversal kernel's HBM bytes per ray from FETCH_SIZE/WRITE_SIZE
    (counters scaled by a plain copy of known size); None when absent."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_info():
    """The host the CPU baseline runs on: CPUs visible to this process
    (nproc) and the CPU model (lscpu). The baseline's thread pool is capped
    at 16, the host-CPU share a one-GPU box grants its jobs."""
    info = {"nproc": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["cpu_model"] = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline(tris, rays36, sample):
    """Oracle (C++ restatement of the reference) on the host cores."""
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build()
    threads = O.nthreads()
    sc = O.Scene()
    sc.add_material(O.material(0, (0.5, 0.5, 0.5)))
    t0 = time.perf_counter()
    sc.add_mesh(tris.reshape(-1, 3), np.arange(3 * len(tris), dtype=np.uint32).reshape(-1, 3), None, 0, BINS, MAX_LEAF)
    build_s = time.perf_counter() - t0
    sub = rays36[:sample]
    t0 = time.perf_counter()
    out = sc.intersect(sub, threads=threads)
    dt = time.perf_counter() - t0
    return sc, out, dict(value=round(sample / dt / 1e6, 3), unit="Mrays/s", cores=threads, kind="port",
                         sample=f"first {sample} closest-hit rays of the workload, oracle BVH "
                                f"(reference build, {BINS} bins, leaf {MAX_LEAF}; build {build_s:.1f} s untimed)",
                         host=host_info())


def bench_cornell(args):
    """Cornell preset (main.rs:365-496), 512x512, 64 spp, depth 8, pinhole:
    the reference's own CPU-runnable case. Step = one full render."""
    import pyre_amd
    ctx = pyre_amd.Context(0)
    dev_name, _ = ctx.device_info()
    W = H = 512
    SPP = 64
    pre = pyre_amd.Preset.cornell()
    scene = pyre_amd.Scene.from_preset(ctx, pre)
    scene.build_bvh()  # reference parameters: 12 bins, leaf 8
    cam = pre.camera(1.0)
    ip = pyre_amd.PathIntegrator(8, 3)
    film = pyre_amd.Film(W, H)
    for _ in range(args.warmup):
        film.render(scene, cam, ip, SPP)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        film.render(scene, cam, ip, SPP)
    wall = time.perf_counter() - t0
    ext, sh = scene.render_stats()
    per = wall / args.steps
    line = {
        "metric": "samples/s (Cornell preset, NEE+MIS path tracer)", "value": round(W * H * SPP / per / 1e6, 2),
        "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(per * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (built-in Cornell preset)",
        "config": {"workload": "Cornell preset 512x512, 64 spp, max depth 8, pinhole", "width": W, "height": H,
                   "spp": SPP, "max_depth": 8},
        "extension_mrays_s": round(ext / per / 1e6, 1), "shadow_mrays_s": round(sh / per / 1e6, 1),
        "rays_per_sample": round((ext + sh) / (W * H * SPP), 3), "device": dev_name,
    }
    # FP64 reference: 16384 spp accumulated in double on the GPU (samples
    # 2^20.. so they are independent of the measured ones)
    ref = None
    if args.ref_spp:
        t0 = time.perf_counter()
        film_ref = pyre_amd.Film(W, H)
        film_ref.render(scene, cam, ip, args.ref_spp, spp_offset=1 << 20, accum64=True)
        ref = film_ref.accum64 / args.ref_spp
        img = film.render(scene, cam, ip, SPP)
        line["relmse_64spp"] = round(rel_mse(img[:, :3], ref), 6)
        line["relmse_eps"] = REL_EPS
        line["reference_spp"] = args.ref_spp
        log(f"cornell reference {args.ref_spp} spp in {time.perf_counter() - t0:.1f}s")
    if not args.no_cpu:
        from oracle import oracle as O
        threads = O.nthreads()
        osc = O.Scene.cornell()
        ocam = O.camera(cam.origin, cam.target, cam.up, cam.vfov_deg, cam.aspect)
        cspp = 8
        t0 = time.perf_counter()
        cimg = osc.render(ocam, W, H, cspp, threads=threads)
        dt = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": round(W * H * cspp / dt / 1e6, 3), "unit": "Msamples/s", "cores": threads,
                                "kind": "port", "sample": f"full 512x512 frame at {cspp} spp (of 64)"}
        # correlated check: the GPU with the same 8 per-sample streams vs the CPU
        gimg = pyre_amd.Film(W, H).render(scene, cam, ip, cspp)
        corr = {"relmse_gpu_vs_cpu_8spp": round(rel_mse(gimg[:, :3], cimg.reshape(-1, 10)[:, :3]), 7)}
        if ref is not None:
            corr["relmse_cpu_8spp_vs_reference"] = round(rel_mse(cimg.reshape(-1, 10)[:, :3], ref), 6)
        line["correlated_check"] = corr
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------- production
PROD_W, PROD_H = 1920, 1080
REL_EPS = 1e-2  # relMSE = mean((I - R)^2 / (R^2 + eps))


def production_scene(ctx):
    """64 level-7 icosphere instances (20,971,520 triangles) with random
    affine transforms and Disney materials, a ground quad, a procedural
    2048x1024 sky + sun HDRI (seed 3); camera auto-framed, 16:9, thin lens
    radius 0.05. BVH with the reference's build parameters (12 bins, leaf 8)."""
    import pyre_amd
    t0 = time.perf_counter()
    pre = pyre_amd.Preset.production(seed=3, n_instances=64, level=7, env_w=2048, env_h=1024)
    scene = pyre_amd.Scene.from_preset(ctx, pre)
    gen_s = time.perf_counter() - t0
    info = scene.build_bvh()
    cam = pre.camera(PROD_W / PROD_H, aperture=0.05)
    log(f"production scene: generated+uploaded in {gen_s:.1f}s, bvh {info['build_ms']:.1f} ms, "
        f"nodes={2 * info['n_leaves'] - 1} depth={info['depth']}")
    return pre, scene, cam, info


def nrc_view(ctx, name):
    """The scene of the cache comparison and of its reference: the production
    view (depth 12), or the Cornell preset at the same 1920x1080 (depth 8;
    a closed, indirect-lit box, where a cache that takes over after the
    first diffuse bounce has most to gain)."""
    import pyre_amd
    if name == "cornell":
        pre = pyre_amd.Preset.cornell()
        scene = pyre_amd.Scene.from_preset(ctx, pre)
        scene.build_bvh()
        return scene, pre.camera(PROD_W / PROD_H), pyre_amd.PathIntegrator(8, 3)
    pre, scene, cam, info = production_scene(ctx)
    return scene, cam, pyre_amd.PathIntegrator(12, 3)


def rel_mse(img, ref):
    img = np.asarray(img, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + REL_EPS)))


def load_reference(path):
    if not path or not os.path.exists(path):
        return None
    ref = np.load(path)
    return ref.reshape(PROD_H * PROD_W, 3).astype(np.float64)


def bench_production(args):
    """Production scene at 1920x1080, 64 spp, max depth 12: one step = one
    full render through a progressive session (a graph-replayed pass per spp)."""
    import pyre_amd
    ctx = pyre_amd.Context(0)
    dev_name, _ = ctx.device_info()
    SPP = 64
    pre, scene, cam, info = production_scene(ctx)
    ip = pyre_amd.PathIntegrator(12, 3)
    sess = pyre_amd.Session(scene, cam, ip, PROD_W, PROD_H)
    for _ in range(args.warmup):
        sess.render(SPP)
        sess.reset()
    walls = []
    for k in range(args.steps):
        sess.reset()
        t0 = time.perf_counter()
        sess.render(SPP)
        walls.append(time.perf_counter() - t0)
        log(f"step {k}: {walls[-1]:.3f}s")
    ext, sh = scene.render_stats()
    per = float(np.mean(walls))
    line = {
        "metric": "samples/s (production scene, 1080p, NEE+MIS path tracer)",
        "value": round(PROD_W * PROD_H * SPP / per / 1e6, 2), "unit": "Msamples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(per * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (procedural scene, seed 3)",
        "config": {"workload": "64 level-7 icosphere instances (20,971,520 tris) + ground, 2048x1024 sky/sun HDRI, "
                               "1920x1080, 64 spp, max depth 12, thin lens 0.05",
                   "triangles": 20971522, "width": PROD_W, "height": PROD_H, "spp": SPP, "max_depth": 12},
        "extension_mrays_s": round(ext / per / 1e6, 1), "shadow_mrays_s": round(sh / per / 1e6, 1),
        "rays_per_sample": round((ext + sh) / (PROD_W * PROD_H * SPP), 3),
        "bvh_build_ms": round(info["build_ms"], 1), "device": dev_name,
    }
    ref = load_reference(args.reference)
    if ref is not None:
        _, r64, spp = sess.read(mean=False, radiance64=True)
        line["relmse_64spp"] = round(rel_mse(r64 / spp, ref), 5)
        line["relmse_eps"] = REL_EPS
    if not args.no_cpu:
        line["cpu_baseline"], osc, ocam = production_cpu_baseline(pre, cam, cspp=2)
        line["roofline"] = production_roofline(osc, ocam, ext, sh, PROD_W * PROD_H * SPP)
    print(json.dumps(line), flush=True)


def shade_byte_model(st, ext, sh, samples):
    """Bytes the shade stage (k_shade_split + k_shade) must move per frame
    by the path pool's layout (fields
    read or written once per vertex; CDF searches and cache-line granularity
    not counted). Vertex kinds: V = extension rays (every one is shaded),
    C = V - samples (vertices past the camera ray), Sh = shadow rays, and
    from the oracle's render of every 8th pixel at 1 spp the fraction of
    extension queries that hit a surface (H = h V) and of camera rays that
    hit something (A = a samples). Per vertex: job id 4, ray 32, hit 16,
    radiance 16 read and 16 written; past the camera ray queue index 4 and
    throughput 16; at a surface the shading normal 36, the sampler state 32
    read (not on camera rays) and 32 written, the shadow count 4; the
    environment 24 (a texel and two pdf values: the miss's emission or the
    hit's light sample); first hits write the AOVs 32; a continuing path
    writes ray 32, throughput 16 and queue index 4; a shadow ray writes ray
    32, contribution 16 and queue index 4."""
    h = st["extension_surface_hits"] / max(st["extension_queries"], 1)
    n_cam = ((PROD_W + 7) // 8) * ((PROD_H + 7) // 8)
    a = st["camera_hits"] / n_cam
    V, S0 = float(ext), float(samples)
    C, Sh = V - S0, float(sh)
    H, A = h * V, a * S0
    reads = V * (4 + 32 + 16 + 16) + C * (4 + 16) + H * 36 + max(H - A, 0.0) * 32 + V * 24
    writes = V * 16 + A * 32 + H * (32 + 4) + C * (32 + 16 + 4) + Sh * (32 + 16 + 4)
    return {"bytes_read": reads, "bytes_written": writes, "surface_hit_fraction": round(h, 4),
            "camera_hit_fraction": round(a, 4)}


def production_roofline(osc, ocam, ext, sh, samples):
    """The shade stage (k_shade, and k_shade_split from the second bounce
    on), the frame's longest, against HBM: the layout byte
    model over the stage's per-frame time and its FETCH_SIZE / WRITE_SIZE
    bytes, both from the committed kernel trace and counter runs of this
    frame (the stage runs inside the library's graph-replayed passes, out of
    reach of events on the caller's stream)."""
    st = osc.render_stats(ocam, PROD_W, PROD_H, spp=1, stride=8, max_depth=12, min_rr=3)
    m = shade_byte_model(st, ext, sh, samples)
    ms = measured_traffic("k_shade_ms_per_frame", PRODUCTION_PROFILE)
    fetch = measured_traffic("k_shade_fetch_bytes_per_frame", PRODUCTION_PROFILE)
    write = measured_traffic("k_shade_write_bytes_per_frame", PRODUCTION_PROFILE)
    model = m["bytes_read"] + m["bytes_written"]
    out = {"bound": "hbm", "kernel": "shade stage: k_shade (12 launches per frame, one per bounce) + k_shade_split "
                                     "(11, from the second bounce); figures per frame",
           "peak": HBM_PEAK_GBS, "unit": "GB/s", "algorithmic_bytes": round(model),
           "traffic": None if fetch is None or write is None else round(fetch + write), "model": m,
           "sample_pixels": "oracle: every 8th pixel in x and y (240 x 135), 1 spp"}
    if ms:
        gbs = model / (ms * 1e-3) / 1e9
        out.update({"achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4), "kernel_ms": ms})
        if out["traffic"]:
            out["traffic_gbs"] = round(out["traffic"] / (ms * 1e-3) / 1e9, 1)
    return out


def production_cpu_baseline(pre, cam, cspp):
    from oracle import oracle as O
    threads = O.nthreads()
    arr = pre.arrays()
    osc = O.Scene()
    for m in arr["materials"]:
        osc.add_material(O.material(m.kind, m.base_color, m.metallic, m.roughness, m.specular))
    t0 = time.perf_counter()
    for m in arr["meshes"]:
        osc.add_mesh(m["positions"], m["indices"], m["normals"], m["material_id"])
    e = arr["env"]
    osc.set_env(O.Hdri(e["width"], e["height"], e["rgb"], e["intensity"]))
    build_s = time.perf_counter() - t0
    log(f"oracle scene built in {build_s:.1f}s")
    ocam = O.camera(cam.origin, cam.target, cam.up, cam.vfov_deg, cam.aspect, cam.aperture_radius,
                    cam.focus_distance if cam.focus_distance > 0 else None)
    t0 = time.perf_counter()
    osc.render(ocam, PROD_W, PROD_H, cspp, max_depth=12, min_rr=3, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(PROD_W * PROD_H * cspp / dt / 1e6, 4), "unit": "Msamples/s", "cores": threads,
            "kind": "port", "sample": f"full 1920x1080 frame at {cspp} spp (of 64); per-mesh BVH build "
                                      f"{build_s:.0f} s untimed"}, osc, ocam


def make_reference(args):
    """FP64-accumulated plain reference of the production view: args.ref_spp
    samples per pixel, saved as float32 radiance (H, W, 3) .npy."""
    import pyre_amd
    ctx = pyre_amd.Context(0)
    scene, cam, ip = nrc_view(ctx, args.nrc_scene)
    sess = pyre_amd.Session(scene, cam, ip, PROD_W, PROD_H)
    t0 = time.perf_counter()
    done = 0
    while done < args.ref_spp:
        n = min(256, args.ref_spp - done)
        sess.render(n)
        done += n
        log(f"reference {done}/{args.ref_spp} spp, {time.perf_counter() - t0:.1f}s")
    _, r64, spp = sess.read(mean=False, radiance64=True)
    img = (r64 / spp).astype(np.float32).reshape(PROD_H, PROD_W, 3)
    os.makedirs(os.path.dirname(os.path.abspath(args.make_reference)), exist_ok=True)
    np.save(args.make_reference, img)
    print(json.dumps({"reference": args.make_reference, "spp": spp, "seconds": round(time.perf_counter() - t0, 1),
                      "mean": float(img.mean())}), flush=True)


def _crossing(curve, target):
    """First time at which relMSE falls to target (log-log interpolation
    between checkpoints); None if it never does."""
    prev = None
    for t, e in curve:
        if e <= target:
            if prev is None:
                return t
            (t0, e0) = prev
            a = (np.log(target) - np.log(e0)) / (np.log(e) - np.log(e0))
            return float(np.exp(np.log(t0) + a * (np.log(t) - np.log(t0))))
        prev = (t, e)
    return None


# Radiance-cache training per frame (the configuration's 2^16 samples): about
# a quarter of the production view's training paths reach a surface vertex
# per bounce, so 2^18 paths give at least 2^16 samples; 4 Adam steps on
# batches of 2^14 use 2^16 of them.
NRC_TRAIN_PATHS = 1 << 18
NRC_BATCH = 1 << 14
NRC_STEPS = 4
# A cache that starts untrained renders its first NRC_COLD_PLAIN frames as
# plain frames (image samples traced to the end while the frame's training
# paths train the cache) and takes no training-only frames: on the
# production view, 1 to 16 plain frames of 4 spp reached relMSE 0.01 in
# 1.56x the plain time, 8 training-only frames then cache frames in 3.06x,
# cache frames from the start in 6.1x; each extra frame is an extra pass
# over the pool, whose deep-bounce tail costs about 2.4 ms, so the cold
# render takes one frame per checkpoint (its spp up to NRC_COLD_SPF).
NRC_COLD_PLAIN = 1
NRC_COLD_SPF = 128
# Frames of at most this many spp for a cache that starts untrained and uses
# it from the start (the --nrc-warmup comparison).
NRC_MAX_SPF = 4
NRC_WARMUP = 0
# Untimed training-only frames of the warm-cache comparison.
NRC_PRETRAIN = 64


def _at_time(curve, t):
    """relMSE at time t on a (time, relMSE) curve, log-log interpolated; None
    outside the measured span."""
    for (t0, e0), (t1, e1) in zip(curve, curve[1:]):
        if t0 <= t <= t1:
            a = (np.log(t) - np.log(t0)) / (np.log(t1) - np.log(t0))
            return float(np.exp(np.log(e0) + a * (np.log(e1) - np.log(e0))))
    return None


# Algorithmic FLOPs of the cache network (110 used inputs, 4 hidden 64x64
# layers, 3 used outputs): forward per query; per training sample the forward,
# dX back to the first hidden layer and dW of every layer.
NRC_FLOP_QUERY = 2 * (110 * 64 + 4 * 64 * 64 + 64 * 3)  # 47,232
NRC_FLOP_SAMPLE = NRC_FLOP_QUERY + 2 * (4 * 64 * 64 + 64 * 3) + NRC_FLOP_QUERY  # 127,616
MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16


def cache_mfma(ctx, lo, hi, args):
    """Inference and training throughput of the cache kernels on device
    arrays (synthetic queries in the scene box): k calls queued back to back
    on one stream between two HIP events, as the renderer queues its training
    steps (the host enqueues a call in about 12 us, less than any call's GPU
    time, so the queue never runs dry); TFLOP/s from the algorithmic FLOPs."""
    import torch
    import pyre_amd
    rng = np.random.default_rng(8)

    def queries(n):
        p = rng.uniform(lo, hi, (n, 3))
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        m = rng.normal(size=(n, 3))
        m /= np.linalg.norm(m, axis=1, keepdims=True)
        return np.concatenate([p, d, m, rng.random((n, 4))], 1).astype(np.float32)

    cache = pyre_amd.RadianceCache(ctx, lo, hi, seed=5)
    stream = torch.cuda.Stream()
    out = {"unit": "TFLOP/s", "peak": MFMA_BF16_PEAK_TFLOPS, "flop_per_query": NRC_FLOP_QUERY,
           "flop_per_training_sample": NRC_FLOP_SAMPLE}

    def timed(fn, k=20, warm=3):
        for _ in range(warm):
            fn()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record(stream)
        for _ in range(k):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / k

    n = 1 << 20
    with torch.cuda.stream(stream):
        dq = torch.from_numpy(queries(n)).to("cuda")
        dt = torch.from_numpy(rng.random((n, 3)).astype(np.float32)).to("cuda")
        rgb = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ms = timed(lambda: cache.infer_dev(dq.data_ptr(), n, rgb.data_ptr(), stream.cuda_stream))
    tf = n * NRC_FLOP_QUERY / (ms * 1e-3) / 1e12
    out["infer"] = {"queries": n, "ms": round(ms, 4), "achieved": round(tf, 1), "frac": round(tf / MFMA_BF16_PEAK_TFLOPS, 4)}
    for lg in (14, 16, 18):
        b = 1 << lg
        ms = timed(lambda: cache.train_step_dev(dq.data_ptr(), dt.data_ptr(), b, 0, stream.cuda_stream))
        tf = b * NRC_FLOP_SAMPLE / (ms * 1e-3) / 1e12
        out[f"train_step_2e{lg}"] = {"samples": b, "ms": round(ms, 4), "achieved": round(tf, 1),
                                     "frac": round(tf / MFMA_BF16_PEAK_TFLOPS, 4)}
    cache.close()
    log(f"cache MFMA: {out}")
    return out


def render_reference(ctx, scene, cam, ip, ref_spp):
    """FP64-accumulated plain render of ref_spp samples per pixel (samples
    2^20.., independent of the measured ones) on the GPU: radiance (H*W, 3)."""
    import pyre_amd
    t0 = time.perf_counter()
    film = pyre_amd.Film(PROD_W, PROD_H)
    done = 0
    while done < ref_spp:
        n = min(256, ref_spp - done)
        film.render(scene, cam, ip, n, spp_offset=(1 << 20) + done, accum64=True)
        done += n
    log(f"reference {ref_spp} spp in {time.perf_counter() - t0:.1f}s")
    return film.accum64 / ref_spp


def nrc_comparison(ctx, scene, cam, ip, ref, max_spp=128, warmup=0, pre_frames=0, spf_max=None, plain_frames=0):
    """Seconds to relMSE 0.01 against `ref`, plain path tracer vs radiance
    cache (paths ending in the cache after the first diffuse bounce). Every
    checkpoint N spp is its own render from a cold start, timed whole from
    the first launch to the last result (pools are reserved before the
    clock starts): plain = one session render of N spp; cache = a fresh
    cache and frames of min(N, 16) spp, each tracing its samples and 2^18
    training paths in one pass, taking 4 Adam steps of 2^14 of the frame's
    training samples, then adding the updated cache's radiance to the paths
    that ended in it; `warmup` training-only frames come first, timed, and
    the first `plain_frames` frames after them trace their image samples to the end
    (the plain renderer's samples) while their training paths train the
    cache.
    pre_frames > 0: the cache is first trained, untimed, for that many
    training-only frames (a warm cache, as a renderer that keeps its cache
    across frames of a still view has one), then the timed render starts
    from it. relMSE is taken outside the timed region."""
    import pyre_amd
    lo, hi = scene.bounds()
    target = 0.01
    spf_max = spf_max or NRC_MAX_SPF
    checkpoints = [c for c in (4, 6, 8, 10, 12, 16, 24, 32, 48, 64, 96, 128) if c <= max_spp]
    stats = {}

    def render(kind, n):
        cache = pyre_amd.RadianceCache(ctx, lo, hi, seed=5) if kind == "nrc" else None
        if cache is not None:
            cache.reserve(NRC_BATCH)
            if pre_frames:
                pre = pyre_amd.Session(scene, cam, ip, PROD_W, PROD_H, nrc_train_paths=NRC_TRAIN_PATHS)
                pre.render_nrc(cache, frames=pre_frames, spp_per_frame=1, train_paths=NRC_TRAIN_PATHS,
                               train_batch=NRC_BATCH, train_steps=NRC_STEPS, warmup_frames=pre_frames)
                pre.close()
        sess = pyre_amd.Session(scene, cam, ip, PROD_W, PROD_H, nrc_train_paths=NRC_TRAIN_PATHS if cache else 0)
        spf = min(n, spf_max)
        sess.reserve(n if cache is None else spf)
        t0 = time.perf_counter()
        if cache is None:
            sess.render(n)
        else:
            sess.render_nrc(cache, frames=warmup + n // spf, spp_per_frame=spf, train_paths=NRC_TRAIN_PATHS,
                            train_batch=NRC_BATCH, train_steps=NRC_STEPS, warmup_frames=warmup,
                            plain_frames=plain_frames)
        dt = time.perf_counter() - t0  # render calls return after the stream is idle
        _, r64, spp = sess.read(mean=False, radiance64=True)
        assert spp == n
        e = rel_mse(r64 / spp, ref)
        stats[kind] = dict(sess.stats(), samples=n * PROD_W * PROD_H, spp=n)
        sess.close()
        if cache is not None:
            cache.close()
        log(f"{kind}: {n} spp from a cold start, {dt:.4f}s, relMSE {e:.5f}")
        return dt, e

    def run(kind):
        render(kind, checkpoints[0])  # untimed: first-use costs (graph capture, code load)
        curve = []
        for n in checkpoints:
            if kind == "nrc" and n > spf_max and n % spf_max:
                continue
            curve.append(render(kind, n))
            if curve[-1][1] <= target * 0.7:
                break
        return curve

    plain = run("plain")
    nrc = run("nrc")
    tp, tn = _crossing(plain, target), _crossing(nrc, target)
    eq = _at_time(nrc, tp) if tp else None
    return {
        "plain_seconds": round(tp, 4) if tp else None, "nrc_seconds": round(tn, 4) if tn else None,
        "nrc_over_plain": round(tn / tp, 4) if (tn and tp) else None,
        "relmse_at_plain_seconds": {"plain": target, "nrc": round(eq, 5) if eq else None},
        "training": {"train_paths_per_frame": NRC_TRAIN_PATHS, "adam_steps_per_frame": NRC_STEPS,
                     "batch": NRC_BATCH, "samples_per_frame": NRC_STEPS * NRC_BATCH, "lr": 1e-2,
                     "max_spp_per_frame": spf_max, "timed_warmup_frames": warmup, "plain_frames": plain_frames,
                     "untimed_pretraining_frames": pre_frames},
        "plain_curve": [(round(t, 4), round(e, 5)) for t, e in plain],
        "nrc_curve": [(round(t, 4), round(e, 5)) for t, e in nrc], "path_stats_last": stats,
    }


def cache_best_case(ctx, scene, cam, ip, ref, plain_seconds, n_spp=8, k_cont=256):
    """The most any cache that takes over after the first diffuse bounce can
    gain on this view. Its error: the mean of k_cont renders of the same
    n_spp samples whose paths continue on k_cont independent streams from the
    vertex a cache render queries the cache at (pyre_render_split) is what a
    perfect cache gives those samples (up to the suffix variance / k_cont);
    an unbiased render's relMSE falls as 1/spp, so a perfect cache needs
    relMSE * n_spp / 0.01 spp. Its cost: a cache render that does not train
    (paths end at the query vertex, the cache is evaluated there), per spp,
    measured on a 16-spp frame. Bound = that many spp at that cost, against
    the plain path tracer's seconds to relMSE 0.01."""
    import pyre_amd
    t0 = time.perf_counter()
    film = pyre_amd.Film(PROD_W, PROD_H)
    for k in range(k_cont):
        film.render_split(scene, cam, ip, n_spp, continuation=k)
    perfect = film.accum64 / (n_spp * k_cont)
    e_perfect = rel_mse(perfect, ref)
    plain = pyre_amd.Film(PROD_W, PROD_H)
    plain.render(scene, cam, ip, n_spp, accum64=True)
    e_plain = rel_mse(plain.accum64 / n_spp, ref)
    need = e_perfect * n_spp / 0.01
    lo, hi = scene.bounds()
    cache = pyre_amd.RadianceCache(ctx, lo, hi, seed=5)
    sess = pyre_amd.Session(scene, cam, ip, PROD_W, PROD_H, nrc_train_paths=1 << 10)
    sess.reserve(16)
    kw = dict(frames=1, spp_per_frame=16, train_paths=1 << 10, train_batch=NRC_BATCH, train_steps=0)
    sess.render_nrc(cache, **kw)  # untimed: first use
    t1 = time.perf_counter()
    sess.render_nrc(cache, **kw)
    per_spp = (time.perf_counter() - t1) / 16
    sess.close()
    cache.close()
    bound = need * per_spp
    out = {"continuations": k_cont, "spp": n_spp, "relmse_perfect_cache": round(e_perfect, 6),
           "mean_radiance_perfect_cache": round(float(perfect.mean()), 5),
           "mean_radiance_reference": round(float(ref.mean()), 5),
           "relmse_plain": round(e_plain, 6), "spp_needed_perfect_cache": round(need, 2),
           "seconds_per_spp_cache_render_without_training": round(per_spp, 6), "seconds": round(bound, 5),
           "over_plain": round(bound / plain_seconds, 4) if plain_seconds else None,
           "wall_s": round(time.perf_counter() - t0, 1)}
    log(f"cache best case: {out}")
    return out


def nrc_block(ctx, args):
    """The second half of the metric on the production view at 1920x1080:
    the FP64 reference rendered here, plain vs cache seconds to relMSE 0.01,
    and the cache kernels' MFMA throughput."""
    import pyre_amd
    t0 = time.perf_counter()
    pre, scene, cam, info = production_scene(ctx)
    ip = pyre_amd.PathIntegrator(12, 3)
    ref = render_reference(ctx, scene, cam, ip, args.ref_spp)
    out = nrc_comparison(ctx, scene, cam, ip, ref, max_spp=args.max_spp, warmup=args.nrc_warmup,
                         plain_frames=NRC_COLD_PLAIN, spf_max=NRC_COLD_SPF)
    warm = nrc_comparison(ctx, scene, cam, ip, ref, max_spp=args.max_spp, warmup=0, pre_frames=NRC_PRETRAIN,
                          spf_max=16)
    out["warm_cache"] = {k: warm[k] for k in ("nrc_seconds", "nrc_over_plain", "relmse_at_plain_seconds",
                                              "training", "nrc_curve", "path_stats_last")}
    out["best_case"] = cache_best_case(ctx, scene, cam, ip, ref, out["plain_seconds"], n_spp=16, k_cont=256)
    lo, hi = scene.bounds()
    out["cache_mfma"] = cache_mfma(ctx, lo, hi, args)
    out["config"] = {"workload": "production scene 1920x1080 (64 level-7 icospheres, 2048x1024 sky/sun HDRI), "
                                 "max depth 12, thin lens 0.05; cache: 5x64 ReLU MLP, bf16 MFMA",
                     "reference_spp": args.ref_spp, "relmse_eps": REL_EPS}
    out["seconds"] = round(time.perf_counter() - t0, 1)
    scene.close()
    return out


def bench_nrc(args):
    """Seconds to relMSE 0.01 at 1080p, plain vs radiance cache, on its own
    (--reference: a saved reference .npy; otherwise one is rendered here)."""
    # torch first: it and the library load a HIP runtime of the same soname,
    # and the first one loaded serves both
    import torch  # noqa: F401
    import pyre_amd
    ctx = pyre_amd.Context(0)
    dev_name, _ = ctx.device_info()
    scene, cam, ip = nrc_view(ctx, args.nrc_scene)
    ref = load_reference(args.reference)
    if ref is None:
        ref = render_reference(ctx, scene, cam, ip, args.ref_spp)
    out = nrc_comparison(ctx, scene, cam, ip, ref, max_spp=args.max_spp, warmup=args.nrc_warmup,
                         plain_frames=NRC_COLD_PLAIN, spf_max=NRC_COLD_SPF)
    lo, hi = scene.bounds()
    out["cache_mfma"] = cache_mfma(ctx, lo, hi, args)
    tn = out["nrc_seconds"]
    line = {
        "metric": "seconds to relMSE 0.01 at 1080p, plain vs NRC", "value": tn, "unit": "s (NRC)", "n_gpus": 1,
        "steps": 1, "warmup": 0, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (cache: bf16 MFMA, f32 accumulate)", "data": "synthetic (procedural scene, seed 3)",
        "config": {"workload": ("production scene 1920x1080, max depth 12, thin lens 0.05" if args.nrc_scene ==
                                "production" else "Cornell preset 1920x1080, max depth 8, pinhole"),
                   "relmse_eps": REL_EPS, "reference_spp": args.ref_spp},
        "device": dev_name, **out,
    }
    print(json.dumps(line), flush=True)


def bench_cache(args):
    """The radiance cache's kernels alone: inference over 2^20 queries and
    training steps of 2^14, 2^16 and 2^18 samples, on synthetic queries in
    the production scene's box (the scene is built for its bounds only)."""
    import torch  # noqa: F401  (first: see bench_nrc)
    import pyre_amd
    ctx = pyre_amd.Context(0)
    dev_name, _ = ctx.device_info()
    _, scene, _, _ = production_scene(ctx)
    lo, hi = scene.bounds()
    scene.close()
    out = cache_mfma(ctx, lo, hi, args)
    t = out["train_step_2e16"]
    line = {
        "metric": "radiance-cache training step, 2^16 samples", "value": t["achieved"], "unit": "TFLOP/s",
        "n_gpus": 1, "steps": 20, "warmup": 3, "ms_per_step": t["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 (f32 accumulate, f32 Adam)", "data": "synthetic queries",
        "config": {"workload": "5x64 ReLU MLP (128-wide encoding in, RGB out), relative-L2 loss, Adam"},
        "device": dev_name, "cache_mfma": out,
    }
    print(json.dumps(line), flush=True)


# ------------------------------------------------------------------- batch
def batch_jobs(ctx, n_jobs=96, seed=4):
    """96 jobs: square images with side log-uniform in [128, 1024] (128^2 to
    1024^2 pixels), spp from {4, 16, 64, 256}, scenes of log-uniform
    10k..2M triangles built from instanced icospheres (level and count
    chosen to hit the triangle count), half Lambertian / half Disney
    materials, the production scene's procedural 2048x1024 sky + sun HDRI;
    pinhole, max depth 8."""
    import pyre_amd
    rng = np.random.default_rng(seed)
    jobs, scenes = [], []
    t0 = time.perf_counter()
    for j in range(n_jobs):
        side = int(round(np.exp(rng.uniform(np.log(128), np.log(1024)))))
        spp = int(rng.choice([4, 16, 64, 256]))
        tris = float(np.exp(rng.uniform(np.log(1e4), np.log(2e6))))
        level = 1
        while tris / (20 * 4 ** level) > 64:
            level += 1
        n_inst = int(min(64, max(1, round(tris / (20 * 4 ** level)))))
        pre = pyre_amd.Preset.instances(1000 + j, n_inst, level, 2048, 1024, 0.5)
        sc = pyre_amd.Scene.from_preset(ctx, pre)
        sc.build_bvh()
        scenes.append((pre, sc, n_inst * 20 * 4 ** level + 2))
        jobs.append(pyre_amd.Job(sc, pre.camera(1.0), pyre_amd.PathIntegrator(8, 3), side, side, spp))
    log(f"batch: {n_jobs} scenes generated and built in {time.perf_counter() - t0:.1f}s")
    return jobs, scenes


def bench_batch(args):
    """All 96 jobs in one pyre_render_batch call (shared path pool, graph-
    replayed passes, autotuned launch shape); step = the whole batch, its
    means written to device buffers (results resident in HBM). One more
    call with host buffers gives the PCIe-inclusive time."""
    # torch first: it and the library load a HIP runtime of the same soname,
    # and the first one loaded serves both
    import torch
    import pyre_amd
    ctx = pyre_amd.Context(0)
    dev_name, _ = ctx.device_info()
    jobs, scenes = batch_jobs(ctx)
    total = sum(j.width * j.height * j.spp for j in jobs)
    dev_out = [torch.empty((j.width * j.height, 10), dtype=torch.float32, device="cuda") for j in jobs]
    ptrs = [t.data_ptr() for t in dev_out]
    torch.cuda.synchronize()
    _, _, st0 = pyre_amd.render_batch(ctx, jobs, autotune=True, device_out=ptrs)  # warmup + tuning
    tune = dict(pool_paths=st0["pool_paths"], trace_blocks_per_cu=st0["trace_blocks_per_cu"],
                stack_ring=st0["stack_ring"])
    log(f"autotuned: {tune}")
    walls, stats = [], None
    for k in range(args.steps):
        t0 = time.perf_counter()
        _, _, stats = pyre_amd.render_batch(ctx, jobs, **tune, device_out=ptrs)
        walls.append(time.perf_counter() - t0)
        log(f"batch step {k}: {walls[-1]:.3f}s ({stats['passes']} passes, gpu {stats['gpu_ms']:.1f} ms)")
    per = float(np.mean(walls))
    t0 = time.perf_counter()
    host_outs, _, _ = pyre_amd.render_batch(ctx, jobs, **tune)
    host_wall = time.perf_counter() - t0
    # the device means are the host means
    for k in (0, len(jobs) - 1):
        assert np.array_equal(dev_out[k].cpu().numpy().view(np.uint32), host_outs[k].view(np.uint32))
    out_bytes = sum(o.nbytes for o in host_outs)
    del host_outs
    # the same jobs one pyre_render call after another
    t0 = time.perf_counter()
    for j in jobs:
        pyre_amd.Film(j.width, j.height).render(j.scene, j.camera, j.integrator, j.spp)
    seq = time.perf_counter() - t0
    line = {
        "metric": "samples/s (96 batched uneven jobs)", "value": round(total / per / 1e6, 2), "unit": "Msamples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": round(per * 1e3, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (job list seed 4)",
        "config": {"workload": "96 jobs: 128^2..1024^2 px (log-uniform), spp {4,16,64,256}, 10k..2M-triangle "
                               "icosphere-instance scenes (log-uniform), Lambertian/Disney 50/50, 2048x1024 sky+sun "
                               "HDRI, pinhole, depth 8", "jobs": len(jobs), "samples": total,
                   "triangles": int(sum(t for _, _, t in scenes))},
        "tuning": tune, "passes": stats["passes"], "gpu_ms": round(stats["gpu_ms"], 1),
        "host_output_ms_per_step": round(host_wall * 1e3, 1), "host_output_bytes": out_bytes,
        "extension_mrays_s": round(stats["extension_rays"] / per / 1e6, 1),
        "shadow_mrays_s": round(stats["shadow_rays"] / per / 1e6, 1),
        "sequential_gpu_s": round(seq, 3),
        # both legs return host arrays (pyre_render has host outputs only)
        "batched_over_sequential_speedup": round(seq / host_wall, 3), "device": dev_name,
    }
    if args.ref_spp:
        line["relmse_subset"] = batch_relmse(ctx, jobs, args.ref_spp)
    if not args.no_cpu:
        line["cpu_baseline"] = batch_cpu_baseline(jobs, scenes)
    print(json.dumps(line), flush=True)


def batch_relmse(ctx, jobs, ref_spp, n=8):
    """relMSE of the n jobs with the fewest pixels against references of
    ref_spp samples per pixel (FP64 sums, samples 2^20.., one batch call)."""
    import pyre_amd
    order = sorted(range(len(jobs)), key=lambda k: jobs[k].width * jobs[k].height)[:n]
    outs, _, _ = pyre_amd.render_batch(ctx, [jobs[k] for k in order])
    refs = [pyre_amd.Job(jobs[k].scene, jobs[k].camera, jobs[k].integrator, jobs[k].width, jobs[k].height, ref_spp,
                         spp_offset=1 << 20) for k in order]
    _, acc, _ = pyre_amd.render_batch(ctx, refs, accum64=True)
    res = []
    for k, o, a in zip(order, outs, acc):
        res.append({"job": k, "pixels": jobs[k].width * jobs[k].height, "spp": jobs[k].spp,
                    "relmse": round(rel_mse(o[:, :3], a / ref_spp), 6)})
    return {"reference_spp": ref_spp, "relmse_eps": REL_EPS, "jobs": res}


def batch_cpu_baseline(jobs, scenes, n=4):
    """Oracle on the n jobs with the fewest triangles, one after another, at
    1 spp each (per-mesh BVH builds untimed)."""
    from oracle import oracle as O
    threads = O.nthreads()
    order = sorted(range(len(jobs)), key=lambda k: scenes[k][2])[:n]
    samples, secs = 0, 0.0
    for k in order:
        pre, _, _ = scenes[k]
        arr = pre.arrays()
        osc = O.Scene()
        for m in arr["materials"]:
            osc.add_material(O.material(m.kind, m.base_color, m.metallic, m.roughness, m.specular))
        for m in arr["meshes"]:
            osc.add_mesh(m["positions"], m["indices"], m["normals"], m["material_id"])
        e = arr["env"]
        osc.set_env(O.Hdri(e["width"], e["height"], e["rgb"], e["intensity"]))
        cam = jobs[k].camera
        ocam = O.camera(cam.origin, cam.target, cam.up, cam.vfov_deg, cam.aspect)
        t0 = time.perf_counter()
        osc.render(ocam, jobs[k].width, jobs[k].height, 1, max_depth=8, min_rr=3, threads=threads)
        secs += time.perf_counter() - t0
        samples += jobs[k].width * jobs[k].height
    return {"value": round(samples / secs / 1e6, 4), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": f"the {n} jobs with the fewest triangles at 1 spp each, sequentially"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and parity spot check")
    ap.add_argument("--no-nrc", action="store_true",
                    help="traversal: skip the plain-vs-radiance-cache comparison at 1080p")
    ap.add_argument("--config", default="traversal",
                    choices=["traversal", "cornell", "production", "nrc", "cache", "batch"],
                    help="traversal: the headline microbench; cornell: the Cornell render; production: the "
                         "1080p production render; nrc: seconds to relMSE 0.01, plain vs radiance cache; "
                         "cache: the radiance cache's kernels alone; batch: 96 uneven jobs in one batched call")
    ap.add_argument("--reference", default=None, help="production reference .npy (relMSE)")
    ap.add_argument("--make-reference", default=None, help="render the production reference to this .npy")
    ap.add_argument("--ref-spp", type=int, default=16384,
                    help="spp of the FP64 references (0 = skip the relMSE figures of cornell / batch)")
    ap.add_argument("--max-spp", type=int, default=128, help="nrc: largest spp tried")
    ap.add_argument("--nrc-warmup", type=int, default=NRC_WARMUP,
                    help="nrc: timed training-only frames before the first sample")
    ap.add_argument("--nrc-scene", default="production", choices=["production", "cornell"],
                    help="nrc / --make-reference: the view of the cache comparison")
    ap.add_argument("--tuning", action="append", default=[], metavar="FIELD=VALUE",
                    help="traversal: a launch setting of the context (pyre_tuning field), e.g. traversal=3; "
                         "repeatable (A/B runs of the kernel forms)")
    ap.add_argument("--ray-order", default="generated", choices=["generated", "octant-morton"],
                    help="traversal: order of the rays in HBM; octant-morton sorts the closest-hit rays on the "
                         "host (untimed) by direction octant, then the Morton code of the point where each "
                         "enters the scene box (an experiment on ray coherence; the workload's own order is "
                         "'generated')")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path here "
              "(pixel samples are independent, so rays and pixels would shard across GPUs with no exchange)",
              file=sys.stderr)
        sys.exit(2)
    if args.make_reference:
        return make_reference(args)
    if args.config == "cornell":
        return bench_cornell(args)
    if args.config == "production":
        return bench_production(args)
    if args.config == "nrc":
        return bench_nrc(args)
    if args.config == "cache":
        return bench_cache(args)
    if args.config == "batch":
        return bench_batch(args)

    import torch
    import pyre_amd

    ctx = pyre_amd.Context(0)
    dev_name, cus = ctx.device_info()
    log(f"device {dev_name} cus={cus}")
    tuning = {}
    for kv in args.tuning:
        k, v = kv.split("=", 1)
        tuning[k] = float(v) if k == "memory_fraction" else int(v)
    if tuning:
        ctx.set_tuning(**tuning)
        log(f"tuning {ctx.get_tuning()}")
    tris = pyre_amd.random_triangles(1, N_TRIS, 0.01)
    rays32 = pyre_amd.random_rays(2, N_RAYS, N_CLOSEST, "ray32")
    scene = pyre_amd.Scene.from_arrays(
        ctx, [dict(positions=tris.reshape(-1, 3), indices=np.arange(3 * N_TRIS, dtype=np.uint32).reshape(-1, 3),
                   material_id=0)], [pyre_amd.Material(pyre_amd.MAT_LAMBERTIAN, (0.5, 0.5, 0.5))])
    info = scene.build_bvh(bins=BINS, max_leaf=MAX_LEAF)
    build_ms = [info["build_ms"]]
    for _ in range(2):
        build_ms.append(scene.build_bvh(bins=BINS, max_leaf=MAX_LEAF)["build_ms"])
    info = scene.bvh_info()
    log(f"bvh nodes={2 * info['n_leaves'] - 1} leaves={info['n_leaves']} depth={info['depth']} build_ms={build_ms}")
    order = None
    if args.ray_order == "octant-morton":
        order = coherent_order(rays32[:N_CLOSEST], *scene.bounds())
        rays32[:N_CLOSEST] = rays32[:N_CLOSEST][order]
        log("closest-hit rays sorted by octant, then the Morton code of their entry point")

    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        d_rays = torch.from_numpy(rays32).to("cuda")
        closest = d_rays[:N_CLOSEST]
        shadow = d_rays[N_CLOSEST:]
        hits = torch.empty((N_CLOSEST, 2), dtype=torch.int32, device="cuda")
        occ = torch.empty(N_RAYS - N_CLOSEST, dtype=torch.int32, device="cuda")
    sh = stream.cuda_stream
    torch.cuda.synchronize()

    def closest_pass():
        scene.trace_closest_dev(closest.data_ptr(), N_CLOSEST, hits.data_ptr(), sh)

    def any_pass():
        scene.trace_any_dev(shadow.data_ptr(), N_RAYS - N_CLOSEST, occ.data_ptr(), sh)

    for _ in range(args.warmup):
        closest_pass()
        any_pass()
    torch.cuda.synchronize()

    def timed(fn, k):
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(k)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for a, b in evs:
            a.record(stream)
            fn()
            b.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return wall, float(np.mean([a.elapsed_time(b) for a, b in evs]))

    wall_c, kern_c = timed(closest_pass, args.steps)
    log(f"closest: wall {wall_c:.3f}s kernel {kern_c:.3f} ms/launch")
    wall_a, kern_a = timed(any_pass, args.steps)
    log(f"any: wall {wall_a:.3f}s kernel {kern_a:.3f} ms/launch")

    value = N_CLOSEST * args.steps / wall_c / 1e6
    achieved = N_CLOSEST * BYTES_CLOSEST / (kern_c * 1e-3) / 1e9
    tpr = measured_traffic("traffic_bytes_per_ray_closest")
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_c / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded triangle soup and rays; no dataset)",
        "config": {"workload": "raw traversal microbench: 2^20 random triangles, GPU binned-SAH BVH (32 bins, "
                               "leaf<=4), 2^24 rays: 2^23 closest-hit + 2^23 any-hit",
                   "n_triangles": N_TRIS, "n_rays": N_RAYS, "bins": BINS, "max_leaf": MAX_LEAF},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if tpr is None else round(tpr * N_CLOSEST),
                     "bytes_per_ray": round(BYTES_CLOSEST, 1), "kernel_ms": round(kern_c, 4)},
        "anyhit_mrays_s": round((N_RAYS - N_CLOSEST) * args.steps / wall_a / 1e6, 2),
        "anyhit_kernel_ms": round(kern_a, 4),
        "bvh_build_ms": round(min(build_ms), 2),
        "bvh": {"nodes": 2 * info["n_leaves"] - 1, "leaves": info["n_leaves"], "depth": info["depth"]},
        "device": dev_name,
    }
    if tuning:
        line["tuning"] = tuning
    if order is not None:
        line["ray_order"] = args.ray_order
    if not args.no_cpu:
        rays36 = pyre_amd.random_rays(2, N_CLOSEST, N_CLOSEST)
        osc, o, cb = cpu_baseline(tris, rays36, CPU_SAMPLE)
        line["cpu_baseline"] = cb
        # parity spot check on the first 2^16 rays: hit ids and t bit-exact
        h = hits.cpu().numpy()
        if order is not None:  # hits of ray order[k] sit at k
            inv = np.empty_like(order)
            inv[order] = np.arange(len(order))
            h = h[inv]
        h = h[: 1 << 16]
        g_prim = h[:, 0].view(np.uint32)
        g_t = h[:, 1].view(np.uint32)
        o_prim = np.where(o["kind"][: 1 << 16] == 1, o["tri"][: 1 << 16], 0xFFFFFFFF).astype(np.uint32)
        o_t = np.where(o["kind"][: 1 << 16] == 1, o["t"][: 1 << 16], 0).astype(np.float32).view(np.uint32)
        line["parity_spot_check"] = {"rays": 1 << 16, "prim_mismatch": int((g_prim != o_prim).sum()),
                                     "t_mismatch": int(((g_t != o_t) & (g_prim == o_prim)).sum())}
    del d_rays, closest, shadow, hits, occ
    scene.close()
    if not args.no_nrc:
        # the metric's second half: seconds to relMSE 0.01 at 1080p, plain vs
        # the radiance cache, on the production view
        try:
            nrc = nrc_block(ctx, args)
            line["plain_seconds_to_relmse_0.01"] = nrc["plain_seconds"]
            line["nrc_seconds_to_relmse_0.01"] = nrc["nrc_seconds"]
            line["nrc_over_plain"] = nrc["nrc_over_plain"]
            line["nrc_warm_cache_over_plain"] = nrc["warm_cache"]["nrc_over_plain"]
            line["nrc_best_case_over_plain"] = nrc["best_case"]["over_plain"]
            line["nrc"] = nrc
        except Exception as e:  # the traversal line is still reported
            line["nrc_error"] = f"{type(e).__name__}: {e}"
            log(f"nrc block failed: {line['nrc_error']}")
    print(json.dumps(line), flush=True)


def coherent_order(rays, lo, hi):
    """Permutation of ray32 records by (misses the scene box, direction
    octant, 30-bit Morton code of the point where the ray enters the box):
    neighbouring rays start near each other and point the same way."""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    o, d = rays[:, 0:3].astype(np.float64), rays[:, 4:7].astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / d
        t1, t2 = (lo - o) * inv, (hi - o) * inv
    tn = np.maximum(np.nanmax(np.minimum(t1, t2), axis=1), 0.0)
    tf = np.nanmin(np.maximum(t1, t2), axis=1)
    miss = ~(tn <= tf)
    p = o + d * np.where(miss, 0.0, tn)[:, None]
    q = np.clip(((p - lo) / (hi - lo) * 1024).astype(np.int64), 0, 1023).astype(np.uint64)

    def spread(v):
        v = (v | (v << 16)) & 0x30000FF
        v = (v | (v << 8)) & 0x300F00F
        v = (v | (v << 4)) & 0x30C30C3
        return (v | (v << 2)) & 0x9249249

    octant = ((d[:, 0] < 0).astype(np.uint64) << 2) | ((d[:, 1] < 0).astype(np.uint64) << 1) | (d[:, 2] < 0)
    key = (miss.astype(np.uint64) << 33) | (octant.astype(np.uint64) << 30) | \
        (spread(q[:, 0]) << 2) | (spread(q[:, 1]) << 1) | spread(q[:, 2])
    return np.argsort(key, kind="stable")


def copy_probe():
    """A plain device copy of known size (1 GiB read + 1 GiB write), for
    scaling the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": x.numel() * 4}))


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "--copy-probe":
        copy_probe()
        sys.exit(0)
    main()
