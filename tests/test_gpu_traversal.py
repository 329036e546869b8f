"""GPU parity: device BVH build and traversal vs the oracle.

The oracle restates geometry/bvh.rs, geometry/mesh.rs and scene.rs; the
device path goes through the C ABI (pyre_bvh_build / pyre_bvh_export /
pyre_intersect_closest / pyre_intersect_any).
Bar: identical BVH topology and bounds; hit ids and t bit-exact, with rays
whose hit lies within 1e-6 of a triangle edge excluded and counted.
"""
import numpy as np
import pytest

import pyre_amd
from tests.helpers import (canonical_gpu_tree, canonical_oracle_tree, edge_distance, near_edge_hits,
                           tree_stats)

pytestmark = pytest.mark.gpu


def _soup_scene(ctx, tris):
    n = len(tris)
    pos = tris.reshape(-1, 3)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)
    mats = [pyre_amd.Material(pyre_amd.MAT_LAMBERTIAN, (0.5, 0.5, 0.5))]
    return pyre_amd.Scene.from_arrays(ctx, [dict(positions=pos, indices=idx, material_id=0)], mats), pos, idx


@pytest.mark.parametrize("n,bins,leaf,seed", [(1, 12, 8, 5), (7, 12, 8, 6), (3000, 12, 8, 7), (20000, 32, 4, 8),
                                              (70000, 12, 8, 9)])
def test_bvh_topology_matches_reference_build(ctx, oracle, n, bins, leaf, seed):
    print(f"[bvh topology] n={n} bins={bins} leaf={leaf} seed={seed}", flush=True)
    tris = pyre_amd.random_triangles(seed, n, 0.01)
    sc, pos, idx = _soup_scene(ctx, tris)
    info = sc.build_bvh(bins=bins, max_leaf=leaf)
    g = canonical_gpu_tree(sc.bvh_export(n))
    o_sc = oracle.Scene()
    o_sc.add_material(oracle.material(0, (0.5, 0.5, 0.5)))
    o_sc.add_mesh(pos, idx, None, 0, bins, leaf)
    o = canonical_oracle_tree(o_sc.bvh(0))
    assert g == o
    nodes, leaves, depth = tree_stats(o)
    assert info["n_leaves"] == leaves
    assert info["depth"] == depth


def test_bvh_builds_in_sequence(ctx, oracle):
    """Scenes of growing size built one after another in one process, so the
    builder's scratch memory is recycled between builds (a scan sentinel once
    read a stale value there): topology equal to the reference build for the
    smaller ones, node/leaf/depth counts for the larger ones."""
    for lg in (12, 14, 16, 18, 20):
        n = 1 << lg
        print(f"[bvh sequence] n=2^{lg}", flush=True)
        tris = pyre_amd.random_triangles(30 + lg, n, 0.01 * 2 ** ((20 - lg) / 3))
        sc, pos, idx = _soup_scene(ctx, tris)
        info = sc.build_bvh(bins=32, max_leaf=4)
        o_sc = oracle.Scene()
        o_sc.add_material(oracle.material(0, (0.5, 0.5, 0.5)))
        o_sc.add_mesh(pos, idx, None, 0, 32, 4)
        o = canonical_oracle_tree(o_sc.bvh(0))
        nodes, leaves, depth = tree_stats(o)
        assert (info["n_leaves"], info["depth"]) == (leaves, depth)
        if lg <= 16:
            assert canonical_gpu_tree(sc.bvh_export(n)) == o
        sc.close()


def test_bvh_topology_uv_sphere(ctx, oracle):
    """The Cornell preset's gold sphere (main.rs:455), reference parameters."""
    o_cb = oracle.Scene.cornell()
    m = o_cb.mesh(5)
    sc = pyre_amd.Scene.from_arrays(ctx, [dict(positions=m["positions"], indices=m["indices"],
                                               normals=m["normals"], material_id=0)],
                                    [pyre_amd.Material(pyre_amd.MAT_LAMBERTIAN, (0.5, 0.5, 0.5))])
    sc.build_bvh(bins=12, max_leaf=8)
    assert canonical_gpu_tree(sc.bvh_export(len(m["indices"]))) == canonical_oracle_tree(o_cb.bvh(5))


def _closest_parity(ctx, oracle, n_tris, n_rays, bins, leaf, seed):
    tris = pyre_amd.random_triangles(seed, n_tris, 0.01)
    rays = pyre_amd.random_rays(seed + 1000, n_rays, n_rays // 2)
    sc, pos, idx = _soup_scene(ctx, tris)
    sc.build_bvh(bins=bins, max_leaf=leaf)
    g = sc.intersect(rays)
    o_sc = oracle.Scene()
    o_sc.add_material(oracle.material(0, (0.5, 0.5, 0.5)))
    o_sc.add_mesh(pos, idx, None, 0, bins, leaf)
    o = o_sc.intersect(rays)
    o_hit = o["kind"] == 1
    # rays whose oracle hit is within 1e-6 of an edge are excluded (and counted)
    t3 = tris.reshape(-1, 3, 3)
    hi = np.where(o_hit)[0]
    ed = edge_distance(t3[o["tri"][hi], 0], t3[o["tri"][hi], 1], t3[o["tri"][hi], 2], o["u"][hi], o["v"][hi])
    excl = np.zeros(len(rays), bool)
    excl[hi[ed < 1e-6]] = True
    keep = ~excl
    g_hit = g["inst"] != pyre_amd.MISS
    assert np.array_equal(g_hit[keep], o_hit[keep])
    both = keep & o_hit
    assert np.array_equal(g["tri"][both], o["tri"][both])
    assert np.array_equal(g["t"][both].view(np.uint32), o["t"][both].view(np.uint32))
    assert np.array_equal(g["u"][both].view(np.uint32), o["u"][both].view(np.uint32))
    assert np.array_equal(g["v"][both].view(np.uint32), o["v"][both].view(np.uint32))
    assert np.all(g["inst"][both] == 0)
    return int(excl.sum()), int(o_hit.sum())


@pytest.mark.parametrize("n_tris,n_rays,bins,leaf,seed", [(50000, 40000, 32, 4, 11), (20000, 20000, 12, 8, 12)])
def test_closest_hit_bit_exact(ctx, oracle, n_tris, n_rays, bins, leaf, seed):
    print(f"[closest] tris={n_tris} rays={n_rays} bins={bins} leaf={leaf}", flush=True)
    excluded, hits = _closest_parity(ctx, oracle, n_tris, n_rays, bins, leaf, seed)
    print(f"  hits={hits} excluded_near_edge={excluded}")
    assert hits > n_rays // 20
    assert excluded < n_rays // 1000 + 5


def test_any_hit_matches_occluded(ctx, oracle):
    print("[any-hit] tris=50000 rays=40000", flush=True)
    tris = pyre_amd.random_triangles(21, 50000, 0.01)
    rays = pyre_amd.random_rays(22, 40000, 0)  # all rays have t_max ~ U[0.1, 3]
    sc, pos, idx = _soup_scene(ctx, tris)
    sc.build_bvh(bins=32, max_leaf=4)
    g = sc.occluded(rays)
    o_sc = oracle.Scene()
    o_sc.add_material(oracle.material(0, (0.5, 0.5, 0.5)))
    o_sc.add_mesh(pos, idx, None, 0, 32, 4)
    o = o_sc.any_hit(rays)
    # any-hit must agree with closest-hit existence on the same ray (scene.rs:102-140)
    c = sc.intersect(rays)
    assert np.array_equal(g.astype(bool), c["inst"] != pyre_amd.MISS)
    # no mismatch outside rays whose closest hit lies within 1e-6 of an edge
    excl = near_edge_hits(o_sc, rays, [tris.reshape(-1, 3, 3)])
    mism = np.nonzero((g.astype(bool) != o.astype(bool)) & ~excl)[0]
    print(f"  any-hit: excluded near an edge {excl.sum()}, mismatches {len(mism)}")
    assert len(mism) == 0, f"{len(mism)} any-hit mismatches"
    assert excl.sum() < len(rays) // 1000 + 5
    assert 0.05 < g.mean() < 0.95


def _adversarial_rays(seed, n):
    """Rays aimed at the quantised slab test's edge cases (n per kind): axis-
    parallel directions (1/d infinite: an axis without a bound), directions
    with one component below 1e-6 (1/d huge), origins 10^3..10^4 away (the
    widest pads), origins on the root box's faces, and short any-hit ranges."""
    rng = np.random.default_rng(seed)
    out = []
    axis = np.eye(3)[rng.integers(0, 3, n)] * rng.choice([-1.0, 1.0], (n, 1))
    o = rng.uniform(-1.2, 1.2, (n, 3)) - 1.3 * axis
    out.append(np.c_[o, axis, np.full(n, 1e-4), np.full(n, np.inf), np.zeros(n)])
    d = axis + rng.uniform(-1e-6, 1e-6, (n, 3)) * (1 - np.abs(axis))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    out.append(np.c_[rng.uniform(-1.0, 1.0, (n, 3)) - 1.3 * axis, d, np.full(n, 1e-4), np.full(n, np.inf),
                     np.zeros(n)])
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    far = u * rng.uniform(1e3, 1e4, (n, 1))
    aim = rng.uniform(-0.8, 0.8, (n, 3)) - far
    aim /= np.linalg.norm(aim, axis=1, keepdims=True)
    out.append(np.c_[far, aim, np.full(n, 1e-4), np.full(n, np.inf), np.zeros(n)])
    face = rng.uniform(-1.0, 1.0, (n, 3))
    k = rng.integers(0, 3, n)
    face[np.arange(n), k] = rng.choice([-1.0, 1.0], n) * 1.0
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    out.append(np.c_[face, v, np.zeros(n), rng.uniform(0.01, 0.5, n), np.zeros(n)])
    return np.concatenate(out).astype(np.float32)


def test_quantised_boxes_on_edge_case_rays(ctx, oracle):
    """The default trace kernel tests 16-bit quantised child boxes widened by
    a per-axis rounding pad; it must accept every box the exact test accepts.
    On rays chosen for the pad's edge cases its hits equal the exact-box
    kernel's on every ray and the reference's bit for bit (near-edge rays
    excluded and counted), closest and any."""
    print("[quantised boxes] edge-case rays", flush=True)
    tris = pyre_amd.random_triangles(61, 30000, 0.01)
    rays = _adversarial_rays(62, 4000)
    sc, pos, idx = _soup_scene(ctx, tris)
    sc.build_bvh(bins=32, max_leaf=4)
    g = sc.intersect(rays)
    g_occ = sc.occluded(rays)
    with ctx.tuning(traversal=4):
        e = sc.intersect(rays)
        e_occ = sc.occluded(rays)
    for key in ("inst", "tri", "t", "u", "v"):
        assert np.array_equal(g[key].view(np.uint32), e[key].view(np.uint32)), key
    assert np.array_equal(g_occ, e_occ)
    o_sc = oracle.Scene()
    o_sc.add_material(oracle.material(0, (0.5, 0.5, 0.5)))
    o_sc.add_mesh(pos, idx, None, 0, 32, 4)
    o = o_sc.intersect(rays)
    excl = near_edge_hits(o_sc, rays, [tris.reshape(-1, 3, 3)])
    o_hit = o["kind"] == 1
    keep = ~excl
    g_hit = g["inst"] != pyre_amd.MISS
    assert np.array_equal(g_hit[keep], o_hit[keep])
    both = keep & o_hit
    assert np.array_equal(g["tri"][both], o["tri"][both])
    assert np.array_equal(g["t"][both].view(np.uint32), o["t"][both].view(np.uint32))
    mism = (g_occ.astype(bool) != o_sc.any_hit(rays).astype(bool)) & keep
    print(f"  hits {int(o_hit.sum())} of {len(rays)}, excluded near an edge {int(excl.sum())}, "
          f"any-hit mismatches {int(mism.sum())}", flush=True)
    assert not mism.any()
    assert o_hit.sum() > len(rays) // 10


def test_quantised_boxes_on_the_cornell_box(ctx, oracle):
    """Axis-aligned, coplanar geometry (the Cornell preset: walls whose
    leaf boxes are flat, blocks touching the floor, two uv spheres; main.rs:
    353-496) through the trace entry points' quantised boxes: hits equal the
    reference's bit for bit away from triangle edges, closest and any-hit,
    and equal the exact-box kernel's on every ray counted here."""
    print("[quantised boxes] cornell preset", flush=True)
    pre = pyre_amd.Preset.cornell()
    sc = pyre_amd.Scene.from_preset(ctx, pre)
    sc.build_bvh()
    o_sc = oracle.Scene.cornell()
    rng = np.random.default_rng(63)
    n = 20000
    o = rng.uniform(-0.99, 0.99, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    closest = np.c_[o, d, np.full(n, 1e-4), np.full(n, np.inf), np.zeros(n)].astype(np.float32)
    target = np.c_[rng.uniform(-0.3, 0.3, n), np.full(n, 0.999), rng.uniform(-0.3, 0.3, n)]
    sd = target - o
    dist = np.linalg.norm(sd, axis=1)
    shadow = np.c_[o, sd / dist[:, None], np.full(n, 1e-3), dist - 1e-3, np.zeros(n)].astype(np.float32)
    rays = np.concatenate([closest, shadow])
    g = sc.intersect(rays)
    g_occ = sc.occluded(rays)
    with ctx.tuning(traversal=4):
        e = sc.intersect(rays)
        e_occ = sc.occluded(rays)
    same = np.ones(len(rays), bool)
    for key in ("inst", "tri", "t", "u", "v"):
        same &= g[key].view(np.uint32) == e[key].view(np.uint32)
    r = o_sc.intersect(rays)
    tri_arrays = []
    for p in range(int(r["id"][r["kind"] == 1].max()) + 1):
        m = o_sc.mesh(p)
        tri_arrays.append(m["positions"][m["indices"].reshape(-1)].reshape(-1, 3, 3))
    excl = near_edge_hits(o_sc, rays, tri_arrays)
    # the trace entry points see triangles only: rays whose reference hit is
    # the quad light (HitKind::Light, scene.rs:36-45) are left out
    keep = ~excl & (r["kind"] != 2)
    r_hit = r["kind"] == 1
    both = keep & r_hit
    print(f"  hits {int(r_hit.sum())} of {len(rays)}, excluded near an edge {int(excl.sum())}, light hits "
          f"{int((r['kind'] == 2).sum())}, differing from the exact boxes {int((~same).sum())}", flush=True)
    assert np.array_equal((g["inst"] != pyre_amd.MISS)[keep], r_hit[keep])
    assert np.array_equal(g["inst"][both], r["id"][both])
    assert np.array_equal(g["tri"][both], r["tri"][both])
    assert np.array_equal(g["t"][both].view(np.uint32), r["t"][both].view(np.uint32))
    assert not ((g_occ.astype(bool) != o_sc.any_hit(rays).astype(bool)) & keep).any()
    assert same[keep].all() and np.array_equal(g_occ[keep], e_occ[keep])
    assert r_hit.sum() > n // 2


def test_empty_and_degenerate_inputs(ctx):
    tris = pyre_amd.random_triangles(31, 100, 0.01)
    sc, _, _ = _soup_scene(ctx, tris)
    sc.build_bvh(bins=12, max_leaf=8)
    out = sc.intersect(np.zeros((0, 9), np.float32))
    assert len(out["t"]) == 0
    # zero-length range and t_max < t_min never hit; NaN direction never hits
    rays = np.array([[0, 0, -5, 0, 0, 1, 1e-4, 1e-4, 0],
                     [0, 0, -5, 0, 0, 1, 1.0, 0.5, 0],
                     [0, 0, -5, np.nan, 0, 1, 1e-4, np.inf, 0]], np.float32)
    h = sc.intersect(rays)
    assert np.all(h["inst"] == pyre_amd.MISS)
    assert np.all(sc.occluded(rays) == 0)


def test_degenerate_coincident_triangles(ctx, oracle):
    """Coincident centroids force the extent < 1e-9 leaf (bvh.rs:205-208)."""
    tri = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32)
    tris = np.tile(tri, (40, 1))
    sc, pos, idx = _soup_scene(ctx, tris)
    sc.build_bvh(bins=12, max_leaf=8)
    g = canonical_gpu_tree(sc.bvh_export(40))
    o_sc = oracle.Scene()
    o_sc.add_material(oracle.material(0, (0.5, 0.5, 0.5)))
    o_sc.add_mesh(pos, idx, None, 0, 12, 8)
    assert g == canonical_oracle_tree(o_sc.bvh(0))
    r = np.array([[0.2, 0.2, 1, 0, 0, -1, 1e-4, np.inf, 0]], np.float32)
    h = sc.intersect(r)
    assert h["inst"][0] == 0 and h["tri"][0] == 0  # exact tie -> lowest triangle id
    assert h["t"][0] == np.float32(1.0)


@pytest.mark.parametrize("switch", [dict(traversal=1), dict(traversal=2), dict(traversal=3), dict(traversal=4),
                                    dict(stack_ring=4), dict(stack_ring=16), dict(trace_blocks_per_cu=2)])
def test_traversal_variants_bit_exact(ctx, oracle, switch):
    """Every selectable traversal form (binary tree with an LDS stack;
    grid-stride persistent; leaf triangles tested one per step; the exact
    float boxes instead of the default's quantised ones; LDS stack rings of 4
    and 16; a small persistent grid) gives the reference's hits bit for bit,
    closest and any."""
    print(f"[traversal variant] {switch}", flush=True)
    with ctx.tuning(**switch):
        excluded, hits = _closest_parity(ctx, oracle, 30000, 20000, 32, 4, 41)
        assert hits > 1000 and excluded < 25
        tris = pyre_amd.random_triangles(43, 30000, 0.01)
        rays = pyre_amd.random_rays(44, 20000, 0)
        sc, pos, idx = _soup_scene(ctx, tris)
        sc.build_bvh(bins=32, max_leaf=4)
        g = sc.occluded(rays).astype(bool)
        c = sc.intersect(rays)
        assert np.array_equal(g, c["inst"] != pyre_amd.MISS)
        sc.close()
