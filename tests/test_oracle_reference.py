"""The reference's own hot-path unit tests, restated against the oracle.

    geometry/mesh.rs:197-234   hits/misses/closest-of-two
    material.rs:332-387        Disney sample/eval/pdf consistency
    distribution.rs:178-220    Distribution1D/2D
plus the published test vectors of the third-party generators the reference
samples with (rand_xoshiro 0.7 xoshiro256++ / SplitMix64, rand 0.9 f32).
"""
import numpy as np

import pytest


def _tri_scene(O, positions, indices):
    sc = O.Scene()
    sc.add_material(O.material(0, (0.5, 0.5, 0.5)))
    sc.add_mesh(np.asarray(positions, np.float32), np.asarray(indices, np.uint32), None, 0)
    return sc


def _ray(o, d, tmin=1e-4, tmax=np.inf):
    d = np.asarray(d, np.float32)
    d = d * np.float32(1.0) / np.sqrt(np.float32(d @ d))  # Ray::new normalises
    return np.array([*o, *d, tmin, tmax, 0.0], np.float32)


XY_TRI = ([[-1, 0, 0], [1, 0, 0], [0, 1, 0]], [[0, 1, 2]])


def test_hits_triangle_along_minus_z(oracle):  # mesh.rs:197-204
    sc = _tri_scene(oracle, *XY_TRI)
    h = sc.intersect(_ray((0, 0.3, 1), (0, 0, -1))[None], with_pn=True)
    assert h["kind"][0] == 1
    assert abs(h["t"][0] - 1.0) < 1e-4
    assert h["normal"][0] @ np.array([0, 0, 1], np.float32) > 0.99


def test_misses_triangle_below(oracle):  # mesh.rs:206-211
    sc = _tri_scene(oracle, *XY_TRI)
    assert sc.intersect(_ray((0, -1, 1), (0, 0, -1))[None])["kind"][0] == 0


def test_bvh_finds_closest_of_two(oracle):  # mesh.rs:213-234
    pos = [[-1, -1, 0], [1, -1, 0], [0, 1, 0], [-1, -1, -1], [1, -1, -1], [0, 1, -1]]
    sc = _tri_scene(oracle, pos, [[0, 1, 2], [3, 4, 5]])
    h = sc.intersect(_ray((0, 0, 2), (0, 0, -1))[None])
    assert abs(h["t"][0] - 2.0) < 1e-4
    assert h["tri"][0] == 0


DISNEY_CASES = [((0.7, 0.3, 0.2), 0.0, 0.4, 0.5), ((1.0, 0.766, 0.336), 1.0, 0.1, 0.5), ((0.5, 0.5, 0.5), 0.5, 0.7, 0.5)]


def test_disney_sample_eval_pdf_consistency(oracle):  # material.rs:332-387
    """Same draws as the reference test: one Xoshiro256PlusPlus stream seeded
    0xDEADBEEF, five f32 per iteration (cos_t, phi, u.x, u.y)."""
    O = oracle
    stream = O.rng_f32(0xDEADBEEF, 3 * 500 * 4)
    k = 0
    checked = 0
    for base, m, r, s in DISNEY_CASES:
        mat = O.material(1, base, m, r, s)
        for _ in range(500):
            cos_t = np.float32(stream[k] * np.float32(0.99) + np.float32(0.01))
            phi = np.float32(stream[k + 1] * np.float32(2 * np.pi))
            u = np.array([stream[k + 2], stream[k + 3]], np.float32)
            k += 4
            sin_t = np.sqrt(np.float32(1) - cos_t * cos_t)
            wo = np.array([sin_t * np.cos(phi), sin_t * np.sin(phi), cos_t], np.float32)
            wi, f, pdf, ok = O.bsdf_sample(mat, wo, u)
            if not ok[0]:
                continue
            fc, pc = O.bsdf_eval(mat, wo, wi[0])
            assert np.linalg.norm(fc[0] - f[0]) < 1e-5
            assert abs(pc[0] - pdf[0]) < 1e-5
            assert pdf[0] > 0 and wi[0, 2] > 0
            checked += 1
    assert checked > 1000


def test_distribution1d_uniform_samples_uniformly(oracle):  # distribution.rs:178-190
    d = oracle.Distribution1D([1.0] * 4)
    t, pdf, b = d.sample_continuous([0.0, 0.999999])
    assert abs(t[0]) < 1e-6 and abs(pdf[0] - 1) < 1e-6 and b[0] == 0
    assert 0.75 <= t[1] < 1.0 and abs(pdf[1] - 1) < 1e-3 and b[1] == 3


def test_distribution1d_concentrates_mass(oracle):  # distribution.rs:192-201
    d = oracle.Distribution1D([0.05, 0.9, 0.05])
    t, pdf, b = d.sample_continuous(0.5)
    assert b[0] == 1 and 1 / 3 <= t[0] < 2 / 3
    assert abs(pdf[0] - 2.7) < 1e-3


def test_distribution2d_pdf_normalizes(oracle):  # distribution.rs:203-220
    d = oracle.Distribution2D(np.arange(16, dtype=np.float32) + 1, 4, 4)
    n = 32
    g = (np.arange(n, dtype=np.float32) + np.float32(0.5)) / np.float32(n)
    uv = np.stack(np.meshgrid(g, g, indexing="xy"), -1).reshape(-1, 2)
    total = float(d.pdf_at(uv).astype(np.float32).sum()) / (n * n)
    assert abs(total - 1.0) < 1e-2


def test_xoshiro256pp_published_vectors(oracle):
    """xoshiro256++ from state {1, 2, 3, 4} (Blackman & Vigna reference
    implementation; the first two values also follow by hand:
    rotl(1 + 4, 23) + 1 = 41943041)."""
    out = oracle.rng_u64_from_state([1, 2, 3, 4], 10)
    assert [int(x) for x in out] == [41943041, 58720359, 3588806011781223, 3591011842654386,
                                     9228616714210784205, 9973669472204895162, 14011001112246962877,
                                     12406186145184390807, 15849039046786891736, 10450023813501588000]


def test_splitmix64_seed_fill(oracle):
    """seed_from_u64 fills the state with SplitMix64 outputs; seed 0 gives the
    published first outputs of splitmix64.c."""
    assert [int(x) for x in oracle.rng_state(0)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                                     0x06C45D188009454F, 0xF88BB8A8724C81EC]


def test_f32_from_upper_bits(oracle):
    """random::<f32>() = (next_u64 >> 32 >> 8) * 2^-24 on the same stream."""
    u = oracle.rng_u64(12345, 64)
    f = oracle.rng_f32(12345, 64)
    expect = ((u >> np.uint64(40)).astype(np.float64) * 2.0 ** -24).astype(np.float32)
    assert np.array_equal(f, expect)
    assert f.min() >= 0 and f.max() < 1


def test_pixel_seed_spreads_and_is_deterministic(oracle):
    s = {oracle.pixel_seed(x, y, k) for x in range(8) for y in range(8) for k in range(8)}
    assert len(s) == 512
    assert oracle.pixel_seed(3, 5, 7) == oracle.pixel_seed(3, 5, 7)
    # closed form at the origin: h = 0 -> the finaliser of 0 is 0
    assert oracle.pixel_seed(0, 0, 0) == 0


@pytest.mark.parametrize("bins,leaf", [(12, 8), (32, 4)])
def test_bvh_invariants(oracle, bins, leaf):
    """Every triangle in exactly one leaf, node bounds contain children, leaf
    sizes obey the leaf rule, depth below the 64-entry stack (bvh.rs:99)."""
    rng = np.random.default_rng(bins)
    n = 4000
    c = rng.uniform(-1, 1, (n, 1, 3))
    tris = (c + rng.normal(0, 0.02, (n, 3, 3))).astype(np.float32)
    sc = _tri_scene(oracle, tris.reshape(-1, 3), np.arange(3 * n).reshape(-1, 3))
    t = sc.bvh(0)
    seen = np.zeros(n, int)

    def walk(i, depth):
        b = t["bounds"][i]
        if t["count"][i] > 0:
            s, k = t["offset"][i], t["count"][i]
            assert k <= leaf
            prims = t["prim"][s:s + k]
            seen[prims] += 1
            lo = tris[prims].reshape(-1, 3).min(0)
            hi = tris[prims].reshape(-1, 3).max(0)
            assert np.array_equal(lo, b[:3]) and np.array_equal(hi, b[3:])
            return depth
        l, r = i + 1, t["offset"][i]
        for ch in (l, r):
            cb = t["bounds"][ch]
            assert np.all(cb[:3] >= b[:3]) and np.all(cb[3:] <= b[3:])
        return max(walk(l, depth + 1), walk(r, depth + 1))

    depth = walk(0, 0)
    assert np.all(seen == 1)
    assert depth == t["depth"] and depth < 63


def test_bvh_closest_equals_brute_force(oracle):
    rng = np.random.default_rng(5)
    n = 3000
    c = rng.uniform(-1, 1, (n, 1, 3))
    tris = (c + rng.normal(0, 0.05, (n, 3, 3))).astype(np.float32)
    sc = _tri_scene(oracle, tris.reshape(-1, 3), np.arange(3 * n).reshape(-1, 3))
    o = rng.uniform(-1.5, 1.5, (3000, 3)).astype(np.float32)
    d = rng.normal(size=(3000, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d, np.full((3000, 1), 1e-4), np.full((3000, 1), np.inf), np.zeros((3000, 1))],
                          1).astype(np.float32)
    h = sc.intersect(rays)
    tri, t, ties = sc.brute_closest(rays)
    ok = ties <= 1
    assert ok.mean() > 0.99
    hit = (h["kind"] == 1) & ok
    assert np.array_equal(h["kind"][ok] == 1, tri[ok] != 0xFFFFFFFF)
    assert np.array_equal(h["tri"][hit], tri[hit])
    assert np.array_equal(h["t"][hit], t[hit])
    # any-hit (scene.rs:102-140) is closest-hit existence
    assert np.array_equal(sc.any_hit(rays).astype(bool), h["kind"] == 1)


def test_scene_tie_goes_to_lower_primitive(oracle):
    """Two coincident quads in different primitives: scene.rs:65 keeps the
    first primitive on an exact tie."""
    sc = oracle.Scene()
    sc.add_material(oracle.material(0, (0.5, 0.5, 0.5)))
    quad = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    sc.add_mesh(quad, idx, None, 0)
    sc.add_mesh(quad, idx, None, 0)
    h = sc.intersect(_ray((0.3, 0.2, 1), (0, 0, -1))[None])
    assert h["kind"][0] == 1 and h["id"][0] == 0
