"""Oracle pins for the emissive triangle mesh light (the general emitter that
the reference's PBRT builder drops, io/pbrt/builder.rs:463-485,513-516, put
behind the Light trait of light.rs:36-42).

The reference has no implementation to compare with, so the oracle's
MeshLight is pinned by what the mathematics fixes:
- the mean of 1/pdf over its samples is the solid angle of its front-facing
  triangles (Van Oosterom-Strackee), which fails on a dropped cos or dist^2
  term, a non-uniform point map or a wrong area;
- triangles are chosen in proportion to their area;
- pdf(from, wi) at a sampled direction equals the sample's pdf;
- two triangles of a quad behave as DiffuseAreaQuadLight (light.rs:46-162);
- emission is one-sided;
- the Cornell box lit by its light as a two-triangle mesh renders to the
  same expected image as with the quad light.
"""
import numpy as np
import pytest


def solid_angle(p, a, b, c):
    """Van Oosterom & Strackee (1983): solid angle of triangle abc seen from p."""
    r1, r2, r3 = (np.asarray(v, np.float64) - np.asarray(p, np.float64) for v in (a, b, c))
    l1, l2, l3 = (np.linalg.norm(r) for r in (r1, r2, r3))
    num = abs(np.dot(r1, np.cross(r2, r3)))
    den = l1 * l2 * l3 + np.dot(r1, r2) * l3 + np.dot(r1, r3) * l2 + np.dot(r2, r3) * l1
    return 2.0 * np.arctan2(num, den)


def octahedron(center, radius):
    c = np.asarray(center, np.float32)
    v = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32) * radius + c
    # outward-facing (counter-clockwise seen from outside)
    f = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]], np.uint32)
    return v, f


def strata(n, seed):
    rng = np.random.default_rng(seed)
    k = int(np.sqrt(n))
    g = (np.stack(np.meshgrid(np.arange(k), np.arange(k), indexing="ij"), -1).reshape(-1, 2) + rng.random((k * k, 2)))
    return (g / k).astype(np.float32)


def test_single_triangle_solid_angle_and_pdf(oracle):
    v = np.array([[-0.5, 0.2, 1.0], [0.7, -0.1, 1.2], [0.1, 0.8, 0.9]], np.float32)
    # orient the triangle toward the origin
    n = np.cross(v[1] - v[0], v[2] - v[0])
    idx = np.array([[0, 1, 2]] if np.dot(n, -v[0]) > 0 else [[0, 2, 1]], np.uint32)
    L = oracle.MeshLight(v, idx, (2.0, 3.0, 4.0))
    assert L.area == pytest.approx(0.5 * np.linalg.norm(n), rel=1e-6)
    u = strata(250_000, 1)
    p = np.zeros((len(u), 3), np.float32)
    pos, wi, dist, pdf, ok, tri = L.sample(p, u)
    assert ok.all()
    est = np.mean(1.0 / pdf.astype(np.float64))
    assert est == pytest.approx(solid_angle((0, 0, 0), *v), rel=2e-3)
    # sampled points lie in the triangle: barycentrics in [0, 1]
    e1, e2 = (v[1] - v[0]).astype(np.float64), (v[2] - v[0]).astype(np.float64)
    if idx[0, 1] == 2:
        e1, e2 = e2, e1
    d = pos.astype(np.float64) - v[0]
    m = np.stack([e1, e2], 1)
    bc, *_ = np.linalg.lstsq(m, d.T, rcond=None)
    assert (bc > -1e-5).all() and (bc.sum(0) < 1 + 1e-5).all()
    assert np.abs(m @ bc - d.T).max() < 1e-5
    # uniform in area: the sampled points' mean is the centroid
    assert np.allclose(pos.astype(np.float64).mean(0), v.astype(np.float64).mean(0), atol=2e-3)
    # wi and dist reach the sampled point; pdf(from, wi) equals the sample's pdf
    assert np.allclose(p + wi * dist[:, None], pos, atol=1e-5)
    q = L.pdf(p[:4096], wi[:4096])
    assert np.allclose(q, pdf[:4096], rtol=2e-4)


def test_octahedron_front_faces_and_area_selection(oracle):
    v, f = octahedron((0.3, -0.2, 2.5), 0.6)
    L = oracle.MeshLight(v, f, (1.0, 1.0, 1.0))
    p0 = np.array([0.0, 0.0, 0.0], np.float32)
    u = strata(250_000, 2)
    p = np.repeat(p0[None], len(u), 0)
    pos, wi, dist, pdf, ok, tri = L.sample(p, u)
    front = []
    for t in f:
        a, b, c = v[t]
        n = np.cross(b - a, c - a)
        front.append(np.dot(n, p0 - a) > 0)
    front = np.array(front)
    # back-facing triangles never give a sample; every front-facing one does
    assert not ok[~front[tri]].any() and ok[front[tri]].all()
    omega = sum(solid_angle(p0, *v[t]) for t, fr in zip(f, front) if fr)
    est = np.mean(np.where(ok, 1.0 / np.where(ok, pdf, 1).astype(np.float64), 0.0))
    assert est == pytest.approx(omega, rel=3e-3)
    # pdf at the sampled directions: the nearest hit along wi is the sampled point (convex emitter)
    sel = np.where(ok)[0][:4096]
    assert np.allclose(L.pdf(p[sel], wi[sel]), pdf[sel], rtol=2e-4)
    # equal areas: every triangle chosen about 1/8 of the time
    counts = np.bincount(tri, minlength=8)
    assert np.all(np.abs(counts - len(u) / 8) < 5 * np.sqrt(len(u) / 8))


def test_unequal_areas_selected_by_area(oracle):
    v = np.array([[0, 0, 1], [1, 0, 1], [0, 1, 1], [2, 0, 1], [5, 0, 1], [2, 3, 1]], np.float32)
    f = np.array([[0, 2, 1], [3, 5, 4]], np.uint32)  # areas 0.5 and 4.5, facing -z
    L = oracle.MeshLight(v, f, (1.0, 1.0, 1.0))
    assert L.area == pytest.approx(5.0, rel=1e-6)
    u = strata(200_000, 3)
    _, _, _, _, ok, tri = L.sample(np.zeros((len(u), 3), np.float32), u)
    assert ok.all()
    frac = np.mean(tri == 1)
    assert frac == pytest.approx(0.9, abs=5 * np.sqrt(0.09 / len(u)))


def test_quad_as_two_triangles_equals_quad_light(oracle):
    p0, eu, ev = np.array([-0.3, 0.999, -0.3], np.float32), np.array([0.6, 0, 0], np.float32), \
        np.array([0, 0, 0.6], np.float32)
    quad = np.concatenate([p0, eu, ev, [15, 15, 15]]).astype(np.float32)
    verts = np.array([p0, p0 + eu, p0 + eu + ev, p0 + ev], np.float32)
    L = oracle.MeshLight(verts, np.array([[0, 1, 2], [0, 2, 3]], np.uint32), (15, 15, 15))
    n_quad = np.cross(eu, ev) / np.linalg.norm(np.cross(eu, ev))
    assert L.area == pytest.approx(0.36, rel=1e-6)
    rng = np.random.default_rng(4)
    frm = (rng.random((20000, 3)) * [1.6, 1.0, 1.6] - [0.8, 1.0, 0.8]).astype(np.float32)  # below the light
    tgt = verts[0] + rng.random((20000, 1)).astype(np.float32) * eu + rng.random((20000, 1)).astype(np.float32) * ev
    wi = tgt - frm
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    q = oracle.quad_pdf(quad, frm, wi)
    m = L.pdf(frm, wi)
    good = q > 0
    assert good.mean() > 0.99
    assert np.allclose(m[good], q[good], rtol=1e-5)
    # the quad emits along eu x ev; both see nothing from above
    above = frm + np.array([0, 2.0, 0], np.float32)
    _, _, _, _, ok, _ = L.sample(above, rng.random((20000, 2)).astype(np.float32))
    if n_quad[1] < 0:
        assert not ok.any()
    else:
        assert ok.all()


def _cornell_with(oracle, as_mesh):
    import pyre_amd
    pre = pyre_amd.Preset.cornell().arrays()
    sc = oracle.Scene()
    for m in pre["materials"]:
        sc.add_material(oracle.material(m.kind, m.base_color, m.metallic, m.roughness, m.specular))
    for m in pre["meshes"]:
        sc.add_mesh(m["positions"], m["indices"], m["normals"], m["material_id"])
    for l in pre["lights"]:
        if as_mesh:
            p0, eu, ev = l[0:3], l[3:6], l[6:9]
            verts = np.array([p0, p0 + eu, p0 + eu + ev, p0 + ev], np.float32)
            sc.add_mesh_light(verts, np.array([[0, 1, 2], [0, 2, 3]], np.uint32), l[9:12])
        else:
            sc.add_light(l)
    return sc


@pytest.mark.slow
def test_cornell_mesh_light_renders_the_quad_lights_image(oracle):
    import pyre_amd
    cam = pyre_amd.Preset.cornell().camera(1.0)
    ocam = oracle.camera(cam.origin, cam.target, cam.up, cam.vfov_deg, cam.aspect)
    w = h = 24
    spp = 96
    a = _cornell_with(oracle, False).render(ocam, w, h, spp, max_depth=4).reshape(h, w, 10)
    b = _cornell_with(oracle, True).render(ocam, w, h, spp, max_depth=4).reshape(h, w, 10)
    # first-hit AOVs agree wherever the camera ray does not end on the light
    assert np.mean(np.all(a[..., 3:] == b[..., 3:], -1)) > 0.95
    # the same expectation: image means within a few standard errors
    ma, mb = a[..., :3].mean(), b[..., :3].mean()
    se = np.sqrt(a[..., :3].var() / a[..., :3].size + b[..., :3].var() / b[..., :3].size) * np.sqrt(spp)
    assert abs(ma - mb) < max(5 * se, 0.01 * ma), (ma, mb, se)
