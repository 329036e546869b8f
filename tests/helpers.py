"""Shared helpers for the parity tests (oracle vs HIP path)."""
from __future__ import annotations

import numpy as np


def canonical_oracle_tree(t, i=0):
    """Oracle BVH (DFS, left = i + 1, right = offset) -> nested tuples."""
    b = tuple(t["bounds"][i].tolist())
    if t["count"][i] > 0:
        s, n = int(t["offset"][i]), int(t["count"][i])
        return ("L", b, tuple(sorted(t["prim"][s:s + n].tolist())))
    return ("N", b, int(t["axis"][i]), canonical_oracle_tree(t, i + 1), canonical_oracle_tree(t, int(t["offset"][i])))


def canonical_gpu_tree(t, i=0):
    """Exported device BVH (explicit child ids) -> nested tuples."""
    b = tuple(t["bounds"][i].tolist())
    if t["is_leaf"][i] == 1:
        s, n = int(t["a"][i]), int(t["b"][i])
        return ("L", b, tuple(sorted(t["prim"][s:s + n].tolist())))
    assert t["is_leaf"][i] == 0, f"node {i} unused"
    return ("N", b, int(t["axis"][i]), canonical_gpu_tree(t, int(t["a"][i])), canonical_gpu_tree(t, int(t["b"][i])))


def tree_stats(tree):
    """(nodes, leaves, depth) of a canonical tree."""
    if tree[0] == "L":
        return 1, 1, 0
    a = tree_stats(tree[3])
    b = tree_stats(tree[4])
    return 1 + a[0] + b[0], a[1] + b[1], 1 + max(a[2], b[2])


def edge_distance(v0, v1, v2, u, v):
    """Distance (float64) from the barycentric point to the nearest triangle edge."""
    v0, v1, v2 = (np.asarray(x, np.float64) for x in (v0, v1, v2))
    u = np.asarray(u, np.float64)
    v = np.asarray(v, np.float64)
    w = 1.0 - u - v
    area2 = np.linalg.norm(np.cross(v1 - v0, v2 - v0), axis=-1)
    e_opp0 = np.linalg.norm(v2 - v1, axis=-1)  # opposite v0 (weight w)
    e_opp1 = np.linalg.norm(v2 - v0, axis=-1)  # opposite v1 (weight u)
    e_opp2 = np.linalg.norm(v1 - v0, axis=-1)  # opposite v2 (weight v)
    with np.errstate(divide="ignore", invalid="ignore"):
        d = np.minimum(np.minimum(w * area2 / e_opp0, u * area2 / e_opp1), v * area2 / e_opp2)
    return np.nan_to_num(d, nan=0.0)


def near_edge_hits(o_scene, rays, tri_arrays):
    """Rays whose oracle closest hit in (t_min, t_max) lies within 1e-6 of a
    triangle edge (tri_arrays[p] = (n, 3, 3) vertices of primitive p), as a
    bool mask: excluded from any-hit parity and counted."""
    o = o_scene.intersect(rays)
    hit = np.where(o["kind"] == 1)[0]
    ed = np.empty(len(hit))
    for p, t3 in enumerate(tri_arrays):
        sel = o["id"][hit] == p
        tri = o["tri"][hit][sel]
        ed[sel] = edge_distance(t3[tri, 0], t3[tri, 1], t3[tri, 2], o["u"][hit][sel], o["v"][hit][sel])
    mask = np.zeros(len(rays), bool)
    mask[hit[ed < 1e-6]] = True
    return mask


def rel_err(a, b, floor=1e-30):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), floor)


def rel_mse(img, ref, eps=1e-2):
    img = np.asarray(img, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + eps)))


def oracle_camera(oracle, cam):
    """The oracle's camera for a pyre_amd.Camera."""
    return oracle.camera(cam.origin, cam.target, cam.up, cam.vfov_deg, cam.aspect, cam.aperture_radius,
                         cam.focus_distance if cam.focus_distance > 0 else None)


def oracle_scene(oracle, pre):
    """The oracle's Scene built from a preset's arrays: its meshes (with
    their motion), materials, quad lights and environment map."""
    arr = pre.arrays()
    s = oracle.Scene()
    for m in arr["materials"]:
        s.add_material(oracle.material(m.kind, m.base_color, m.metallic, m.roughness, m.specular))
    for m in arr["meshes"]:
        prim = s.add_mesh(m["positions"], m["indices"], m["normals"], m["material_id"])
        if m.get("motion") is not None:
            s.set_motion(prim, m["motion"][0], m["motion"][1])
    for l in arr["lights"]:
        s.add_light(l)
    if arr["env"] is not None:
        e = arr["env"]
        s.set_env(oracle.Hdri(e["width"], e["height"], e["rgb"], e["intensity"]))
    return s


def assert_render_parity(g, o, exact):
    """GPU vs oracle PixelSample images (h, w, 10) of the same samples:
    exact -> every field of >= 99 % of the pixels bit-equal (scenes without
    an environment map or thin lens); otherwise first-hit AOVs bit-equal and
    radiance within 1e-5 relative on >= 99 % of the pixels. Returns the
    fractions."""
    same = np.all(g.view(np.uint32) == o.view(np.uint32), axis=-1).mean()
    aov = np.all(g[..., 3:].view(np.uint32) == o[..., 3:].view(np.uint32), axis=-1).mean()
    rel = (np.abs(g[..., :3] - o[..., :3]) / np.maximum(np.abs(o[..., :3]), 1e-6)).max(-1)
    close = float(np.mean(rel <= 1e-5))
    if exact:
        assert same >= 0.99, f"bit-equal pixels {same:.4f}"
    else:
        assert aov >= 0.95, f"AOV bit-equal pixels {aov:.4f}"
        assert close >= 0.99, f"radiance within 1e-5 on {close:.4f} of the pixels"
    assert rel_mse(g[..., :3], o[..., :3]) < 1e-8
    return dict(bit_equal=float(same), aov_bit_equal=float(aov), radiance_1e5=close)

