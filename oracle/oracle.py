"""TEST INFRASTRUCTURE ONLY: ctypes view of the oracle (liboracle.so).

The oracle is a CPU restatement of the Pyre reference (crates/pyre/src/*.rs,
crates/pyre-cli/src/main.rs); see pyre_oracle.hpp for the rules it follows.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module. The product (pyre_amd) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_u32 = C.c_uint32


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        sig = {
            "orc_pixel_seed": (C.c_uint64, [_u32, _u32, _u32]),
            "orc_rng_state": (None, [C.c_uint64, _u64p]),
            "orc_rng_f32": (None, [C.c_uint64, _u32, _f32p]),
            "orc_rng_u64": (None, [C.c_uint64, _u32, _u64p]),
            "orc_rng_u64_from_state": (None, [_u64p, _u32, _u64p]),
            "orc_concentric_disk": (None, [_f32p, _u32, _f32p]),
            "orc_dist1d_new": (_vp, [_f32p, _u32]),
            "orc_dist1d_free": (None, [_vp]),
            "orc_dist1d_cdf": (None, [_vp, _f32p]),
            "orc_dist1d_integral": (C.c_float, [_vp]),
            "orc_dist1d_sample": (None, [_vp, _f32p, _u32, _f32p, _f32p, _u32p]),
            "orc_dist1d_pdf": (None, [_vp, _f32p, _u32, _f32p]),
            "orc_dist2d_new": (_vp, [_f32p, _u32, _u32]),
            "orc_dist2d_free": (None, [_vp]),
            "orc_dist2d_sample": (None, [_vp, _f32p, _u32, _f32p, _f32p]),
            "orc_dist2d_pdf": (None, [_vp, _f32p, _u32, _f32p]),
            "orc_bsdf_eval": (None, [_f32p, _f32p, _f32p, _u32, _f32p, _f32p]),
            "orc_bsdf_sample": (None, [_f32p, _f32p, _f32p, _u32, _f32p, _f32p, _f32p, _u8p]),
            "orc_quad_sample": (None, [_f32p, _f32p, _f32p, _u32, _f32p, _f32p, _f32p, _f32p, _u8p]),
            "orc_quad_pdf": (None, [_f32p, _f32p, _f32p, _u32, _f32p]),
            "orc_hdri_new": (_vp, [_u32, _u32, _f32p, C.c_float]),
            "orc_hdri_gradient": (_vp, [_f32p, _f32p, C.c_float]),
            "orc_hdri_free": (None, [_vp]),
            "orc_hdri_texels": (None, [_vp, _u32p, _vp]),
            "orc_hdri_sample": (None, [_vp, _f32p, _u32, _f32p, _f32p, _f32p]),
            "orc_hdri_pdf": (None, [_vp, _f32p, _u32, _f32p]),
            "orc_hdri_le": (None, [_vp, _f32p, _u32, _f32p]),
            "orc_hdri_uv": (None, [_f32p, _u32, _f32p]),
            "orc_hdri_dir": (None, [_f32p, _u32, _f32p]),
            "orc_scene_new": (_vp, []),
            "orc_scene_free": (None, [_vp]),
            "orc_scene_cornell": (_vp, [C.c_int, _u32, _u32]),
            "orc_scene_add_material": (_u32, [_vp, _f32p]),
            "orc_scene_add_mesh": (_u32, [_vp, _f32p, _u32, _u32p, _u32, _vp, _u32, _u32, _u32]),
            "orc_scene_set_motion": (None, [_vp, _u32, _f32p, _f32p]),
            "orc_scene_add_light": (None, [_vp, _f32p]),
            "orc_scene_add_mesh_light": (None, [_vp, _f32p, _u32, _u32p, _u32, _f32p]),
            "orc_mesh_light_new": (_vp, [_f32p, _u32, _u32p, _u32, _f32p]),
            "orc_mesh_light_free": (None, [_vp]),
            "orc_mesh_light_area": (C.c_float, [_vp]),
            "orc_mesh_light_sample": (None, [_vp, _f32p, _f32p, _u32, _u8p, _f32p, _f32p, _f32p, _f32p, _u32p]),
            "orc_mesh_light_pdf": (None, [_vp, _f32p, _f32p, _u32, _f32p]),
            "orc_scene_set_env": (None, [_vp, _vp]),
            "orc_scene_num_prims": (_u32, [_vp]),
            "orc_scene_num_materials": (_u32, [_vp]),
            "orc_scene_material": (None, [_vp, _u32, _f32p]),
            "orc_scene_num_lights": (_u32, [_vp]),
            "orc_scene_light": (None, [_vp, _u32, _f32p]),
            "orc_scene_mesh_info": (None, [_vp, _u32, _u32p]),
            "orc_scene_mesh_data": (None, [_vp, _u32, _f32p, _u32p, _vp]),
            "orc_scene_bounds": (None, [_vp, _f32p]),
            "orc_scene_bvh_info": (None, [_vp, _u32, C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32)]),
            "orc_scene_bvh_export": (None, [_vp, _u32, _f32p, _u32p, _u32p, _u32p, _u32p]),
            "orc_scene_intersect": (None, [_vp, _f32p, _u32, _u32, _i32p, _u32p, _u32p, _f32p, _vp]),
            "orc_scene_any_hit": (None, [_vp, _f32p, _u32, _u32, _u8p]),
            "orc_scene_occluded": (None, [_vp, _f32p, _f32p, _f32p, _u32, _u8p]),
            "orc_mesh_traversal_stats": (None, [_vp, _f32p, _u32, _u32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
            "orc_mesh_brute_closest": (None, [_vp, _f32p, _u32, _u32, _u32p, _f32p, _u32p]),
            "orc_camera_rays": (None, [_f32p, _f32p, _u32, _f32p]),
            "orc_auto_frame": (None, [_f32p, C.c_float, C.c_float, _f32p, _f32p]),
            "orc_render": (None, [_vp, _f32p, _u32, _u32, _u32, _u32, _u32, _u32, _u32, _u32, _u32, _f32p, _vp]),
            "orc_render_stats": (None, [_vp, _f32p, _u32, _u32, _u32, _u32, _u32, _u32, _u32, _u32,
                                        C.POINTER(C.c_uint64)]),
            "orc_encode_png": (None, [_f32p, _u32, _u32, _u32, _u8p]),
            "orc_frame": (None, [_f32p, _u32, _f32p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def nthreads() -> int:
    try:
        return max(1, min(len(os.sched_getaffinity(0)), 16))
    except AttributeError:  # pragma: no cover
        return max(1, min(os.cpu_count() or 1, 16))


# ------------------------------------------------------------------ sampling
def pixel_seed(x, y, s) -> int:
    return int(lib().orc_pixel_seed(x, y, s))


def rng_f32(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.float32)
    lib().orc_rng_f32(seed, n, out)
    return out


def rng_u64(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.uint64)
    lib().orc_rng_u64(seed, n, out)
    return out


def rng_u64_from_state(state4, n: int) -> np.ndarray:
    out = np.empty(n, np.uint64)
    lib().orc_rng_u64_from_state(np.ascontiguousarray(state4, np.uint64), n, out)
    return out


def rng_state(seed: int) -> np.ndarray:
    out = np.empty(4, np.uint64)
    lib().orc_rng_state(seed, out)
    return out


def concentric_disk(u):
    u = f32(u).reshape(-1, 2)
    out = np.empty_like(u)
    lib().orc_concentric_disk(u, len(u), out)
    return out


class Distribution1D:
    def __init__(self, weights):
        w = f32(weights)
        self.n = len(w)
        self.h = lib().orc_dist1d_new(w, self.n)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_dist1d_free(self.h)

    @property
    def cdf(self):
        out = np.empty(self.n + 1, np.float32)
        lib().orc_dist1d_cdf(self.h, out)
        return out

    @property
    def integral(self):
        return float(lib().orc_dist1d_integral(self.h))

    def sample_continuous(self, u):
        u = f32(np.atleast_1d(u))
        t = np.empty_like(u)
        pdf = np.empty_like(u)
        b = np.empty(len(u), np.uint32)
        lib().orc_dist1d_sample(self.h, u, len(u), t, pdf, b)
        return t, pdf, b

    def pdf_at(self, t):
        t = f32(np.atleast_1d(t))
        out = np.empty_like(t)
        lib().orc_dist1d_pdf(self.h, t, len(t), out)
        return out


class Distribution2D:
    def __init__(self, weights, width, height):
        w = f32(weights)
        assert w.size == width * height
        self.h = lib().orc_dist2d_new(w, width, height)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_dist2d_free(self.h)

    def sample_continuous(self, u):
        u = f32(u).reshape(-1, 2)
        uv = np.empty_like(u)
        pdf = np.empty(len(u), np.float32)
        lib().orc_dist2d_sample(self.h, u, len(u), uv, pdf)
        return uv, pdf

    def pdf_at(self, uv):
        uv = f32(uv).reshape(-1, 2)
        out = np.empty(len(uv), np.float32)
        lib().orc_dist2d_pdf(self.h, uv, len(uv), out)
        return out


# ------------------------------------------------------------------ shading
LAMBERTIAN, DISNEY = 0, 1


def material(kind, base, metallic=0.0, roughness=0.0, specular=0.0):
    return f32([kind, *base, metallic, roughness, specular])


def bsdf_eval(mat, wo, wi):
    wo = f32(wo).reshape(-1, 3)
    wi = f32(wi).reshape(-1, 3)
    f = np.empty_like(wo)
    pdf = np.empty(len(wo), np.float32)
    lib().orc_bsdf_eval(f32(mat), wo, wi, len(wo), f, pdf)
    return f, pdf


def bsdf_sample(mat, wo, u):
    wo = f32(wo).reshape(-1, 3)
    u = f32(u).reshape(-1, 2)
    wi = np.empty_like(wo)
    f = np.empty_like(wo)
    pdf = np.empty(len(wo), np.float32)
    ok = np.empty(len(wo), np.uint8)
    lib().orc_bsdf_sample(f32(mat), wo, u, len(wo), wi, f, pdf, ok)
    return wi, f, pdf, ok.astype(bool)


def quad_sample(light12, frm, u):
    frm = f32(frm).reshape(-1, 3)
    u = f32(u).reshape(-1, 2)
    n = len(frm)
    pos = np.empty((n, 3), np.float32)
    wi = np.empty((n, 3), np.float32)
    dist = np.empty(n, np.float32)
    pdf = np.empty(n, np.float32)
    ok = np.empty(n, np.uint8)
    lib().orc_quad_sample(f32(light12), frm, u, n, pos, wi, dist, pdf, ok)
    return pos, wi, dist, pdf, ok.astype(bool)


def quad_pdf(light12, frm, wi):
    frm = f32(frm).reshape(-1, 3)
    wi = f32(wi).reshape(-1, 3)
    out = np.empty(len(frm), np.float32)
    lib().orc_quad_pdf(f32(light12), frm, wi, len(frm), out)
    return out


class MeshLight:
    """Emissive triangle mesh behind the Light trait (light.rs:36-42); the
    general emitter the reference's PBRT builder drops (io/pbrt/builder.rs:463-485)."""

    def __init__(self, positions, indices, emission):
        self._pos = f32(positions).reshape(-1)
        self._idx = u32(indices).reshape(-1)
        self.h = lib().orc_mesh_light_new(self._pos, self._pos.size // 3, self._idx, self._idx.size // 3,
                                          f32(emission))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_light_free(self.h)
            self.h = None

    @property
    def area(self):
        return float(lib().orc_mesh_light_area(self.h))

    def sample(self, frm, u):
        frm = f32(frm).reshape(-1, 3)
        u = f32(u).reshape(-1, 2)
        n = len(frm)
        ok = np.empty(n, np.uint8)
        pos = np.empty((n, 3), np.float32)
        wi = np.empty((n, 3), np.float32)
        dist = np.empty(n, np.float32)
        pdf = np.empty(n, np.float32)
        tri = np.empty(n, np.uint32)
        lib().orc_mesh_light_sample(self.h, frm, u, n, ok, pos, wi, dist, pdf, tri)
        return pos, wi, dist, pdf, ok.astype(bool), tri

    def pdf(self, frm, wi):
        frm = f32(frm).reshape(-1, 3)
        wi = f32(wi).reshape(-1, 3)
        out = np.empty(len(frm), np.float32)
        lib().orc_mesh_light_pdf(self.h, frm, wi, len(frm), out)
        return out


class Hdri:
    def __init__(self, width=None, height=None, rgb=None, intensity=1.0, gradient=None):
        if gradient is not None:
            z, h = gradient
            self.h = lib().orc_hdri_gradient(f32(z), f32(h), intensity)
        else:
            rgb = f32(rgb).reshape(-1)
            assert rgb.size == width * height * 3
            self.h = lib().orc_hdri_new(width, height, rgb, intensity)
        self.owned = True

    def __del__(self):
        if getattr(self, "owned", False) and self.h:
            lib().orc_hdri_free(self.h)

    def sample(self, u):
        u = f32(u).reshape(-1, 2)
        n = len(u)
        wi = np.empty((n, 3), np.float32)
        li = np.empty((n, 3), np.float32)
        pdf = np.empty(n, np.float32)
        lib().orc_hdri_sample(self.h, u, n, wi, li, pdf)
        return wi, li, pdf

    def pdf(self, wi):
        wi = f32(wi).reshape(-1, 3)
        out = np.empty(len(wi), np.float32)
        lib().orc_hdri_pdf(self.h, wi, len(wi), out)
        return out

    def le(self, wi):
        wi = f32(wi).reshape(-1, 3)
        out = np.empty_like(wi)
        lib().orc_hdri_le(self.h, wi, len(wi), out)
        return out

    def texels(self):
        """The map's pixels, (height, width, 3), before the intensity."""
        wh = np.empty(2, np.uint32)
        lib().orc_hdri_texels(self.h, wh, None)
        out = np.empty((int(wh[1]), int(wh[0]), 3), np.float32)
        lib().orc_hdri_texels(self.h, wh, out.ctypes.data)
        return out


def dir_to_uv(wi):
    wi = f32(wi).reshape(-1, 3)
    out = np.empty((len(wi), 2), np.float32)
    lib().orc_hdri_uv(wi, len(wi), out)
    return out


def uv_to_dir(uv):
    uv = f32(uv).reshape(-1, 2)
    out = np.empty((len(uv), 3), np.float32)
    lib().orc_hdri_dir(uv, len(uv), out)
    return out


def frame_from_normal(n):
    n = f32(n).reshape(-1, 3)
    out = np.empty((len(n), 9), np.float32)
    lib().orc_frame(n, len(n), out)
    return out.reshape(-1, 3, 3)


# -------------------------------------------------------------------- scenes
class Scene:
    """Scene (scene.rs:24-34) with per-primitive BVHs (bvh.rs)."""

    def __init__(self, handle=None):
        self.h = handle if handle is not None else lib().orc_scene_new()
        self._env = None

    @classmethod
    def cornell(cls, motion_blur=False, bins=12, max_leaf=8):
        return cls(lib().orc_scene_cornell(int(motion_blur), bins, max_leaf))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_free(self.h)

    def add_material(self, mat) -> int:
        return int(lib().orc_scene_add_material(self.h, f32(mat)))

    def add_mesh(self, positions, indices, normals=None, material_id=0, bins=12, max_leaf=8) -> int:
        pos = f32(positions).reshape(-1)
        idx = u32(indices).reshape(-1)
        nrm = None
        if normals is not None:
            nrm_arr = f32(normals).reshape(-1)
            self._keep = nrm_arr
            nrm = nrm_arr.ctypes.data
        return int(lib().orc_scene_add_mesh(self.h, pos, pos.size // 3, idx, idx.size // 3, nrm,
                                            material_id, bins, max_leaf))

    def set_motion(self, prim, off0, off1):
        lib().orc_scene_set_motion(self.h, prim, f32(off0), f32(off1))

    def add_light(self, light12):
        lib().orc_scene_add_light(self.h, f32(light12))

    def add_mesh_light(self, positions, indices, emission):
        pos = f32(positions).reshape(-1)
        idx = u32(indices).reshape(-1)
        lib().orc_scene_add_mesh_light(self.h, pos, pos.size // 3, idx, idx.size // 3, f32(emission))

    def set_env(self, hdri: Hdri):
        hdri.owned = False  # ownership moves to the scene
        lib().orc_scene_set_env(self.h, hdri.h)
        self._env = hdri

    @property
    def num_prims(self):
        return int(lib().orc_scene_num_prims(self.h))

    def materials(self):
        n = lib().orc_scene_num_materials(self.h)
        out = []
        for i in range(n):
            m = np.empty(7, np.float32)
            lib().orc_scene_material(self.h, i, m)
            out.append(m)
        return out

    def lights(self):
        out = []
        for i in range(lib().orc_scene_num_lights(self.h)):
            m = np.empty(12, np.float32)
            lib().orc_scene_light(self.h, i, m)
            out.append(m)
        return out

    def mesh(self, i):
        info = np.empty(4, np.uint32)
        lib().orc_scene_mesh_info(self.h, i, info)
        nv, nt, has_n, mat = (int(x) for x in info)
        pos = np.empty((nv, 3), np.float32)
        idx = np.empty((nt, 3), np.uint32)
        nrm = np.empty((nv, 3), np.float32) if has_n else None
        lib().orc_scene_mesh_data(self.h, i, pos, idx, nrm.ctypes.data if has_n else None)
        return dict(positions=pos, indices=idx, normals=nrm, material_id=mat)

    def bounds(self):
        out = np.empty(6, np.float32)
        lib().orc_scene_bounds(self.h, out)
        return out

    def bvh(self, i):
        nn, npr, dep = _u32(), _u32(), _u32()
        lib().orc_scene_bvh_info(self.h, i, C.byref(nn), C.byref(npr), C.byref(dep))
        nn, npr = nn.value, npr.value
        b = np.empty((nn, 6), np.float32)
        off = np.empty(nn, np.uint32)
        cnt = np.empty(nn, np.uint32)
        ax = np.empty(nn, np.uint32)
        prim = np.empty(npr, np.uint32)
        lib().orc_scene_bvh_export(self.h, i, b, off, cnt, ax, prim)
        return dict(bounds=b, offset=off, count=cnt, axis=ax, prim=prim, depth=dep.value)

    def intersect(self, rays, threads=None, with_pn=False):
        rays = f32(rays).reshape(-1, 9)
        n = len(rays)
        kind = np.empty(n, np.int32)
        pid = np.empty(n, np.uint32)
        tri = np.empty(n, np.uint32)
        tuv = np.empty((n, 3), np.float32)
        pn = np.empty((n, 6), np.float32) if with_pn else None
        lib().orc_scene_intersect(self.h, rays, n, threads or nthreads(), kind, pid, tri, tuv,
                                  pn.ctypes.data if with_pn else None)
        out = dict(kind=kind, id=pid, tri=tri, t=tuv[:, 0], u=tuv[:, 1], v=tuv[:, 2])
        if with_pn:
            out["position"] = pn[:, :3]
            out["normal"] = pn[:, 3:]
        return out

    def any_hit(self, rays, threads=None):
        rays = f32(rays).reshape(-1, 9)
        out = np.empty(len(rays), np.uint8)
        lib().orc_scene_any_hit(self.h, rays, len(rays), threads or nthreads(), out)
        return out

    def occluded(self, frm, to, time):
        frm = f32(frm).reshape(-1, 3)
        to = f32(to).reshape(-1, 3)
        time = f32(np.broadcast_to(time, (len(frm),)))
        out = np.empty(len(frm), np.uint8)
        lib().orc_scene_occluded(self.h, frm, to, time, len(frm), out)
        return out

    def traversal_stats(self, rays, threads=None):
        rays = f32(rays).reshape(-1, 9)
        nn, nt = C.c_uint64(), C.c_uint64()
        lib().orc_mesh_traversal_stats(self.h, rays, len(rays), threads or nthreads(), C.byref(nn), C.byref(nt))
        return nn.value, nt.value

    def brute_closest(self, rays, threads=None):
        rays = f32(rays).reshape(-1, 9)
        n = len(rays)
        tri = np.empty(n, np.uint32)
        t = np.empty(n, np.float32)
        ties = np.empty(n, np.uint32)
        lib().orc_mesh_brute_closest(self.h, rays, n, threads or nthreads(), tri, t, ties)
        return tri, t, ties

    def render(self, cam13, width, height, spp, spp_offset=0, max_depth=8, min_rr=3, rows=None,
               threads=None, per_sample=False):
        y0, y1 = rows if rows is not None else (0, height)
        out = np.empty(((y1 - y0) * width, 10), np.float32)
        ps = np.empty(((y1 - y0) * width, spp, 3), np.float32) if per_sample else None
        lib().orc_render(self.h, f32(cam13), width, height, spp, spp_offset, max_depth, min_rr, y0, y1,
                         threads or nthreads(), out, ps.ctypes.data if per_sample else None)
        return (out, ps) if per_sample else out


    def render_stats(self, cam13, width, height, spp=1, stride=8, max_depth=8, min_rr=3, threads=None):
        """Reference traversal work per scene query in a render of every
        stride-th pixel in x and y: per extension and per shadow query, the
        mean node fetches and triangle tests (bvh.rs:86-143 over every
        primitive, scene.rs:57-140); extension queries that hit a surface;
        samples whose camera ray hit a surface or a light."""
        out = (C.c_uint64 * 8)()
        lib().orc_render_stats(self.h, f32(cam13), width, height, spp, 0, max_depth, min_rr, stride,
                               threads or nthreads(), out)
        ne, nn_e, nt_e, ns, nn_s, nt_s, n_surf, n_first = (int(v) for v in out)
        return {"extension_queries": ne, "n_node_extension": nn_e / max(ne, 1), "n_tri_extension": nt_e / max(ne, 1),
                "shadow_queries": ns, "n_node_shadow": nn_s / max(ns, 1), "n_tri_shadow": nt_s / max(ns, 1),
                "extension_surface_hits": n_surf, "camera_hits": n_first}


def camera(origin, target, up=(0, 1, 0), vfov=45.0, aspect=1.0, aperture=0.0, focus=None):
    origin = np.asarray(origin, np.float32)
    target = np.asarray(target, np.float32)
    if focus is None:
        d = (target - origin).astype(np.float32)
        focus = np.float32(np.sqrt(np.float32(d @ d)))
    return f32([*origin, *target, *up, vfov, aspect, aperture, focus])


def camera_rays(cam13, samples):
    s = f32(samples).reshape(-1, 5)
    out = np.empty((len(s), 9), np.float32)
    lib().orc_camera_rays(f32(cam13), s, len(s), out)
    return out


def auto_frame(bounds6, vfov, aspect):
    o = np.empty(3, np.float32)
    t = np.empty(3, np.float32)
    lib().orc_auto_frame(f32(bounds6), vfov, aspect, o, t)
    return o, t


def encode_png(samples, width, height, channel):
    """channel: 0 beauty, 1 albedo, 2 normal (RGB u8), 3 depth (L u8)."""
    s = f32(samples).reshape(-1, 10)
    out = np.empty((height, width, 1 if channel == 3 else 3), np.uint8)
    lib().orc_encode_png(s, width, height, channel, out)
    return out
