"""pyre_amd -- MI355X (gfx950) path-tracing backend for Pyre.

Python view of the C ABI in include/pyre_hip.h (libpyre_hip.so). The names
mirror the reference's public surface (crates/pyre/src/lib.rs:22-35):
Scene / PathIntegrator / Film / camera / Bsdf probes. Every call runs the HIP
implementation; when the library is missing, import succeeds but the first
call raises.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import weakref
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import PyreError

__all__ = [
    "Context", "Scene", "Preset", "PathIntegrator", "Film", "Camera", "Material", "PyreError",
    "random_triangles", "random_rays", "auto_frame", "MAT_LAMBERTIAN", "MAT_DISNEY", "MISS", "RadianceCache",
    "load_hdr", "write_png", "write_exr",
]

MAT_LAMBERTIAN = L.MAT_LAMBERTIAN
MAT_DISNEY = L.MAT_DISNEY
MISS = L.PYRE_MISS


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Context:
    """pyre_ctx: one GPU, one stream."""

    def __init__(self, device: int = 0):
        self.lib = L.load()
        h = C.c_void_p()
        rc = self.lib.pyre_ctx_create(device, C.byref(h))
        if rc != 0:
            raise PyreError(rc, f"pyre_ctx_create({device}) failed")
        self.h = h
        # scenes, caches and sessions hold device memory of this context and
        # must be destroyed before it
        self._children = weakref.WeakSet()

    def _adopt(self, obj):
        self._children.add(obj)

    def check(self, rc: int, what: str):
        if rc != 0:
            msg = self.lib.pyre_last_error(self.h)
            raise PyreError(rc, f"{what}: {msg.decode() if msg else ''}")

    def device_info(self):
        buf = C.create_string_buffer(128)
        cu = C.c_int()
        self.check(self.lib.pyre_ctx_device_info(self.h, buf, 128, C.byref(cu)), "device_info")
        return buf.value.decode(), cu.value

    def get_tuning(self) -> dict:
        """pyre_ctx_get_tuning: the context's launch settings (0 = default)."""
        t = L.Tuning()
        self.check(self.lib.pyre_ctx_get_tuning(self.h, C.byref(t)), "pyre_ctx_get_tuning")
        return {name: getattr(t, name) for name, _ in L.Tuning._fields_}

    def set_tuning(self, **kw) -> None:
        """pyre_ctx_set_tuning: pool_paths, trace_blocks_per_cu, stack_ring,
        shade_waves, traversal, eager, memory_fraction (unnamed fields keep
        their current values)."""
        cur = self.get_tuning()
        unknown = set(kw) - set(cur)
        if unknown:
            raise ValueError(f"unknown tuning fields {sorted(unknown)}")
        cur.update(kw)
        t = L.Tuning(**cur)
        self.check(self.lib.pyre_ctx_set_tuning(self.h, C.byref(t)), "pyre_ctx_set_tuning")

    @contextlib.contextmanager
    def tuning(self, **kw):
        """Launch settings for the duration of a with-block."""
        old = self.get_tuning()
        self.set_tuning(**kw)
        try:
            yield self
        finally:
            self.set_tuning(**old)

    def close(self):
        if getattr(self, "h", None):
            kids = list(getattr(self, "_children", ()))
            for kind in (Session, RadianceCache, Scene):
                for k in kids:
                    if isinstance(k, kind):
                        k.close()
            self.lib.pyre_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Material:
    """pyre_material: LAMBERTIAN (material.rs:42-77) or DISNEY (:83-226)."""
    kind: int
    base_color: tuple
    metallic: float = 0.0
    roughness: float = 0.0
    specular: float = 0.0

    def to_c(self) -> L.Material:
        return L.Material(self.kind, (C.c_float * 3)(*self.base_color), self.metallic, self.roughness, self.specular)


@dataclass
class Camera:
    """pyre_camera_desc (camera.rs): aperture_radius > 0 -> thin lens."""
    origin: tuple
    target: tuple
    up: tuple = (0.0, 1.0, 0.0)
    vfov_deg: float = 45.0
    aspect: float = 1.0
    aperture_radius: float = 0.0
    focus_distance: float = 0.0

    def to_c(self) -> L.CameraDesc:
        f3 = C.c_float * 3
        return L.CameraDesc(f3(*self.origin), f3(*self.target), f3(*self.up), self.vfov_deg, self.aspect,
                            self.aperture_radius, self.focus_distance)

    @classmethod
    def from_c(cls, c: L.CameraDesc) -> "Camera":
        return cls(tuple(c.origin), tuple(c.target), tuple(c.up), c.vfov_deg, c.aspect, c.aperture_radius,
                   c.focus_distance)


@dataclass
class PathIntegrator:
    """PathIntegrator (integrator.rs:36-48)."""
    max_depth: int = 8
    min_rr_depth: int = 3


def load_hdr(path) -> np.ndarray:
    """load_hdri's decode (io/hdr.rs:24-50): a Radiance .hdr/.pic as an
    (height, width, 3) float32 array, top row first."""
    lib = L.load()
    w, h = L._u32(), L._u32()
    err = C.create_string_buffer(512)
    rc = lib.pyre_load_hdr(os.fsencode(path), C.byref(w), C.byref(h), None, 0, err, 512)
    if rc:
        raise PyreError(rc, f"pyre_load_hdr({path}): {err.value.decode(errors='replace')}")
    out = np.empty((h.value, w.value, 3), np.float32)
    rc = lib.pyre_load_hdr(os.fsencode(path), C.byref(w), C.byref(h), out.ctypes.data, out.size, err, 512)
    if rc:
        raise PyreError(rc, f"pyre_load_hdr({path}): {err.value.decode(errors='replace')}")
    return out


class Preset:
    """Host-built scenes (pyre_gen.h): the CLI's Cornell box, the synthetic
    production scene, and scene files (.pbrt, .gltf/.glb, .hdr environments)."""

    def __init__(self, handle, lib):
        self.h = handle
        self.lib = lib

    @classmethod
    def cornell(cls, motion_blur: bool = False) -> "Preset":
        """cornell_box (main.rs:365-496); motion_blur moves the gold sphere
        by (0, 0.25, 0) over the shutter."""
        lib = L.load()
        h = C.c_void_p()
        rc = lib.pyre_preset_cornell(int(motion_blur), C.byref(h))
        if rc:
            raise PyreError(rc, "pyre_preset_cornell")
        return cls(h, lib)

    @classmethod
    def studio(cls, motion_blur: bool = False, env_intensity: float = 1.0) -> "Preset":
        """studio_scene (main.rs:502-562) with the CLI's default gradient sky."""
        lib = L.load()
        h = C.c_void_p()
        rc = lib.pyre_preset_studio(int(motion_blur), env_intensity, C.byref(h))
        if rc:
            raise PyreError(rc, "pyre_preset_studio")
        return cls(h, lib)

    @classmethod
    def production(cls, seed=3, n_instances=64, level=7, env_w=2048, env_h=1024) -> "Preset":
        lib = L.load()
        h = C.c_void_p()
        rc = lib.pyre_preset_production(seed, n_instances, level, env_w, env_h, C.byref(h))
        if rc:
            raise PyreError(rc, "pyre_preset_production")
        return cls(h, lib)

    @classmethod
    def instances(cls, seed, n_instances, level, env_w=512, env_h=256, lambert_fraction=0.5) -> "Preset":
        """Instanced icospheres like production with a Lambertian share of the
        materials (the batched-jobs scene family)."""
        lib = L.load()
        h = C.c_void_p()
        rc = lib.pyre_preset_instances(seed, n_instances, level, env_w, env_h, lambert_fraction, C.byref(h))
        if rc:
            raise PyreError(rc, "pyre_preset_instances")
        return cls(h, lib)

    @classmethod
    def _load(cls, fn, path) -> "Preset":
        lib = L.load()
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = getattr(lib, fn)(os.fsencode(path), C.byref(h), err, 512)
        if rc:
            raise PyreError(rc, f"{fn}({path}): {err.value.decode(errors='replace')}")
        return cls(h, lib)

    @classmethod
    def from_pbrt(cls, path, mesh_emitters: bool = False) -> "Preset":
        """load_pbrt (io/pbrt/mod.rs:56-76): the reference's PBRT subset; the
        file's LookAt/fov and Film resolution are in .view. mesh_emitters:
        area lights on non-quad shapes become emissive mesh lights instead of
        being dropped (PYRE_PBRT_MESH_EMITTERS)."""
        if not mesh_emitters:
            return cls._load("pyre_preset_load_pbrt", path)
        lib = L.load()
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = lib.pyre_preset_load_pbrt_ex(os.fsencode(path), 1, C.byref(h), err, 512)
        if rc:
            raise PyreError(rc, f"pyre_preset_load_pbrt_ex({path}): {err.value.decode(errors='replace')}")
        return cls(h, lib)

    @classmethod
    def from_gltf(cls, path) -> "Preset":
        """The CLI's glTF scene (main.rs:128-141 over io/gltf.rs:20-92): every
        triangle primitive of the default scene, Lambertian 0.73."""
        return cls._load("pyre_preset_load_gltf", path)

    def set_env_hdr(self, path, intensity: float = 1.0) -> None:
        """The CLI's --env (main.rs:162-166): a Radiance .hdr environment."""
        err = C.create_string_buffer(512)
        rc = self.lib.pyre_preset_set_env_hdr(self.h, os.fsencode(path), intensity, err, 512)
        if rc:
            raise PyreError(rc, f"pyre_preset_set_env_hdr({path}): {err.value.decode(errors='replace')}")

    @property
    def view(self) -> dict:
        """What the scene file says about the render: camera (origin, target,
        up, vfov_deg) and film (width, height), each None when absent."""
        v = L.View()
        self.lib.pyre_preset_view(self.h, C.byref(v))
        cam = dict(origin=tuple(v.origin), target=tuple(v.target), up=tuple(v.up),
                   vfov_deg=v.vfov_deg) if v.has_camera else None
        film = (v.width, v.height) if v.has_film else None
        return dict(camera=cam, film=film)

    def desc(self) -> L.SceneDesc:
        d = L.SceneDesc()
        self.lib.pyre_preset_desc(self.h, C.byref(d))
        return d

    def camera(self, aspect: float, aperture: float = 0.0) -> Camera:
        c = L.CameraDesc()
        self.lib.pyre_preset_camera(self.h, aspect, aperture, C.byref(c))
        return Camera.from_c(c)

    def arrays(self):
        """Copies of the preset's meshes/materials/lights/env as numpy arrays."""
        d = self.desc()
        meshes = []
        for i in range(d.n_meshes):
            m = d.meshes[i]
            pos = np.ctypeslib.as_array(m.positions, (m.n_vertices, 3)).copy()
            idx = np.ctypeslib.as_array(m.indices, (m.n_triangles, 3)).copy()
            nrm = np.ctypeslib.as_array(m.normals, (m.n_vertices, 3)).copy() if m.normals else None
            mot = np.ctypeslib.as_array(m.motion, (2, 3)).copy() if m.motion else None
            meshes.append(dict(positions=pos, indices=idx, normals=nrm, material_id=m.material_id, motion=mot))
        mats = [Material(d.materials[i].kind, tuple(d.materials[i].base_color), d.materials[i].metallic,
                         d.materials[i].roughness, d.materials[i].specular) for i in range(d.n_materials)]
        lights = [np.array([*d.lights[i].p0, *d.lights[i].edge_u, *d.lights[i].edge_v, *d.lights[i].emission],
                           np.float32) for i in range(d.n_lights)]
        env = None
        if d.env:
            e = d.env.contents
            env = dict(width=e.width, height=e.height, intensity=e.intensity,
                       rgb=np.ctypeslib.as_array(e.rgb, (e.height, e.width, 3)).copy())
        mesh_lights = []
        for i in range(d.n_mesh_lights):
            m = d.mesh_lights[i]
            mesh_lights.append(dict(positions=np.ctypeslib.as_array(m.positions, (m.n_vertices, 3)).copy(),
                                    indices=np.ctypeslib.as_array(m.indices, (m.n_triangles, 3)).copy(),
                                    emission=tuple(m.emission)))
        return dict(meshes=meshes, materials=mats, lights=lights, env=env, mesh_lights=mesh_lights)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.pyre_preset_destroy(self.h)
            self.h = None


class Scene:
    """A scene uploaded to the device (pyre_scene) plus its BVH."""

    def __init__(self, ctx: Context, desc: L.SceneDesc, keep=None):
        self.ctx = ctx
        self.lib = ctx.lib
        h = C.c_void_p()
        ctx.check(self.lib.pyre_scene_upload(ctx.h, C.byref(desc), C.byref(h)), "pyre_scene_upload")
        self.h = h
        self._keep = keep
        ctx._adopt(self)

    @classmethod
    def from_preset(cls, ctx: Context, preset: Preset) -> "Scene":
        return cls(ctx, preset.desc(), keep=preset)

    @classmethod
    def from_arrays(cls, ctx: Context, meshes, materials, lights=(), env=None, mesh_lights=()) -> "Scene":
        """meshes: dicts with positions (n,3), indices (m,3), optional normals,
        material_id and motion ((offset_t0, offset_t1), InstanceMotion).
        mesh_lights: dicts with positions, indices and emission (emissive
        triangle meshes, lit after the quad lights)."""
        keep = []
        md = (L.MeshDesc * max(1, len(meshes)))()
        for i, m in enumerate(meshes):
            pos = _f32(m["positions"]).reshape(-1)
            idx = _u32(m["indices"]).reshape(-1)
            nrm = _f32(m["normals"]).reshape(-1) if m.get("normals") is not None else None
            mot = _f32(m["motion"]).reshape(6) if m.get("motion") is not None else None
            keep += [pos, idx, nrm, mot]
            md[i] = L.MeshDesc(pos.ctypes.data_as(L._fp), idx.ctypes.data_as(L._u32p),
                               nrm.ctypes.data_as(L._fp) if nrm is not None else None,
                               pos.size // 3, idx.size // 3, int(m.get("material_id", 0)),
                               mot.ctypes.data_as(L._fp) if mot is not None else None)
        mats = (L.Material * max(1, len(materials)))(*[m.to_c() for m in materials])
        lts = (L.QuadLight * max(1, len(lights)))()
        for i, l in enumerate(lights):
            l = _f32(l)
            lts[i] = L.QuadLight((C.c_float * 3)(*l[0:3]), (C.c_float * 3)(*l[3:6]), (C.c_float * 3)(*l[6:9]),
                                 (C.c_float * 3)(*l[9:12]))
        envp = None
        if env is not None:
            rgb = _f32(env["rgb"]).reshape(-1)
            keep.append(rgb)
            ed = L.EnvDesc(env["width"], env["height"], rgb.ctypes.data_as(L._fp), env.get("intensity", 1.0))
            keep.append(ed)
            envp = C.pointer(ed)
        mls = (L.MeshLight * max(1, len(mesh_lights)))()
        for i, m in enumerate(mesh_lights):
            pos = _f32(m["positions"]).reshape(-1)
            idx = _u32(m["indices"]).reshape(-1)
            keep += [pos, idx]
            mls[i] = L.MeshLight(pos.ctypes.data_as(L._fp), idx.ctypes.data_as(L._u32p), pos.size // 3, idx.size // 3,
                                 (C.c_float * 3)(*[float(x) for x in m["emission"]]))
        d = L.SceneDesc(md, len(meshes), mats, len(materials), lts, len(lights), envp, mls, len(mesh_lights))
        keep += [md, mats, lts, mls]
        return cls(ctx, d, keep=keep)

    def bounds(self):
        lo = (C.c_float * 3)()
        hi = (C.c_float * 3)()
        self.ctx.check(self.lib.pyre_scene_bounds(self.ctx.h, self.h, lo, hi), "pyre_scene_bounds")
        return np.array(lo, np.float32), np.array(hi, np.float32)

    def build_bvh(self, bins=12, max_leaf=8, traversal_cost=1.0, intersect_cost=2.0):
        p = L.BvhParams(bins, max_leaf, traversal_cost, intersect_cost)
        self.ctx.check(self.lib.pyre_bvh_build(self.ctx.h, self.h, C.byref(p)), "pyre_bvh_build")
        return self.bvh_info()

    def bvh_info(self):
        nn, nl, dep = C.c_uint32(), C.c_uint32(), C.c_uint32()
        ms = C.c_double()
        self.ctx.check(self.lib.pyre_bvh_info(self.ctx.h, self.h, C.byref(nn), C.byref(nl), C.byref(dep),
                                              C.byref(ms)), "pyre_bvh_info")
        return dict(n_nodes=nn.value, n_leaves=nl.value, depth=dep.value, build_ms=ms.value)

    def bvh_export(self, n_tris: int):
        info = self.bvh_info()
        n = info["n_nodes"]
        b = np.empty((n, 6), np.float32)
        a = np.empty(n, np.uint32)
        bb = np.empty(n, np.uint32)
        leaf = np.empty(n, np.uint32)
        ax = np.empty(n, np.uint32)
        prim = np.empty(n_tris, np.uint32)
        self.ctx.check(self.lib.pyre_bvh_export(self.ctx.h, self.h, _ptr(b), _ptr(a), _ptr(bb), _ptr(leaf), _ptr(ax),
                                                _ptr(prim)), "pyre_bvh_export")
        return dict(bounds=b, a=a, b=bb, is_leaf=leaf, axis=ax, prim=prim)

    def intersect(self, rays):
        """rays: (n, 9) float32 (o, d, t_min, t_max, time). Returns dict of arrays."""
        rays = _f32(rays).reshape(-1, 9)
        out = np.empty((len(rays), 5), np.uint32)
        self.ctx.check(self.lib.pyre_intersect_closest(self.ctx.h, self.h, _ptr(rays), len(rays), _ptr(out)),
                       "pyre_intersect_closest")
        f = out.view(np.float32)
        return dict(inst=out[:, 0].copy(), tri=out[:, 1].copy(), t=f[:, 2].copy(), u=f[:, 3].copy(),
                    v=f[:, 4].copy())

    def occluded(self, rays):
        rays = _f32(rays).reshape(-1, 9)
        out = np.empty(len(rays), np.uint32)
        self.ctx.check(self.lib.pyre_intersect_any(self.ctx.h, self.h, _ptr(rays), len(rays), _ptr(out)),
                       "pyre_intersect_any")
        return out

    def trace_closest_dev(self, rays_ptr: int, n: int, hits_ptr: int, stream: int = 0):
        self.ctx.check(self.lib.pyre_trace_closest_dev(self.ctx.h, self.h, C.c_void_p(rays_ptr), n,
                                                       C.c_void_p(hits_ptr), C.c_void_p(stream or None)),
                       "pyre_trace_closest_dev")

    def trace_closest_dev_noalloc(self, rays_ptr: int, n: int, hits_ptr: int, stream: int = 0):
        """pyre_trace_closest_dev on the library's fixed scratch (allocates nothing)."""
        self.ctx.check(self.lib.pyre_trace_closest_dev_noalloc(self.ctx.h, self.h, C.c_void_p(rays_ptr), n,
                                                               C.c_void_p(hits_ptr), C.c_void_p(stream or None)),
                       "pyre_trace_closest_dev_noalloc")

    def trace_any_dev(self, rays_ptr: int, n: int, occ_ptr: int, stream: int = 0):
        self.ctx.check(self.lib.pyre_trace_any_dev(self.ctx.h, self.h, C.c_void_p(rays_ptr), n,
                                                   C.c_void_p(occ_ptr), C.c_void_p(stream or None)),
                       "pyre_trace_any_dev")

    # -- shade-stage probes
    def env_sample(self, u):
        u = _f32(u).reshape(-1, 2)
        n = len(u)
        wi, li, pdf = np.empty((n, 3), np.float32), np.empty((n, 3), np.float32), np.empty(n, np.float32)
        self.ctx.check(self.lib.pyre_env_sample_batch(self.ctx.h, self.h, _ptr(u), n, _ptr(wi), _ptr(li), _ptr(pdf)),
                       "pyre_env_sample_batch")
        return wi, li, pdf

    def env_pdf(self, wi):
        wi = _f32(wi).reshape(-1, 3)
        pdf = np.empty(len(wi), np.float32)
        self.ctx.check(self.lib.pyre_env_pdf_batch(self.ctx.h, self.h, _ptr(wi), len(wi), _ptr(pdf)),
                       "pyre_env_pdf_batch")
        return pdf

    def env_le(self, wi):
        wi = _f32(wi).reshape(-1, 3)
        le = np.empty_like(wi)
        self.ctx.check(self.lib.pyre_env_le_batch(self.ctx.h, self.h, _ptr(wi), len(wi), _ptr(le)),
                       "pyre_env_le_batch")
        return le

    def light_sample(self, light, frm, u):
        frm = _f32(frm).reshape(-1, 3)
        u = _f32(u).reshape(-1, 2)
        n = len(frm)
        pos, wi = np.empty((n, 3), np.float32), np.empty((n, 3), np.float32)
        dist, pdf, ok = np.empty(n, np.float32), np.empty(n, np.float32), np.empty(n, np.uint32)
        self.ctx.check(self.lib.pyre_light_sample_batch(self.ctx.h, self.h, light, _ptr(frm), _ptr(u), n, _ptr(pos),
                                                        _ptr(wi), _ptr(dist), _ptr(pdf), _ptr(ok)),
                       "pyre_light_sample_batch")
        return pos, wi, dist, pdf, ok.astype(bool)

    def light_pdf(self, light, frm, wi):
        frm = _f32(frm).reshape(-1, 3)
        wi = _f32(wi).reshape(-1, 3)
        pdf = np.empty(len(frm), np.float32)
        self.ctx.check(self.lib.pyre_light_pdf_batch(self.ctx.h, self.h, light, _ptr(frm), _ptr(wi), len(frm),
                                                     _ptr(pdf)), "pyre_light_pdf_batch")
        return pdf

    def render_stats(self):
        e, sh = C.c_uint64(), C.c_uint64()
        self.ctx.check(self.lib.pyre_render_stats(self.ctx.h, self.h, C.byref(e), C.byref(sh)), "pyre_render_stats")
        return e.value, sh.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.pyre_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bsdf_eval(ctx: Context, mat: Material, wo, wi):
    wo = _f32(wo).reshape(-1, 3)
    wi = _f32(wi).reshape(-1, 3)
    f = np.empty_like(wo)
    pdf = np.empty(len(wo), np.float32)
    m = mat.to_c()
    ctx.check(ctx.lib.pyre_bsdf_eval_batch(ctx.h, C.byref(m), _ptr(wo), _ptr(wi), len(wo), _ptr(f), _ptr(pdf)),
              "pyre_bsdf_eval_batch")
    return f, pdf


def bsdf_sample(ctx: Context, mat: Material, wo, u):
    wo = _f32(wo).reshape(-1, 3)
    u = _f32(u).reshape(-1, 2)
    n = len(wo)
    wi, f = np.empty((n, 3), np.float32), np.empty((n, 3), np.float32)
    pdf, ok = np.empty(n, np.float32), np.empty(n, np.uint32)
    m = mat.to_c()
    ctx.check(ctx.lib.pyre_bsdf_sample_batch(ctx.h, C.byref(m), _ptr(wo), _ptr(u), n, _ptr(wi), _ptr(f), _ptr(pdf),
                                             _ptr(ok)), "pyre_bsdf_sample_batch")
    return wi, f, pdf, ok.astype(bool)


def _pack_shape(what, w, h):
    if what == L.PACK_EXR:
        return (10, h, w), np.float32
    return ((h, w) if what == L.PACK_DEPTH_PNG else (h, w, 3)), np.uint8


def write_png(path, pixels) -> None:
    """An 8-bit PNG of (h, w) grey or (h, w, 3) RGB bytes (pyre_write_png)."""
    px = np.ascontiguousarray(pixels, np.uint8)
    err = C.create_string_buffer(512)
    ch = 1 if px.ndim == 2 else 3
    rc = L.load().pyre_write_png(os.fsencode(path), px.shape[1], px.shape[0], ch, _ptr(px), err, 512)
    if rc:
        raise PyreError(rc, f"pyre_write_png({path}): {err.value.decode(errors='replace')}")


def write_exr(path, planes, albedo=True, normal=True, depth=True) -> None:
    """An OpenEXR of PACK_EXR planes (10, h, w): R, G, B and the chosen AOV
    channels in save_exr's sorted order (pyre_write_exr)."""
    pl = np.ascontiguousarray(planes, np.float32)
    mask = (L.AOV_ALBEDO if albedo else 0) | (L.AOV_NORMAL if normal else 0) | (L.AOV_DEPTH if depth else 0)
    err = C.create_string_buffer(512)
    rc = L.load().pyre_write_exr(os.fsencode(path), pl.shape[2], pl.shape[1], mask, _ptr(pl), err, 512)
    if rc:
        raise PyreError(rc, f"pyre_write_exr({path}): {err.value.decode(errors='replace')}")


class _Sinks:
    """Film::save_png / save_aov_pngs / save_exr (film.rs:117-264) over
    self.pack(what)."""

    def save_png(self, path) -> None:
        write_png(path, self.pack(L.PACK_BEAUTY_PNG))

    def save_aov_pngs(self, base_path, albedo=True, normal=True, depth=True) -> None:
        """<stem>.albedo.png, <stem>.normal.png, <stem>.depth.png beside base_path."""
        d = os.path.dirname(os.fspath(base_path)) or "."
        stem = os.path.splitext(os.path.basename(os.fspath(base_path)))[0] or "render"
        for on, what, name in ((albedo, L.PACK_ALBEDO_PNG, "albedo"), (normal, L.PACK_NORMAL_PNG, "normal"),
                               (depth, L.PACK_DEPTH_PNG, "depth")):
            if on:
                write_png(os.path.join(d, f"{stem}.{name}.png"), self.pack(what))

    def save_exr(self, path, albedo=True, normal=True, depth=True) -> None:
        write_exr(path, self.pack(L.PACK_EXR), albedo, normal, depth)


class Film(_Sinks):
    """Film (film.rs:38-109): per-pixel PixelSample means, row-major, (0,0) top-left."""

    def __init__(self, width: int, height: int):
        self.width = width
        self.height = height
        self.samples = np.zeros((height * width, 10), np.float32)
        self.accum64 = None

    def render(self, scene: Scene, camera: Camera, integrator: PathIntegrator, spp: int, spp_offset: int = 0,
               accum64: bool = False):
        """Replaces main.rs:253-287 (pyre_render)."""
        cam = camera.to_c()
        ip = L.IntegratorParams(integrator.max_depth, integrator.min_rr_depth)
        if accum64 and self.accum64 is None:
            self.accum64 = np.zeros((self.height * self.width, 3), np.float64)
        scene.ctx.check(scene.lib.pyre_render(scene.ctx.h, scene.h, C.byref(cam), C.byref(ip), self.width,
                                              self.height, spp, spp_offset, _ptr(self.samples),
                                              _ptr(self.accum64) if accum64 else None), "pyre_render")
        return self.samples

    def render_split(self, scene: Scene, camera: Camera, integrator: PathIntegrator, spp: int, continuation: int,
                     spp_offset: int = 0, diffuse_roughness: float = 0.25):
        """pyre_render_split: adds to accum64 the radiance sums of `spp`
        samples whose paths continue on stream `continuation` from the vertex
        a cache render would query the radiance cache at."""
        cam = camera.to_c()
        ip = L.IntegratorParams(integrator.max_depth, integrator.min_rr_depth)
        if self.accum64 is None:
            self.accum64 = np.zeros((self.height * self.width, 3), np.float64)
        scene.ctx.check(scene.lib.pyre_render_split(scene.ctx.h, scene.h, C.byref(cam), C.byref(ip), self.width,
                                                    self.height, spp, spp_offset, diffuse_roughness, continuation,
                                                    _ptr(self.accum64)), "pyre_render_split")
        return self.accum64

    def pack(self, what: int) -> np.ndarray:
        """The film packed for a sink on the host (pyre_film_pack)."""
        shape, dt = _pack_shape(what, self.width, self.height)
        out = np.empty(shape, dt)
        px = np.ascontiguousarray(self.samples, np.float32)
        rc = L.load().pyre_film_pack(_ptr(px), self.width, self.height, what, _ptr(out))
        if rc:
            raise PyreError(rc, "pyre_film_pack")
        return out


def random_triangles(seed: int, n: int, sigma: float = 0.01) -> np.ndarray:
    out = np.empty((n, 9), np.float32)
    rc = L.load().pyre_gen_random_triangles(seed, n, sigma, _ptr(out))
    if rc:
        raise PyreError(rc, "pyre_gen_random_triangles")
    return out


def random_rays(seed: int, n: int, n_closest: int, layout: str = "ray36") -> np.ndarray:
    """layout 'ray36' -> (n, 9) host rays; 'ray32' -> (n, 8) device records."""
    if layout == "ray36":
        out = np.empty((n, 9), np.float32)
        rc = L.load().pyre_gen_random_rays(seed, n, n_closest, _ptr(out), None)
    else:
        out = np.empty((n, 8), np.float32)
        rc = L.load().pyre_gen_random_rays(seed, n, n_closest, None, _ptr(out))
    if rc:
        raise PyreError(rc, "pyre_gen_random_rays")
    return out


def auto_frame(bmin, bmax, vfov_deg=45.0, aspect=1.0):
    lo, hi = _f32(bmin), _f32(bmax)
    o, t = (C.c_float * 3)(), (C.c_float * 3)()
    L.load().pyre_auto_frame(lo.ctypes.data_as(L._fp), hi.ctypes.data_as(L._fp), vfov_deg, aspect, o, t)
    return np.array(o, np.float32), np.array(t, np.float32)


class RadianceCache:
    """Neural radiance cache (pyre_nrc_*): 13 query floats -> RGB."""

    def __init__(self, ctx: Context, bounds_min, bounds_max, seed=5, learning_rate=1e-2, beta1=0.9, beta2=0.99,
                 epsilon=1e-15):
        self.ctx = ctx
        self.lib = ctx.lib
        p = L.NrcParams(seed, learning_rate, beta1, beta2, epsilon, (C.c_float * 3)(*bounds_min),
                        (C.c_float * 3)(*bounds_max))
        h = C.c_void_p()
        ctx.check(self.lib.pyre_nrc_create(ctx.h, C.byref(p), C.byref(h)), "pyre_nrc_create")
        self.h = h
        ctx._adopt(self)
        self.n_params = int(self.lib.pyre_nrc_param_count())

    def reserve(self, max_batch: int):
        """Sizes the training buffers for batches of up to max_batch now."""
        self.ctx.check(self.lib.pyre_nrc_reserve(self.ctx.h, self.h, max_batch), "pyre_nrc_reserve")

    def weights(self):
        w = np.empty(self.n_params, np.float32)
        self.ctx.check(self.lib.pyre_nrc_get_weights(self.ctx.h, self.h, _ptr(w)), "pyre_nrc_get_weights")
        return w

    def set_weights(self, w):
        w = _f32(w)
        assert w.size == self.n_params
        self.ctx.check(self.lib.pyre_nrc_set_weights(self.ctx.h, self.h, _ptr(w)), "pyre_nrc_set_weights")

    def infer(self, feats):
        feats = _f32(feats).reshape(-1, 13)
        out = np.empty((len(feats), 3), np.float32)
        self.ctx.check(self.lib.pyre_nrc_infer(self.ctx.h, self.h, _ptr(feats), len(feats), _ptr(out)),
                       "pyre_nrc_infer")
        return out

    def gradients(self, feats, targets):
        feats = _f32(feats).reshape(-1, 13)
        targets = _f32(targets).reshape(-1, 3)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float()
        self.ctx.check(self.lib.pyre_nrc_gradients(self.ctx.h, self.h, _ptr(feats), _ptr(targets), len(feats),
                                                   _ptr(g), C.byref(loss)), "pyre_nrc_gradients")
        return float(loss.value), g

    def train_step(self, feats, targets):
        feats = _f32(feats).reshape(-1, 13)
        targets = _f32(targets).reshape(-1, 3)
        loss = C.c_float()
        self.ctx.check(self.lib.pyre_nrc_train_step(self.ctx.h, self.h, _ptr(feats), _ptr(targets), len(feats),
                                                    C.byref(loss)), "pyre_nrc_train_step")
        return float(loss.value)

    def infer_dev(self, feats_ptr: int, n: int, rgb_ptr: int, stream: int = 0):
        """Inference on device arrays (13 floats per query in, RGB out), queued on stream."""
        self.ctx.check(self.lib.pyre_nrc_infer_dev(self.ctx.h, self.h, C.c_void_p(feats_ptr), n,
                                                   C.c_void_p(rgb_ptr), C.c_void_p(stream)), "pyre_nrc_infer_dev")

    def train_step_dev(self, feats_ptr: int, targets_ptr: int, n: int, loss_ptr: int = 0, stream: int = 0):
        """One Adam step on device arrays, queued on stream; adds the batch loss to *loss_ptr if given."""
        self.ctx.check(self.lib.pyre_nrc_train_step_dev(self.ctx.h, self.h, C.c_void_p(feats_ptr),
                                                        C.c_void_p(targets_ptr), n, C.c_void_p(loss_ptr or None),
                                                        C.c_void_p(stream)), "pyre_nrc_train_step_dev")

    def stats(self):
        """(Adam steps taken, loss of the last training batch)."""
        steps, loss = C.c_uint64(), C.c_float()
        self.ctx.check(self.lib.pyre_nrc_stats(self.ctx.h, self.h, C.byref(steps), C.byref(loss)), "pyre_nrc_stats")
        return int(steps.value), float(loss.value)

    def close(self):
        if getattr(self, "h", None):
            self.lib.pyre_nrc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Session(_Sinks):
    """Progressive render session (pyre_session_*): the main.rs:253-287 loop
    with device-resident sums; samples s = 0, 1, 2, ... per pixel, so N spp
    equal Film.render(spp=N). nrc_train_paths > 0 enables render_nrc."""

    def __init__(self, scene: Scene, camera: Camera, integrator: PathIntegrator, width: int, height: int,
                 nrc_train_paths: int = 0):
        self.scene = scene
        self.ctx = scene.ctx
        self.lib = scene.lib
        self.width, self.height = width, height
        cam = camera.to_c()
        ip = L.IntegratorParams(integrator.max_depth, integrator.min_rr_depth)
        h = C.c_void_p()
        self.ctx.check(self.lib.pyre_session_create(self.ctx.h, scene.h, C.byref(cam), C.byref(ip), width, height,
                                                    nrc_train_paths, C.byref(h)), "pyre_session_create")
        self.h = h
        self.ctx._adopt(self)

    def render(self, spp: int):
        self.ctx.check(self.lib.pyre_session_render(self.ctx.h, self.h, spp), "pyre_session_render")

    def reserve(self, spp_per_pass: int):
        """Sizes the path pool now for passes of up to spp_per_pass samples."""
        self.ctx.check(self.lib.pyre_session_reserve(self.ctx.h, self.h, spp_per_pass), "pyre_session_reserve")

    def render_nrc(self, cache: RadianceCache, frames: int, spp_per_frame: int = 1, train_paths: int = 1 << 14,
                   train_batch: int = 1 << 14, train_steps: int = 4, diffuse_roughness: float = 0.25,
                   warmup_frames: int = 0, plain_frames: int = 0):
        """pyre_session_render_nrc: `warmup_frames` training-only frames, then
        `plain_frames` frames whose image samples run to the end (the plain
        renderer's samples) while their training paths train the cache, then
        frames whose paths stop in the cache. Returns each frame's mean loss."""
        p = L.NrcRenderParams(frames, spp_per_frame, train_paths, train_batch, train_steps, diffuse_roughness,
                              warmup_frames, plain_frames)
        losses = np.zeros(max(frames, 1), np.float32)
        self.ctx.check(self.lib.pyre_session_render_nrc(self.ctx.h, self.h, cache.h, C.byref(p), _ptr(losses)),
                       "pyre_session_render_nrc")
        return losses[:frames]

    def read(self, mean: bool = True, radiance64: bool = False):
        """(PixelSample means (npix, 10) or None, FP64 radiance sums (npix, 3) or None, spp so far)."""
        n = self.width * self.height
        m = np.empty((n, 10), np.float32) if mean else None
        r = np.empty((n, 3), np.float64) if radiance64 else None
        spp = C.c_uint32()
        self.ctx.check(self.lib.pyre_session_read(self.ctx.h, self.h, _ptr(m) if mean else None,
                                                  _ptr(r) if radiance64 else None, C.byref(spp)),
                       "pyre_session_read")
        return m, r, int(spp.value)

    def pack(self, what: int) -> np.ndarray:
        """The session's means packed for a sink on the device (pyre_session_pack)."""
        shape, dt = _pack_shape(what, self.width, self.height)
        out = np.empty(shape, dt)
        self.ctx.check(self.lib.pyre_session_pack(self.ctx.h, self.h, what, _ptr(out)), "pyre_session_pack")
        return out

    def read_viewer(self):
        """The viewer's running-mean radiance (npix, 3) f32 (viewer.rs:165-177)."""
        out = np.empty((self.width * self.height, 3), np.float32)
        self.ctx.check(self.lib.pyre_session_read_viewer(self.ctx.h, self.h, _ptr(out)), "pyre_session_read_viewer")
        return out

    def stats(self):
        """Lifetime totals: extension rays, shadow rays, cache-terminated paths, training samples."""
        out = (C.c_uint64 * 4)()
        self.ctx.check(self.lib.pyre_session_stats(self.ctx.h, self.h, out), "pyre_session_stats")
        return dict(extension_rays=int(out[0]), shadow_rays=int(out[1]), cache_queries=int(out[2]),
                    training_samples=int(out[3]))

    def reset(self):
        self.ctx.check(self.lib.pyre_session_reset(self.ctx.h, self.h), "pyre_session_reset")

    def close(self):
        if getattr(self, "h", None):
            self.lib.pyre_session_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Job:
    """One render of a batch (pyre_job): scene, camera, integrator, image size,
    spp and first sample index."""
    scene: Scene
    camera: Camera
    integrator: PathIntegrator
    width: int
    height: int
    spp: int
    spp_offset: int = 0


def render_batch(ctx: Context, jobs, pool_paths=0, trace_blocks_per_cu=0, stack_ring=0, autotune=False,
                 accum64=False, device_out=None):
    """pyre_render_batch: every job's PixelSample means (npix, 10), its FP64
    radiance sums (npix, 3) if accum64, and the batch statistics.
    device_out: one device address per job (npix * 10 floats on the
    context's device, e.g. a tensor's data_ptr()) that receives the job's
    means in place of a host array; the returned means are then None."""
    n = len(jobs)
    if device_out is not None and len(device_out) != n:
        raise ValueError("device_out needs one address per job")
    outs = None if device_out is not None else [np.zeros((j.width * j.height, 10), np.float32) for j in jobs]
    accs = [np.zeros((j.width * j.height, 3), np.float64) for j in jobs] if accum64 else None
    arr = (L.Job * max(n, 1))()
    for i, j in enumerate(jobs):
        arr[i].scene = j.scene.h
        arr[i].camera = j.camera.to_c()
        arr[i].integrator = L.IntegratorParams(j.integrator.max_depth, j.integrator.min_rr_depth)
        arr[i].width, arr[i].height, arr[i].spp, arr[i].spp_offset = j.width, j.height, j.spp, j.spp_offset
        arr[i].out = C.c_void_p(int(device_out[i])) if device_out is not None else _ptr(outs[i])
        arr[i].accum64 = _ptr(accs[i]) if accum64 else None
    p = L.BatchParams(pool_paths, trace_blocks_per_cu, stack_ring, 1 if autotune else 0)
    st = L.BatchStats()
    ctx.check(ctx.lib.pyre_render_batch(ctx.h, arr, n, C.byref(p), C.byref(st)), "pyre_render_batch")
    stats = dict(extension_rays=int(st.extension_rays), shadow_rays=int(st.shadow_rays), passes=int(st.passes),
                 gpu_ms=float(st.gpu_ms), pool_paths=int(st.pool_paths),
                 trace_blocks_per_cu=int(st.trace_blocks_per_cu), stack_ring=int(st.stack_ring))
    return outs, accs, stats
