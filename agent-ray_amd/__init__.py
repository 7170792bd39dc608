"""agent-ray on MI355X: host-side Python mirror of the reference renderer's
scene / camera / integrator API, driving libagentray.so (hand-written gfx950
HIP kernels behind a C ABI, ``include/agentray.h``).

Names follow the reference (shiinamiyuki/agent-ray):

* ``TriangleMesh``, ``SceneObject``, ``Scene``            (src/scene.rs, src/primitives/mesh.rs)
* ``Lambertian``, ``ConductorBsdf``, ``DielectricBsdf``   (src/surfaces/*)
* ``PointLight``, ``QuadLight``                            (src/lights/*; QuadLight is an extension)
* ``PinholeCamera.from_lookat / from_eye_angle``           (src/cameras.rs)
* ``PathTracer(PathTracerConfig)``                         (src/integrators/path_tracer.rs)
* ``BidirectionalPathTracer(BdptConfig)``, ``MisMode``     (src/integrators/bidirectional_path_tracer.rs)
* ``GradientDomainPathTracer(GdptConfig)``                 (src/integrators/gradient_domain_pt.rs)
* ``Film``                                                 (src/film.rs)

Every compute call goes through the HIP library; there is no CPU fallback.
If the shared library is missing the import raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from enum import IntEnum
from typing import List, Optional, Sequence

import numpy as np

from . import workloads  # noqa: F401  (re-exported)
from .workloads import SceneDesc

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libagentray.so")

AR_OK = 0
AR_HOST_PTRS = 0x1
AR_ASYNC = 0x2
AR_REF_ORDER = 0x4
AR_TUNE_CLOSEST = 0x10
AR_TUNE_ANYHIT = 0x20
AR_COUNT_ANYHIT = 0x40
NO_HIT = 0xFFFFFFFF


class ArError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"agentray error {code}: {msg}")
        self.code = code


# ---------------------------------------------------------------------------
# ctypes mirrors of include/agentray.h
# ---------------------------------------------------------------------------
class ArMaterial(C.Structure):
    _fields_ = [("type", C.c_uint32), ("albedo", C.c_float * 3), ("roughness", C.c_float), ("eta", C.c_float),
                ("albedo_texture", C.c_uint32)]


class ArLight(C.Structure):
    _fields_ = [("type", C.c_uint32), ("position", C.c_float * 3), ("intensity", C.c_float * 3),
                ("edge_u", C.c_float * 3), ("edge_v", C.c_float * 3)]


class ArCamera(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("lower_left", C.c_float * 3), ("horizontal", C.c_float * 3),
                ("vertical", C.c_float * 3), ("forward", C.c_float * 3), ("area", C.c_float)]


class ArRenderConfig(C.Structure):
    _fields_ = [("integrator", C.c_uint32), ("spp", C.c_uint32), ("max_depth", C.c_uint32),
                ("rr_depth", C.c_uint32), ("seed", C.c_uint64), ("schedule", C.c_uint32),
                ("sample_chunk", C.c_uint32), ("mis_mode", C.c_uint32), ("mis_beta", C.c_float),
                ("alpha", C.c_float), ("poisson_iterations", C.c_uint32), ("poisson_sign", C.c_int32),
                ("gdpt_outputs", C.c_uint32)]


class ArRenderStats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("paths", C.c_uint64), ("rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("queue_overflow", C.c_uint32), ("reserved", C.c_uint32)]


class ArSceneInfo(C.Structure):
    _fields_ = [("n_objects", C.c_uint32), ("n_triangles", C.c_uint32), ("n_nodes", C.c_uint32),
                ("n_lights", C.c_uint32), ("max_depth", C.c_uint32), ("device_bytes", C.c_uint64)]


class ArServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_active_jobs", C.c_uint32),
                ("max_concurrent_jobs", C.c_uint32), ("pool_paths", C.c_uint32), ("pool_min_paths", C.c_uint32),
                ("pool_fill_paths", C.c_uint32), ("chunk_paths", C.c_uint32), ("bdpt_chunk_samples", C.c_uint32)]


class ArJobDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("camera", ArCamera), ("config", ArRenderConfig)]


class ArTraceConfig(C.Structure):
    _fields_ = [("workgroup_size", C.c_uint32), ("lds_stack", C.c_uint32), ("blocks_per_cu", C.c_uint32),
                ("refill_min", C.c_uint32), ("search_min", C.c_uint32), ("park_slots", C.c_uint32),
                ("balance", C.c_uint32), ("ms", C.c_float)]

    def as_dict(self) -> dict:
        return {"workgroup_size": self.workgroup_size, "lds_stack": self.lds_stack,
                "blocks_per_cu": self.blocks_per_cu, "refill_min": self.refill_min, "search_min": self.search_min,
                "park_slots": self.park_slots, "balance": self.balance, "ms": self.ms}


def config_key(c: dict) -> tuple:
    """A launch configuration dict as the argument tuple of
    Scene.set_trace_config / set_anyhit_config: (workgroup, LDS stack, blocks
    per CU, refill threshold, search quorum, park slots, balance)."""
    return (c["workgroup_size"], c["lds_stack"], c["blocks_per_cu"], c["refill_min"], c["search_min"],
            c["park_slots"], c["balance"])


class ArTraceCounts(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64),
                ("inst_tests", C.c_uint64), ("spill_writes", C.c_uint64), ("spill_reads", C.c_uint64),
                ("grid", C.c_uint32), ("workgroup_size", C.c_uint32)]


class ArJobStats(C.Structure):
    _fields_ = [("submit_time", C.c_double), ("start_time", C.c_double), ("finish_time", C.c_double),
                ("paths", C.c_uint64), ("rays", C.c_uint64), ("state", C.c_int32)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C agent-ray_amd/csrc)")
    # PyTorch ships its own HIP runtime (same soname, libamdhip64.so.7).  If
    # torch is present it is loaded first so the library binds to that copy:
    # two HIP/HSA runtimes in one process cannot both own the GPU, and device
    # pointers from torch tensors are only valid in the runtime that made them.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    u32, u64, i32, f = C.c_uint32, C.c_uint64, C.c_int32, C.c_float
    sig = {
        "ar_last_error": (C.c_char_p, []),
        "ar_version": (C.c_char_p, []),
        "ar_init": (i32, [C.c_int]),
        "ar_scene_create": (i32, [C.POINTER(P)]),
        "ar_scene_destroy": (i32, [P]),
        "ar_scene_add_material": (i32, [P, C.POINTER(ArMaterial), C.POINTER(u32)]),
        "ar_scene_add_mesh": (i32, [P, P, u32, P, u32, P, P, u32, P, u32, P, C.POINTER(u32)]),
        "ar_scene_add_light": (i32, [P, C.POINTER(ArLight), C.POINTER(u32)]),
        "ar_scene_set_light_distribution": (i32, [P, u32]),
        "ar_scene_commit": (i32, [P]),
        "ar_scene_get_info": (i32, [P, C.POINTER(ArSceneInfo)]),
        "ar_trace": (i32, [P, P, u64, P, P, u32, P]),
        "ar_occluded": (i32, [P, P, u64, P, u32, P]),
        "ar_trace_noalloc": (i32, [P, P, u64, P, P, u32, P]),
        "ar_render_autotune_chunk": (i32, [P, C.POINTER(ArCamera), u32, u32, C.POINTER(ArRenderConfig), P,
                                           C.POINTER(u64)]),
        "ar_trace_count": (i32, [P, P, u64, u32, P, C.POINTER(ArTraceCounts)]),
        "ar_camera_lookat": (i32, [P, P, P, f, f, C.POINTER(ArCamera)]),
        "ar_camera_eye_angle": (i32, [P, f, f, f, f, C.POINTER(ArCamera)]),
        "ar_render": (i32, [P, C.POINTER(ArCamera), u32, u32, C.POINTER(ArRenderConfig), P, P,
                            C.POINTER(ArRenderStats), u32, P]),
        "ar_poisson_solve": (i32, [P, P, P, u32, u32, f, u32, i32, P, C.POINTER(C.c_double), u32, P]),
        "ar_server_create": (i32, [P, C.POINTER(ArServerConfig), C.POINTER(P)]),
        "ar_server_destroy": (i32, [P]),
        "ar_job_submit": (i32, [P, C.POINTER(ArJobDesc), C.POINTER(u64)]),
        "ar_job_poll": (i32, [P, u64, C.POINTER(i32)]),
        "ar_job_wait": (i32, [P, u64, P, C.POINTER(ArJobStats)]),
        "ar_job_cancel": (i32, [P, u64]),
        "ar_server_clock": (C.c_double, [P]),
        "ar_bvh_build_host": (i32, [P, u32, P, P, P, C.POINTER(u32), C.POINTER(u32)]),
        "ar_bvh_build_device": (i32, [P, u32, P, P, P, C.POINTER(u32), C.POINTER(u32)]),
        "ar_scene_set_builder": (i32, [P, u32]),
        "ar_scene_set_trace_config": (i32, [P, C.POINTER(ArTraceConfig)]),
        "ar_scene_get_trace_config": (i32, [P, C.POINTER(ArTraceConfig)]),
        "ar_scene_set_anyhit_config": (i32, [P, C.POINTER(ArTraceConfig)]),
        "ar_scene_get_anyhit_config": (i32, [P, C.POINTER(ArTraceConfig)]),
        "ar_scene_autotune": (i32, [P, P, u64, u32, P, C.POINTER(ArTraceConfig)]),
        "ar_film_to_rgb8": (i32, [P, u32, u32, u32, f, f, f, P, u32, P]),
        "ar_scene_add_texture": (i32, [P, P, u32, u32, C.POINTER(u32)]),
        "ar_scene_set_mesh_texcoords": (i32, [P, u32, P, u32]),
        "ar_image_load_png": (i32, [C.c_char_p, C.POINTER(C.c_void_p), C.POINTER(u32), C.POINTER(u32)]),
        "ar_image_free": (None, [P]),
        "ar_scene_load_obj": (i32, [P, C.c_char_p, P, C.POINTER(u32)]),
        "ar_film_save_exr": (i32, [P, u32, u32, f, C.c_char_p, u32, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = _load()
_initialized = False


def lib():
    return _lib


def _check(code):
    if code != AR_OK:
        raise ArError(code, _lib.ar_last_error().decode(errors="replace"))


def init(device: int = 0):
    global _initialized
    if not _initialized:
        _check(_lib.ar_init(device))
        _initialized = True


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


# ---------------------------------------------------------------------------
# Scene description mirrors of the reference types
# ---------------------------------------------------------------------------
def load_png(path) -> np.ndarray:
    """image::open(path).into_rgb8() for a PNG file: (h, w, 3) uint8, top row
    first (decoded by the library)."""
    buf, w, h = C.c_void_p(), C.c_uint32(), C.c_uint32()
    _check(_lib.ar_image_load_png(os.fsencode(os.fspath(path)), C.byref(buf), C.byref(w), C.byref(h)))
    try:
        out = np.ctypeslib.as_array(C.cast(buf, C.POINTER(C.c_uint8)), shape=(h.value, w.value, 3)).copy()
    finally:
        _lib.ar_image_free(buf)
    return out


class ImageTexture:
    """texture.rs:65-135: an 8-bit sRGB image, linear after loading, sampled
    bilinearly with repeat wrap.  `rgb`: (h, w, 3) uint8, top row first."""

    def __init__(self, rgb: np.ndarray):
        self.rgb = np.ascontiguousarray(rgb, np.uint8).reshape(np.shape(rgb)[0], np.shape(rgb)[1], 3)

    @classmethod
    def load(cls, path) -> "ImageTexture":
        """ImageTexture::load (texture.rs:74-93)."""
        return cls(load_png(path))


@dataclass
class Lambertian:
    """surfaces/lambertian.rs:17-65: a constant albedo, or an ImageTexture
    (Lambertian::with_texture)."""
    albedo: object

    @classmethod
    def with_texture(cls, tex: ImageTexture) -> "Lambertian":
        return cls(tex)

    def to_c(self, texture_id: int = 0):
        a = (0.5, 0.5, 0.5) if isinstance(self.albedo, ImageTexture) else self.albedo
        return ArMaterial(0, (C.c_float * 3)(*a), 1.0, 1.0, texture_id)


@dataclass
class ConductorBsdf:
    """surfaces/microfacet.rs:86-153."""
    f0: Sequence[float]
    roughness: float

    def to_c(self):
        return ArMaterial(1, (C.c_float * 3)(*self.f0), self.roughness, 1.0)


@dataclass
class DielectricBsdf:
    """surfaces/microfacet.rs:163-307."""
    eta: float
    roughness: float

    def to_c(self):
        return ArMaterial(2, (C.c_float * 3)(1.0, 1.0, 1.0), self.roughness, self.eta)


def _mat_from_desc(m, textures=()):
    if m.type == 0:
        if getattr(m, "texture", -1) >= 0:
            return Lambertian.with_texture(textures[m.texture])
        return Lambertian(m.albedo)
    if m.type == 1:
        return ConductorBsdf(m.albedo, m.roughness)
    return DielectricBsdf(m.eta, m.roughness)


@dataclass
class PointLight:
    """lights/point.rs:22-100."""
    position: Sequence[float]
    intensity: Sequence[float]

    def to_c(self):
        z = (C.c_float * 3)(0, 0, 0)
        return ArLight(0, (C.c_float * 3)(*self.position), (C.c_float * 3)(*self.intensity), z, z)


@dataclass
class QuadLight:
    """One-sided quad area light (extension); emits `radiance` along
    normalize(cross(edge_u, edge_v)) and adds its own geometry."""
    corner: Sequence[float]
    edge_u: Sequence[float]
    edge_v: Sequence[float]
    radiance: Sequence[float]

    def to_c(self):
        return ArLight(1, (C.c_float * 3)(*self.corner), (C.c_float * 3)(*self.radiance),
                       (C.c_float * 3)(*self.edge_u), (C.c_float * 3)(*self.edge_v))


@dataclass
class TriangleMesh:
    """primitives/mesh.rs:7-16."""
    positions: np.ndarray
    indices: np.ndarray
    normals: Optional[np.ndarray] = None
    material_slots: Optional[np.ndarray] = None
    tex_coords: Optional[np.ndarray] = None  # (nv, 2) float32


@dataclass
class SceneObject:
    """scene.rs:16-35: mesh, local-to-world transform (column-major 4x4), BSDFs by slot."""
    mesh: TriangleMesh
    transform: Optional[np.ndarray]
    materials: List[object]


LIGHTDIST = {"none": 0, "uniform": 1, "power": 2}


class Scene:
    """scene.rs:113-168.  Builds the BLASes and TLAS on the host with the
    reference's SAH, uploads the flattened scene to HBM and keeps it there."""

    def __init__(self, objects: List[SceneObject], lights: List[object], light_dist: str = "power",
                 obj_files: Sequence = (), builder: str = "split"):
        """`obj_files`: (path, transform or None) pairs loaded with
        load_obj_scene (importer.rs:167-241) after `objects`.  `builder`:
        "sah" (the reference's host SAH), "lbvh" (device linear BVH) or "split"
        (the reference's SAH for reference order, spatial-split SAH for the
        near-first trees) for the meshes' trees."""
        init()
        self._h = C.c_void_p()
        _check(_lib.ar_scene_create(C.byref(self._h)))
        _check(_lib.ar_scene_set_builder(self._h, {"sah": 0, "lbvh": 1, "split": 2}[builder]))
        self._keep = []
        mat_ids, tex_ids = {}, {}
        for obj in objects:
            ids = []
            for m in obj.materials:
                key = id(m)
                if key not in mat_ids:
                    tid = 0
                    if isinstance(getattr(m, "albedo", None), ImageTexture):
                        t = m.albedo
                        if id(t) not in tex_ids:
                            out = C.c_uint32()
                            _check(_lib.ar_scene_add_texture(self._h, _ptr(t.rgb), t.rgb.shape[1], t.rgb.shape[0],
                                                             C.byref(out)))
                            tex_ids[id(t)] = out.value
                        tid = tex_ids[id(t)]
                    out = C.c_uint32()
                    c = m.to_c(tid) if tid else m.to_c()
                    _check(_lib.ar_scene_add_material(self._h, C.byref(c), C.byref(out)))
                    mat_ids[key] = out.value
                ids.append(mat_ids[key])
            self._add_mesh(obj.mesh, ids, obj.transform)
        self.obj_objects = 0
        for path, xf in obj_files:
            n = C.c_uint32()
            xfa = None if xf is None else np.ascontiguousarray(xf, np.float32).reshape(16)
            _check(_lib.ar_scene_load_obj(self._h, os.fsencode(os.fspath(path)), _ptr(xfa), C.byref(n)))
            self.obj_objects += n.value
        for l in lights:
            _check(_lib.ar_scene_add_light(self._h, C.byref(l.to_c()), None))
        _check(_lib.ar_scene_set_light_distribution(self._h, LIGHTDIST[light_dist]))
        _check(_lib.ar_scene_commit(self._h))
        self.info = ArSceneInfo()
        _check(_lib.ar_scene_get_info(self._h, C.byref(self.info)))

    def _add_mesh(self, mesh: TriangleMesh, mat_ids, transform):
        pos = np.ascontiguousarray(mesh.positions, np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(mesh.indices, np.uint32).reshape(-1, 3)
        nrm = None if mesh.normals is None else np.ascontiguousarray(mesh.normals, np.float32).reshape(-1, 3)
        slots = np.zeros(1, np.uint32) if mesh.material_slots is None else \
            np.ascontiguousarray(mesh.material_slots, np.uint32).reshape(-1)
        mats = np.asarray(mat_ids, np.uint32)
        xf = None if transform is None else np.ascontiguousarray(transform, np.float32).reshape(16)
        oid = C.c_uint32()
        _check(_lib.ar_scene_add_mesh(self._h, _ptr(pos), pos.shape[0], _ptr(idx), idx.shape[0], _ptr(nrm),
                                      _ptr(slots), slots.shape[0], _ptr(mats), mats.shape[0], _ptr(xf), C.byref(oid)))
        if mesh.tex_coords is not None:
            uv = np.ascontiguousarray(mesh.tex_coords, np.float32).reshape(-1, 2)
            _check(_lib.ar_scene_set_mesh_texcoords(self._h, oid.value, _ptr(uv), uv.shape[0]))

    @classmethod
    def from_desc(cls, desc: SceneDesc, builder: str = "split") -> "Scene":
        textures = [ImageTexture(t) for t in getattr(desc, "textures", [])]
        mats = [_mat_from_desc(m, textures) for m in desc.materials]
        objs = []
        for m in desc.meshes:
            objs.append(SceneObject(TriangleMesh(m.positions, m.indices, m.normals, m.slots, getattr(m, "uvs", None)),
                                    m.transform, [mats[i] for i in m.materials]))
        lights = []
        for l in desc.lights:
            if l.type == 0:
                lights.append(PointLight(l.position, l.intensity))
            else:
                lights.append(QuadLight(l.position, l.edge_u, l.edge_v, l.intensity))
        return cls(objs, lights, {0: "none", 1: "uniform", 2: "power"}[desc.light_dist], builder=builder)

    @property
    def handle(self):
        return self._h

    def intersect(self, rays, ref_order: bool = False):
        """TLAS::intersect over a batch (bvh.rs:413-447).  `rays`: (n, 8) float32
        (origin, t_min, unit dir, t_max) as numpy (host) or a torch CUDA tensor.
        Returns (t, prim_id, inst_id)."""
        flags = AR_REF_ORDER if ref_order else 0
        if _is_torch_cuda(rays):
            import torch
            n = rays.shape[0]
            hits = torch.empty((n, 2), dtype=torch.float32, device=rays.device)
            inst = torch.empty((n,), dtype=torch.int32, device=rays.device)
            _check(_lib.ar_trace(self._h, C.c_void_p(rays.data_ptr()), n, C.c_void_p(hits.data_ptr()),
                                 C.c_void_p(inst.data_ptr()), flags, None))
            return hits[:, 0], hits[:, 1].view(torch.int32), inst
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = r.shape[0]
        hits = np.empty((n, 2), np.float32)
        inst = np.empty(n, np.uint32)
        _check(_lib.ar_trace(self._h, _ptr(r), n, _ptr(hits), _ptr(inst), flags | AR_HOST_PTRS, None))
        return hits[:, 0].copy(), hits[:, 1].view(np.uint32).copy(), inst

    def trace_counts(self, rays, anyhit: bool = False, ref_order: bool = False) -> dict:
        """Traversal work of the launch ar_trace (or ar_occluded with anyhit)
        makes over these rays (torch CUDA tensor, (n, 8) float32): node,
        triangle and instance records read and stack entries spilled, counted
        by an instrumented build of the same kernel with the same launch."""
        flags = (AR_COUNT_ANYHIT if anyhit else 0) | (AR_REF_ORDER if ref_order else 0)
        c = ArTraceCounts()
        _check(_lib.ar_trace_count(self._h, C.c_void_p(rays.data_ptr()), rays.shape[0], flags, None, C.byref(c)))
        return {k: getattr(c, k) for k, _ in ArTraceCounts._fields_}

    def trace_config(self) -> dict:
        c = ArTraceConfig()
        _check(_lib.ar_scene_get_trace_config(self._h, C.byref(c)))
        return c.as_dict()

    def set_trace_config(self, workgroup_size: int, lds_stack: int, blocks_per_cu: int, refill_min: int = 16,
                         search_min: int = 8, park_slots: int = 3, balance: int = None):
        """Launch configuration of closest-hit and any-hit traversal.  balance
        None keeps the closest-hit launch's current setting."""
        if balance is None:
            balance = self.trace_config()["balance"]
        c = ArTraceConfig(workgroup_size, lds_stack, blocks_per_cu, refill_min, search_min, park_slots, balance, 0.0)
        _check(_lib.ar_scene_set_trace_config(self._h, C.byref(c)))

    def set_anyhit_config(self, workgroup_size: int, lds_stack: int, blocks_per_cu: int, refill_min: int = 16,
                          search_min: int = 8, park_slots: int = 3, balance: int = None):
        """Launch configuration of any-hit traversal only (occlusion, shadow
        rays).  balance None keeps the any-hit launch's current setting."""
        if balance is None:
            balance = self.anyhit_config()["balance"]
        c = ArTraceConfig(workgroup_size, lds_stack, blocks_per_cu, refill_min, search_min, park_slots, balance, 0.0)
        _check(_lib.ar_scene_set_anyhit_config(self._h, C.byref(c)))

    def anyhit_config(self) -> dict:
        c = ArTraceConfig()
        _check(_lib.ar_scene_get_anyhit_config(self._h, C.byref(c)))
        return c.as_dict()

    def autotune(self, rays, stream=None, which: str = "both") -> dict:
        """Times the traversal launch variants on a device ray sample (torch
        CUDA tensor, (n, 8) float32) and keeps the fastest for this scene."""
        c = ArTraceConfig()
        sh = C.c_void_p(stream.cuda_stream) if stream is not None else None
        flags = {"both": 0, "closest": AR_TUNE_CLOSEST, "anyhit": AR_TUNE_ANYHIT}[which]
        _check(_lib.ar_scene_autotune(self._h, C.c_void_p(rays.data_ptr()), rays.shape[0], flags, sh, C.byref(c)))
        d = c.as_dict()
        d["anyhit"] = self.anyhit_config()
        return d

    def occluded(self, rays, ref_order: bool = False):
        """Scene::occluded (scene.rs:166-168) as an any-hit query."""
        flags = AR_REF_ORDER if ref_order else 0
        if _is_torch_cuda(rays):
            import torch
            n = rays.shape[0]
            occ = torch.empty((n,), dtype=torch.int32, device=rays.device)
            _check(_lib.ar_occluded(self._h, C.c_void_p(rays.data_ptr()), n, C.c_void_p(occ.data_ptr()), flags,
                                    None))
            return occ
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        occ = np.empty(r.shape[0], np.uint32)
        _check(_lib.ar_occluded(self._h, _ptr(r), r.shape[0], _ptr(occ), flags | AR_HOST_PTRS, None))
        return occ.astype(bool)

    def __del__(self):
        try:
            if self._h:
                _lib.ar_scene_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass


def _is_torch_cuda(x):
    try:
        import torch
        return isinstance(x, torch.Tensor) and x.is_cuda
    except ImportError:
        return False


# ---------------------------------------------------------------------------
# Camera
# ---------------------------------------------------------------------------
class PinholeCamera:
    """cameras.rs:78-249."""

    def __init__(self, c: ArCamera):
        self.c = c

    @classmethod
    def from_lookat(cls, lookfrom, lookat, vup, vfov, aspect_ratio):
        c = ArCamera()
        a = [np.asarray(v, np.float32) for v in (lookfrom, lookat, vup)]
        _check(_lib.ar_camera_lookat(_ptr(a[0]), _ptr(a[1]), _ptr(a[2]), vfov, aspect_ratio, C.byref(c)))
        return cls(c)

    @classmethod
    def from_eye_angle(cls, eye, yaw, pitch, vfov, aspect_ratio):
        c = ArCamera()
        e = np.asarray(eye, np.float32)
        _check(_lib.ar_camera_eye_angle(_ptr(e), yaw, pitch, vfov, aspect_ratio, C.byref(c)))
        return cls(c)

    @classmethod
    def from_array(cls, cam16):
        c = ArCamera()
        C.memmove(C.addressof(c), np.ascontiguousarray(cam16, np.float32).ctypes.data, 64)
        return cls(c)

    def to_array(self) -> np.ndarray:
        return np.frombuffer(bytes(self.c), dtype=np.float32).copy()


def camera_for(desc: SceneDesc) -> PinholeCamera:
    cd = desc.camera
    return PinholeCamera.from_array(
        workloads.camera_lookat(cd.lookfrom, cd.lookat, cd.vup, cd.vfov, desc.width / desc.height))


# ---------------------------------------------------------------------------
# Film and integrators
# ---------------------------------------------------------------------------
AR_TONEMAP_CLAMP, AR_TONEMAP_REINHARD, AR_TONEMAP_REINHARD_EXTENDED = 0, 1, 2


@dataclass(frozen=True)
class ToneMapper:
    """film.rs:138-151 (ToneMapper): Clamp, Reinhard or ReinhardExtended
    with a white point.  Use ToneMapper.Clamp, ToneMapper.Reinhard and
    ToneMapper.ReinhardExtended(white_point)."""
    kind: int
    white_point: float = 1.0

    @staticmethod
    def ReinhardExtended(white_point: float) -> "ToneMapper":
        return ToneMapper(AR_TONEMAP_REINHARD_EXTENDED, float(white_point))


ToneMapper.Clamp = ToneMapper(AR_TONEMAP_CLAMP)
ToneMapper.Reinhard = ToneMapper(AR_TONEMAP_REINHARD)


def film_to_rgb8(pixels, tone_mapper: ToneMapper = ToneMapper.Reinhard, gamma: float = 2.2, scale: float = 1.0):
    """Film::to_rgb_image (film.rs:283-303) on the GPU for a (h, w, 3)
    float32 film: a numpy array (returns numpy uint8) or a torch CUDA tensor
    (returns a torch uint8 CUDA tensor, nothing leaves the device)."""
    if _is_torch_cuda(pixels):
        import torch
        src = pixels.contiguous()
        h, w = src.shape[0], src.shape[1]
        out = torch.empty((h, w, 3), dtype=torch.uint8, device=src.device)
        st = torch.cuda.current_stream(src.device)
        _check(_lib.ar_film_to_rgb8(C.c_void_p(src.data_ptr()), w, h, tone_mapper.kind, tone_mapper.white_point,
                                    gamma, scale, C.c_void_p(out.data_ptr()), 0, C.c_void_p(st.cuda_stream)))
        return out
    src = np.ascontiguousarray(pixels, np.float32)
    h, w = src.shape[0], src.shape[1]
    out = np.empty((h, w, 3), np.uint8)
    _check(_lib.ar_film_to_rgb8(_ptr(src), w, h, tone_mapper.kind, tone_mapper.white_point, gamma, scale,
                                _ptr(out), AR_HOST_PTRS, None))
    return out


def film_save_exr(pixels, path, scale: float = 1.0):
    """Film::save_exr (film.rs:310-346): half-float OpenEXR (B, G, R; ZIP, 16
    lines per block) of a (h, w, 3) float32 film, numpy or torch CUDA."""
    p = os.fsencode(os.fspath(path))
    if _is_torch_cuda(pixels):
        import torch
        src = pixels.contiguous()
        st = torch.cuda.current_stream(src.device)
        _check(_lib.ar_film_save_exr(C.c_void_p(src.data_ptr()), src.shape[1], src.shape[0], scale, p, 0,
                                     C.c_void_p(st.cuda_stream)))
        return
    src = np.ascontiguousarray(pixels, np.float32)
    _check(_lib.ar_film_save_exr(_ptr(src), src.shape[1], src.shape[0], scale, p, AR_HOST_PTRS, None))


class Film:
    """film.rs:190-274: linear HDR RGB, row-major, top row first."""

    def __init__(self, width, height, pixels: np.ndarray):
        self.width, self.height = width, height
        self.pixels = pixels.reshape(height, width, 3)

    def get_pixel(self, x, y):
        return self.pixels[y, x]

    def to_hdr_vec(self):
        return self.pixels.reshape(-1, 3)

    def to_rgb_image(self, tone_mapper: ToneMapper = ToneMapper.Reinhard, gamma: float = 2.2,
                     scale: float = 1.0) -> np.ndarray:
        """Film::to_rgb_image (film.rs:283-303): (h, w, 3) uint8, computed on
        the GPU."""
        return film_to_rgb8(self.pixels, tone_mapper, gamma, scale)

    def save_exr(self, path, scale: float = 1.0):
        """Film::save_exr (film.rs:310-346)."""
        film_save_exr(self.pixels, path, scale)


class Schedule(IntEnum):
    WAVEFRONT = 0
    MEGAKERNEL = 1


@dataclass
class PathTracerConfig:
    """path_tracer.rs:18-31."""
    spp: int = 64
    max_depth: int = 8
    rr_depth: int = 3


class MisMode(IntEnum):
    """bidirectional_path_tracer.rs:91-99 (Power, Uniform), plus Textbook: the
    power heuristic with every reverse pdf taken once, the camera cosine and
    the light pick's pmf included (no reference counterpart), whose weights
    sum to one, so the image equals the path tracer's in expectation."""
    Power = 0
    Uniform = 1
    Textbook = 2


@dataclass
class BdptConfig:
    """bidirectional_path_tracer.rs:102-133 (debug_strategy_images is not offered)."""
    spp: int = 64
    max_depth: int = 8
    rr_depth: int = 3
    mis_mode: MisMode = MisMode.Power
    mis_beta: float = 2.0


@dataclass
class GdptConfig:
    """gradient_domain_pt.rs:47-76; the Poisson solve is conjugate gradient
    (poisson_iterations CG steps) instead of SOR, `sor_omega` is unused."""
    spp: int = 64
    max_depth: int = 8
    rr_depth: int = 3
    alpha: float = 0.2
    poisson_iterations: int = 50
    sor_omega: float = 1.6
    poisson_sign: int = 1


# Launch candidates the on-the-job refinements try around the ray-sample
# autotuner's pick: (workgroup, LDS stack, blocks per CU, refill threshold,
# search quorum, park slots); each is tried with the stage's current balance
# setting.  The winner is then tried with the other park slot count (3 <-> 4)
# and with balance switched.
TUNE_CANDIDATES = [(256, 12, 8, 8, 16, 3), (256, 12, 8, 8, 8, 4), (256, 12, 8, 8, 32, 4), (256, 12, 8, 8, 32, 3),
                   (256, 8, 8, 8, 32, 4), (128, 16, 8, 8, 32, 3),
                   (256, 8, 8, 8, 16, 3), (256, 8, 8, 8, 8, 3), (256, 8, 8, 8, 16, 4), (256, 8, 8, 8, 8, 4),
                   (256, 8, 8, 24, 16, 3), (128, 8, 16, 8, 16, 3), (128, 16, 8, 8, 16, 3), (128, 16, 8, 8, 8, 3),
                   (64, 16, 16, 8, 16, 3)]


def other_park_slots(c: tuple) -> tuple:
    """The launch configuration tuple `c` with the other park slot count."""
    return tuple(c[:5]) + (7 - c[5],) + tuple(c[6:])


def other_balance(c: tuple) -> tuple:
    """The launch configuration tuple `c` (7 entries) with balance switched."""
    return tuple(c[:6]) + (1 - c[6],)


def with_balance(c: tuple, balance: int) -> tuple:
    """A 6- or 7-entry launch configuration tuple with the given balance."""
    return tuple(c[:6]) + (balance,)


class _Integrator:
    integrator_id = 0

    def __init__(self, config, seed: int = 0, schedule: Schedule = Schedule.WAVEFRONT, sample_chunk: int = 0):
        self.config = config
        self.seed = seed
        self.schedule = schedule
        self.sample_chunk = sample_chunk
        self.last_stats = None
        self.aux = None

    def _cfg(self):
        c = ArRenderConfig()
        c.integrator = self.integrator_id
        c.spp = self.config.spp
        c.max_depth = self.config.max_depth
        c.rr_depth = self.config.rr_depth
        c.seed = self.seed
        c.schedule = int(self.schedule)
        c.sample_chunk = self.sample_chunk
        c.mis_mode = int(getattr(self.config, "mis_mode", 0))
        c.mis_beta = float(getattr(self.config, "mis_beta", 2.0))
        c.alpha = float(getattr(self.config, "alpha", 0.2))
        c.poisson_iterations = int(getattr(self.config, "poisson_iterations", 50))
        c.poisson_sign = int(getattr(self.config, "poisson_sign", 1))
        c.gdpt_outputs = 1 if self.integrator_id == 2 else 0
        return c

    def tune_chunk(self, scene: Scene, camera: PinholeCamera, width: int, height: int) -> int:
        """Queue chunk of this integrator's renders on `scene`: the paths
        (BDPT: samples) one wavefront pass holds, timed per candidate on one
        chunk of this job (ar_render_autotune_chunk) and kept on the scene."""
        c = self._cfg()
        c.gdpt_outputs = 0
        best = C.c_uint64()
        _check(_lib.ar_render_autotune_chunk(scene.handle, C.byref(camera.c), width, height, C.byref(c), None,
                                             C.byref(best)))
        return best.value

    def tune_launch(self, scene: Scene, camera: PinholeCamera, width: int, height: int, spp: int = 16,
                    candidates=None) -> dict:
        """Refines the scene's launch configurations on the job itself: times
        a render of `spp` samples per pixel (one chunk of the real workload,
        wavefront schedule) for each candidate (workgroup, LDS stack, blocks
        per CU, refill threshold, search quorum, park slots, with the stage's
        current balance setting), the scene's current one first; first for the
        closest-hit launch (extension stage), then, with the winner, for the
        any-hit launch (shadow stage); each stage's winner is then timed with
        the other park slot count and with balance switched.  Keeps the
        fastest of each and returns them with the timings.  Results do not
        depend on the launch configuration, so this only changes speed."""
        if candidates is None:
            candidates = TUNE_CANDIDATES
        best_c, best_a = config_key(scene.trace_config()), config_key(scene.anyhit_config())
        saved = (self.config.spp, self.schedule)
        self.config.spp, self.schedule = spp, Schedule.WAVEFRONT
        tried = {}

        def timed(c, a):
            scene.set_trace_config(*c)
            scene.set_anyhit_config(*a)
            self.render(scene, camera, width, height)
            return self.last_stats.seconds

        try:
            for stage in ("closest", "anyhit"):
                base = best_c if stage == "closest" else best_a
                cands = [base] + [with_balance(c, base[6]) for c in candidates if with_balance(c, base[6]) != base]
                times = [timed(c, best_a) if stage == "closest" else timed(best_c, c) for c in cands]
                win = cands[int(np.argmin(times))]
                for other in (other_park_slots, other_balance):
                    flip = other(win)
                    if flip not in cands:
                        cands.append(flip)
                        times.append(timed(flip, best_a) if stage == "closest" else timed(best_c, flip))
                        win = cands[int(np.argmin(times))]
                tried[stage] = {"x".join(map(str, c)): round(t * 1e3, 2) for c, t in zip(cands, times)}
                if stage == "closest":
                    best_c = win
                else:
                    best_a = win
        finally:
            self.config.spp, self.schedule = saved
            scene.set_trace_config(*best_c)
            scene.set_anyhit_config(*best_a)
        return {"closest": list(best_c), "anyhit": list(best_a), "tried": tried}

    def render(self, scene: Scene, camera: PinholeCamera, width: int, height: int) -> Film:
        """Integrator::render (integrators/mod.rs:19-27)."""
        out = np.zeros((height, width, 3), np.float32)
        aux = np.zeros((3, height, width, 3), np.float32) if self.integrator_id == 2 else None
        st = ArRenderStats()
        cfg = self._cfg()
        _check(_lib.ar_render(scene.handle, C.byref(camera.c), width, height, C.byref(cfg), _ptr(out),
                              _ptr(aux) if aux is not None else None, C.byref(st), AR_HOST_PTRS, None))
        self.last_stats = st
        self.aux = aux
        return Film(width, height, out)


class PathTracer(_Integrator):
    integrator_id = 0


class BidirectionalPathTracer(_Integrator):
    integrator_id = 1


class GradientDomainPathTracer(_Integrator):
    integrator_id = 2


def screened_poisson_reconstruct(primal, dx, dy, alpha: float, iterations: int = 50, sign: int = 1):
    """gradient_domain_pt.rs:449-630 as a GPU conjugate-gradient solve.
    Inputs (H, W, 3) float32; returns (image, final residual norm)."""
    p = np.ascontiguousarray(primal, np.float32)
    gx = np.ascontiguousarray(dx, np.float32)
    gy = np.ascontiguousarray(dy, np.float32)
    h, w = p.shape[0], p.shape[1]
    out = np.empty_like(p)
    res = C.c_double()
    init()
    _check(_lib.ar_poisson_solve(_ptr(p), _ptr(gx), _ptr(gy), w, h, alpha, iterations, sign, _ptr(out),
                                 C.byref(res), AR_HOST_PTRS, None))
    return out, res.value


def bvh_build_host(boxes: np.ndarray, device: bool = False):
    """Host SAH build (the library's builder) for topology tests, or with
    device=True the device LBVH builder: returns (boxes (m,6), meta (m,3)
    [left_or_start, right_or_end, count], prims, depth)."""
    b = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
    n = b.shape[0]
    cap = max(1, 2 * n)
    ob = np.zeros((cap, 6), np.float32)
    om = np.zeros((cap, 3), np.uint32)
    op = np.zeros(max(1, n), np.uint32)
    nn = C.c_uint32()
    dep = C.c_uint32()
    fn = _lib.ar_bvh_build_device if device else _lib.ar_bvh_build_host
    _check(fn(_ptr(b), n, _ptr(ob), _ptr(om), _ptr(op), C.byref(nn), C.byref(dep)))
    m = nn.value
    return ob[:m], om[:m], op[:n], dep.value


# ---------------------------------------------------------------------------
# Serving: continuous-batching render server over a resident scene
# ---------------------------------------------------------------------------
JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_CANCELLED, JOB_FAILED = 0, 1, 2, 3, 4


class RenderServer:
    """Accepts render jobs of differing resolution, spp and integrator against
    one resident scene; jobs are admitted into slots by HBM budget and run
    concurrently (see include/agentray.h, ar_server_*)."""

    POOL_OFF = 0xFFFFFFFF

    def __init__(self, scene: Scene, hbm_budget_bytes: int = 200 << 30, max_active_jobs: int = 8,
                 max_concurrent_jobs: int = 0, pool_paths: int = 0, pool_min_paths: int = 0,
                 pool_fill_paths: int = 0, chunk_paths: int = 0, bdpt_chunk_samples: int = 0):
        """pool_paths: paths in flight in the shared path-tracer queues (0 =
        the library default, RenderServer.POOL_OFF = no pool); PT / GDPT jobs
        of at least pool_min_paths paths (0 = default) share them; while one
        drains its last paths the others keep at most pool_fill_paths alive
        (0 = default).  Jobs
        outside the pool run in chunks of chunk_paths paths (BDPT:
        bdpt_chunk_samples samples); 0 = default."""
        self.scene = scene
        cfg = ArServerConfig(hbm_budget_bytes, max_active_jobs, max_concurrent_jobs, pool_paths, pool_min_paths,
                             pool_fill_paths, chunk_paths, bdpt_chunk_samples)
        self._h = C.c_void_p()
        _check(_lib.ar_server_create(scene.handle, C.byref(cfg), C.byref(self._h)))
        self._shapes = {}

    def submit(self, integrator: "_Integrator", camera: PinholeCamera, width: int, height: int) -> int:
        d = ArJobDesc()
        d.width, d.height = width, height
        d.camera = camera.c
        d.config = integrator._cfg()
        d.config.gdpt_outputs = 0
        jid = C.c_uint64()
        _check(_lib.ar_job_submit(self._h, C.byref(d), C.byref(jid)))
        self._shapes[jid.value] = (height, width)
        return jid.value

    def poll(self, job_id: int) -> int:
        st = C.c_int32()
        _check(_lib.ar_job_poll(self._h, job_id, C.byref(st)))
        return st.value

    def wait(self, job_id: int):
        """Blocks until the job ends; returns (Film, stats dict)."""
        h, w = self._shapes.pop(job_id)
        out = np.zeros((h, w, 3), np.float32)
        st = ArJobStats()
        _check(_lib.ar_job_wait(self._h, job_id, _ptr(out), C.byref(st)))
        stats = {"submit": st.submit_time, "start": st.start_time, "finish": st.finish_time,
                 "latency": st.finish_time - st.submit_time, "paths": st.paths, "rays": st.rays}
        return Film(w, h, out), stats

    def cancel(self, job_id: int):
        _check(_lib.ar_job_cancel(self._h, job_id))

    def clock(self) -> float:
        return _lib.ar_server_clock(self._h)

    def close(self):
        if self._h:
            _lib.ar_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
