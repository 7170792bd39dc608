"""Python binding of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  The oracle is a CPU restatement of the reference
renderer's hot path; it is imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's CPU-baseline leg.  The product library never
loads it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        P, u32, u64, f, d = C.c_void_p, C.c_uint32, C.c_uint64, C.c_float, C.c_double
        sig = {
            "orc_scene_new": (P, []),
            "orc_scene_free": (None, [P]),
            "orc_scene_add_material": (C.c_int, [P, u32, P]),
            "orc_scene_add_mesh": (C.c_int, [P, P, u32, P, u32, P, P, u32, P, u32, P]),
            "orc_scene_add_light": (C.c_int, [P, u32, P]),
            "orc_scene_commit": (C.c_int, [P, C.c_int]),
            "orc_bvh_dump": (u32, [P, C.c_int, P, P, P]),
            "orc_trace": (d, [P, P, u64, P, P, P, P, C.c_int]),
            "orc_occluded": (d, [P, P, u64, P, P, C.c_int]),
            "orc_ray_new": (None, [P, P, f, f, P]),
            "orc_triangle_hit": (C.c_int, [P, P, P, P, P]),
            "orc_aabb_hit": (C.c_int, [P, P, P]),
            "orc_camera_lookat": (None, [P, P, P, f, f, P]),
            "orc_camera_ray": (None, [P, f, f, P]),
            "orc_camera_sample_we": (C.c_int, [P, P, u32, u32, P]),
            "orc_philox": (None, [P, P, P]),
            "orc_counter_sample": (f, [u64, u32, u32, u32, u32]),
            "orc_ref_sampler": (None, [u32, u32, u32, P]),
            "orc_bsdf_eval": (None, [u32, P, P, P, P]),
            "orc_bsdf_pdf": (f, [u32, P, P, P]),
            "orc_bsdf_sample": (C.c_int, [u32, P, P, f, f, f, P]),
            "orc_render_pt": (d, [P, P, u32, u32, u32, u32, u32, C.c_int, u64, C.c_int, u32, u32, P, P, P]),
            "orc_render_gdpt": (d, [P, P, u32, u32, u32, u32, u32, C.c_int, u64, C.c_int, P, P, P]),
            "orc_poisson_sor": (d, [P, P, P, u32, u32, f, u32, f, C.c_int, P]),
            "orc_render_bdpt": (d, [P, P, u32, u32, u32, u32, u32, C.c_int, f, C.c_int, u64, C.c_int, P, P]),
            "orc_render_bdpt_ex": (d, [P, P, u32, u32, u32, u32, u32, C.c_int, f, C.c_int, u64, C.c_int,
                                       C.c_int, C.c_int, C.c_int, u32, u32, u32, u32, P, P]),
            "orc_render_gdpt_ex": (d, [P, P, u32, u32, u32, u32, u32, C.c_int, u64, C.c_int, u32, u32, u32, u32,
                                       P, P, P]),
            "orc_light_dist": (u32, [P, P, P]),
            "orc_light_sample_index": (u32, [P, f, P]),
            "orc_render_pt_ex": (d, [P, P, u32, u32, u32, u32, u32, C.c_int, u64, C.c_int, u32, u32, u32, u32,
                                     C.c_int, u32, P, P, P]),
            "orc_film_rgb8": (None, [P, u64, u32, f, f, f, P]),
            "orc_scene_add_texture": (C.c_int, [P, P, u32, u32]),
            "orc_scene_set_material_texture": (None, [P, u32, C.c_int]),
            "orc_scene_set_mesh_uv": (None, [P, u32, P, u32]),
            "orc_texture_sample": (None, [P, u32, P, u32, P]),
            "orc_srgb_to_linear": (f, [f]),
            "orc_f32_to_f16": (None, [P, u64, P]),
            "orc_exr_blocks": (None, [P, u32, u32, f, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def f32(*v):
    return np.asarray(v, np.float32)


class OracleScene:
    """Scene built by the oracle from a workloads.SceneDesc."""

    def __init__(self, desc):
        L = lib()
        self.h = C.c_void_p(L.orc_scene_new())
        self._ntris = [int(m.indices.shape[0]) for m in desc.meshes] + \
            [2 for l in desc.lights if l.type == 1]
        for t in getattr(desc, "textures", []):
            t = np.ascontiguousarray(t, np.uint8)
            L.orc_scene_add_texture(self.h, _p(t), t.shape[1], t.shape[0])
        for k, m in enumerate(desc.materials):
            p = np.array(list(m.albedo) + [m.roughness, m.eta, 0, 0, 0], np.float32)
            L.orc_scene_add_material(self.h, m.type, _p(p))
            if getattr(m, "texture", -1) >= 0:
                L.orc_scene_set_material_texture(self.h, k, m.texture)
        for k, m in enumerate(desc.meshes):
            pos = np.ascontiguousarray(m.positions, np.float32)
            idx = np.ascontiguousarray(m.indices, np.uint32)
            nrm = None if m.normals is None else np.ascontiguousarray(m.normals, np.float32)
            slots = np.ascontiguousarray(m.slots, np.uint32)
            mats = np.asarray(m.materials, np.uint32)
            xf = None if m.transform is None else np.ascontiguousarray(m.transform, np.float32)
            L.orc_scene_add_mesh(self.h, _p(pos), pos.shape[0], _p(idx), idx.shape[0], _p(nrm), _p(slots),
                                 slots.shape[0], _p(mats), mats.shape[0], _p(xf))
            if getattr(m, "uvs", None) is not None:
                uv = np.ascontiguousarray(m.uvs, np.float32)
                L.orc_scene_set_mesh_uv(self.h, k, _p(uv), uv.shape[0])
        for l in desc.lights:
            p = np.array(list(l.position) + list(l.intensity) + list(l.edge_u) + list(l.edge_v), np.float32)
            L.orc_scene_add_light(self.h, l.type, _p(p))
        L.orc_scene_commit(self.h, desc.light_dist)

    def light_dist(self):
        """(pmf, cdf) of the scene's light selection (lights.rs:115-235)."""
        n = lib().orc_light_dist(self.h, None, None)
        pmf, cdf = np.zeros(n, np.float32), np.zeros(n, np.float32)
        lib().orc_light_dist(self.h, _p(pmf), _p(cdf))
        return pmf, cdf

    def light_sample_index(self, u):
        """LightDistribution::sample_index(u) -> (index, pmf)."""
        p = np.zeros(1, np.float32)
        i = lib().orc_light_sample_index(self.h, float(u), _p(p))
        return int(i), float(p[0])

    def texture_sample(self, tex, uv):
        """ImageTexture::sample (texture.rs:97-122) at (n, 2) uv -> (n, 3)."""
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        out = np.zeros((uv.shape[0], 3), np.float32)
        lib().orc_texture_sample(self.h, tex, _p(uv), uv.shape[0], _p(out))
        return out

    def __del__(self):
        try:
            lib().orc_scene_free(self.h)
        except Exception:
            pass

    def bvh(self, obj: int):
        L = lib()
        n = L.orc_bvh_dump(self.h, obj, None, None, None)
        boxes = np.zeros((n, 6), np.float32)
        meta = np.zeros((n, 3), np.uint32)
        # first pass with a one-element prim buffer is unsafe; size it from an
        # upper bound: every leaf holds at least one primitive and the tree has
        # (n + 1) / 2 leaves, but degenerate leaves may hold many, so use the
        # object's triangle count (or instance count for the TLAS).
        cap = self._prim_count(obj)
        prims = np.zeros(max(1, cap), np.uint32)
        L.orc_bvh_dump(self.h, obj, _p(boxes), _p(meta), _p(prims))
        leaves = meta[:, 2] > 0
        total = int(meta[leaves, 2].sum())
        return boxes, meta, prims[:total]

    def _prim_count(self, obj):
        if obj < 0:
            return len(self._ntris)
        return self._ntris[obj]

    def trace(self, rays, nthreads=0):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = r.shape[0]
        t = np.empty(n, np.float32)
        prim = np.empty(n, np.uint32)
        inst = np.empty(n, np.uint32)
        cnt = np.zeros(2, np.uint64)
        secs = lib().orc_trace(self.h, _p(r), n, _p(t), _p(prim), _p(inst), _p(cnt), nthreads)
        return t, prim, inst, cnt, secs

    def occluded(self, rays, nthreads=0):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = r.shape[0]
        occ = np.empty(n, np.uint32)
        cnt = np.zeros(2, np.uint64)
        secs = lib().orc_occluded(self.h, _p(r), n, _p(occ), _p(cnt), nthreads)
        return occ.astype(bool), cnt, secs

    def render_pt(self, cam16, w, h, spp, max_depth=8, rr_depth=3, rng_mode=1, seed=0, nthreads=0,
                  rows=None, want_m2=False, pt_mode=0, only_len=0, cols=None):
        """PathTracer::render.  pt_mode / only_len are test-only estimator
        switches: 0 MIS, 1 light sampling only, 2 BSDF hits only; only_len > 0
        keeps paths of exactly that many vertices (camera and light included).
        rows / cols = (begin, end) restrict the render to a window."""
        mean = np.zeros((h, w, 3), np.float32)
        m2 = np.zeros((h, w, 3), np.float32) if want_m2 else None
        cnt = np.zeros(2, np.uint64)
        r0, r1 = (0, h) if rows is None else rows
        c0, c1 = (0, w) if cols is None else cols
        cam = np.ascontiguousarray(cam16, np.float32)
        secs = lib().orc_render_pt_ex(self.h, _p(cam), w, h, spp, max_depth, rr_depth, rng_mode, seed, nthreads, r0,
                                      r1, c0, c1, pt_mode, only_len, _p(mean), _p(m2), _p(cnt))
        return mean, m2, cnt, secs


def render_gdpt(scene, cam16, w, h, spp, max_depth=8, rr_depth=3, rng_mode=1, seed=0, nthreads=0, window=None):
    """GDPT pass 1: (primal, dx, dy) films, each (h, w, 3), and seconds.
    window = (x0, x1, y0, y1) restricts the render to those pixels."""
    films = [np.zeros((h, w, 3), np.float32) for _ in range(3)]
    cam = np.ascontiguousarray(cam16, np.float32)
    x0, x1, y0, y1 = (0, 0, 0, 0) if window is None else window
    secs = lib().orc_render_gdpt_ex(scene.h, _p(cam), w, h, spp, max_depth, rr_depth, rng_mode, seed, nthreads,
                                    x0, x1, y0, y1, _p(films[0]), _p(films[1]), _p(films[2]))
    return films[0], films[1], films[2], secs


def render_bdpt(scene, cam16, w, h, spp, max_depth=8, rr_depth=3, mis_mode=0, mis_beta=2.0, rng_mode=1, seed=0,
                nthreads=0, strategy=None, unweighted=False, window=None):
    """BidirectionalPathTracer::render: (film, splat film, seconds).
    mis_mode 0 power heuristic (reference walk), 1 uniform, 2 textbook walk
    (test only).  strategy=(s, t) evaluates that strategy alone and
    unweighted=True sets every MIS weight to 1 (test only).  window = (x0, x1,
    y0, y1) traces the camera paths of those pixels only."""
    out = np.zeros((h, w, 3), np.float32)
    spl = np.zeros((h, w, 3), np.float32)
    cam = np.ascontiguousarray(cam16, np.float32)
    only_s, only_t = (-1, -1) if strategy is None else strategy
    x0, x1, y0, y1 = (0, 0, 0, 0) if window is None else window
    secs = lib().orc_render_bdpt_ex(scene.h, _p(cam), w, h, spp, max_depth, rr_depth, mis_mode, mis_beta, rng_mode,
                                    seed, nthreads, only_s, only_t, int(unweighted), x0, x1, y0, y1, _p(out), _p(spl))
    return out, spl, secs


def poisson_sor(primal, dx, dy, alpha, iterations=200, omega=1.6, sign=1):
    """screened_poisson_reconstruct (reference red-black SOR); returns (img, seconds)."""
    p = np.ascontiguousarray(primal, np.float32)
    gx = np.ascontiguousarray(dx, np.float32)
    gy = np.ascontiguousarray(dy, np.float32)
    h, w = p.shape[0], p.shape[1]
    out = np.zeros_like(p)
    secs = lib().orc_poisson_sor(_p(p), _p(gx), _p(gy), w, h, alpha, iterations, omega, sign, _p(out))
    return out, secs


def film_rgb8(film, tonemapper=1, white_point=1.0, gamma=2.2, scale=1.0):
    """Film::to_rgb_image (film.rs:283-303): (h, w, 3) float32 -> (h, w, 3) uint8."""
    f = np.ascontiguousarray(film, np.float32)
    out = np.zeros(f.shape, np.uint8)
    lib().orc_film_rgb8(_p(f), f.size // 3, tonemapper, white_point, gamma, scale, _p(out))
    return out


def f32_to_f16(x):
    """half::f16::from_f32 bit patterns (uint16) of a float32 array."""
    a = np.ascontiguousarray(x, np.float32)
    out = np.zeros(a.shape, np.uint16)
    lib().orc_f32_to_f16(_p(a), a.size, _p(out))
    return out


def exr_blocks(film, scale=1.0):
    """The pre-deflate bytes of each 16-line ZIP block of the B, G, R half
    image of `film` (h, w, 3), blocks back to back."""
    f = np.ascontiguousarray(film, np.float32)
    h, w = f.shape[:2]
    out = np.zeros(h * w * 6, np.uint8)
    lib().orc_exr_blocks(_p(f), w, h, scale, _p(out))
    return out


def ray_new(o, d, tmin, tmax):
    out = np.zeros(8, np.float32)
    lib().orc_ray_new(_p(f32(*o)), _p(f32(*d)), tmin, tmax, _p(out))
    return out


def triangle_hit(v0, v1, v2, ray8):
    out = np.zeros(6, np.float32)
    ok = lib().orc_triangle_hit(_p(f32(*v0)), _p(f32(*v1)), _p(f32(*v2)), _p(np.asarray(ray8, np.float32)), _p(out))
    return (out if ok else None)


def aabb_hit(lo, hi, ray8):
    return bool(lib().orc_aabb_hit(_p(f32(*lo)), _p(f32(*hi)), _p(np.asarray(ray8, np.float32))))


def camera_lookat(lookfrom, lookat, vup, vfov, aspect):
    out = np.zeros(16, np.float32)
    lib().orc_camera_lookat(_p(f32(*lookfrom)), _p(f32(*lookat)), _p(f32(*vup)), vfov, aspect, _p(out))
    return out


def camera_ray(cam16, u, v):
    out = np.zeros(8, np.float32)
    lib().orc_camera_ray(_p(np.asarray(cam16, np.float32)), u, v, _p(out))
    return out


def camera_sample_we(cam16, p, w, h):
    out = np.zeros(8, np.float32)
    ok = lib().orc_camera_sample_we(_p(np.asarray(cam16, np.float32)), _p(f32(*p)), w, h, _p(out))
    return out if ok else None


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox(_p(c), _p(k), _p(out))
    return out


def counter_sample(seed, pixel, sample, stream, dim):
    return lib().orc_counter_sample(seed, pixel, sample, stream, dim)


def ref_sampler(x, y, n):
    out = np.zeros(n, np.float32)
    lib().orc_ref_sampler(x, y, n, _p(out))
    return out


def _mp(albedo=(0.5, 0.5, 0.5), roughness=0.5, eta=1.5):
    return np.array(list(albedo) + [roughness, eta, 0, 0, 0], np.float32)


def bsdf_eval(mtype, wi, wo, **kw):
    out = np.zeros(3, np.float32)
    lib().orc_bsdf_eval(mtype, _p(_mp(**kw)), _p(f32(*wi)), _p(f32(*wo)), _p(out))
    return out


def bsdf_pdf(mtype, wi, wo, **kw):
    return lib().orc_bsdf_pdf(mtype, _p(_mp(**kw)), _p(f32(*wi)), _p(f32(*wo)))


def bsdf_sample(mtype, wo, u_sel, ux, uy, **kw):
    out = np.zeros(7, np.float32)
    ok = lib().orc_bsdf_sample(mtype, _p(_mp(**kw)), _p(f32(*wo)), u_sel, ux, uy, _p(out))
    return out if ok else None
