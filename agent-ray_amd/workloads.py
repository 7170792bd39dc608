"""Synthetic workloads for the traversal and path-sampling hot path.

Every generator is seeded and deterministic (numpy PCG64).  Scenes are
returned as plain descriptions (``SceneDesc``) that both the HIP library
(``Scene.from_desc``) and the CPU oracle consume, so the two always see the
same bits.

Workloads (sizes as named in the benchmark configurations):

* ``cornell_box``       - Cornell box, 10-triangle room + two rotated boxes +
                          one quad light (radiance 15), 256x256, 16 spp, depth 8.
* ``triangle_soup``     - 2^20 random triangles in the unit cube and 2^24
                          rays (half coherent pinhole primaries, half
                          incoherent) for closest-hit / any-hit timing.
* ``procedural_room``   - ~280k triangles: 12 level-5 icospheres, 4 displaced
                          64x64 grids, 100 boxes, mixed Lambertian / GGX
                          conductor / GGX dielectric, 3 quad lights.
* ``serving_jobs``      - Poisson job arrivals with mixed resolution, spp and
                          integrator against a resident scene.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

MAT_LAMBERTIAN, MAT_CONDUCTOR, MAT_DIELECTRIC = 0, 1, 2
LIGHT_POINT, LIGHT_QUAD = 0, 1
F32_MAX = np.float32(3.4028234663852886e38)


@dataclass
class MaterialDesc:
    type: int
    albedo: tuple = (0.5, 0.5, 0.5)  # Lambertian albedo or conductor F0
    roughness: float = 0.5
    eta: float = 1.5
    texture: int = -1                # Lambertian albedo texture (index into SceneDesc.textures), -1: none


@dataclass
class MeshDesc:
    positions: np.ndarray            # (nv, 3) float32
    indices: np.ndarray              # (nt, 3) uint32
    materials: List[int]             # material ids by slot
    slots: np.ndarray = None         # (1,) or (nt,) uint32
    normals: Optional[np.ndarray] = None  # (nv, 3) float32 or None
    transform: Optional[np.ndarray] = None  # (16,) float32 column-major, None = identity
    uvs: Optional[np.ndarray] = None  # (nv, 2) float32 tex_coords or None

    def __post_init__(self):
        self.positions = np.ascontiguousarray(self.positions, dtype=np.float32).reshape(-1, 3)
        self.indices = np.ascontiguousarray(self.indices, dtype=np.uint32).reshape(-1, 3)
        if self.slots is None:
            self.slots = np.zeros(1, np.uint32)
        self.slots = np.ascontiguousarray(self.slots, dtype=np.uint32).reshape(-1)
        if self.uvs is not None:
            self.uvs = np.ascontiguousarray(self.uvs, dtype=np.float32).reshape(-1, 2)
        if self.normals is not None:
            self.normals = np.ascontiguousarray(self.normals, dtype=np.float32).reshape(-1, 3)
        if self.transform is not None:
            self.transform = np.ascontiguousarray(self.transform, dtype=np.float32).reshape(16)


@dataclass
class LightDesc:
    type: int
    position: tuple
    intensity: tuple
    edge_u: tuple = (0.0, 0.0, 0.0)
    edge_v: tuple = (0.0, 0.0, 0.0)


@dataclass
class CameraDesc:
    lookfrom: tuple
    lookat: tuple
    vup: tuple = (0.0, 1.0, 0.0)
    vfov: float = 40.0


@dataclass
class SceneDesc:
    materials: List[MaterialDesc] = field(default_factory=list)
    meshes: List[MeshDesc] = field(default_factory=list)
    textures: List[np.ndarray] = field(default_factory=list)  # (h, w, 3) uint8 sRGB images, top row first
    lights: List[LightDesc] = field(default_factory=list)
    light_dist: int = 2  # 0 none, 1 uniform, 2 power (the reference binaries use power)
    camera: Optional[CameraDesc] = None
    width: int = 0
    height: int = 0
    spp: int = 16
    max_depth: int = 8
    rr_depth: int = 3

    @property
    def n_triangles(self) -> int:
        n = sum(int(m.indices.shape[0]) for m in self.meshes)
        return n + 2 * sum(1 for l in self.lights if l.type == LIGHT_QUAD)

    def bounds(self, shrink: float = 0.01):
        """World AABB of all meshes (transforms applied), shrunk by `shrink`
        of its extent on every side."""
        lo = np.full(3, np.inf)
        hi = np.full(3, -np.inf)
        for m in self.meshes:
            p = m.positions.astype(np.float64)
            if m.transform is not None:
                t = np.asarray(m.transform, np.float64).reshape(4, 4).T  # column-major storage
                p = p @ t[:3, :3].T + t[:3, 3]
            lo = np.minimum(lo, p.min(0))
            hi = np.maximum(hi, p.max(0))
        e = (hi - lo) * shrink
        return (lo + e).astype(np.float32), (hi - e).astype(np.float32)


# ---------------------------------------------------------------------------
# fp32 helpers with the reference's operation order
# ---------------------------------------------------------------------------
f32 = np.float32


def normalize_f32(d: np.ndarray) -> np.ndarray:
    """glam normalize: v * (1 / sqrt((x*x + y*y) + z*z)) in fp32."""
    d = np.asarray(d, dtype=np.float32)
    x, y, z = d[..., 0], d[..., 1], d[..., 2]
    dd = (x * x + y * y) + z * z
    r = f32(1.0) / np.sqrt(dd)
    return d * r[..., None]


def cross_f32(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], axis=-1)


def camera_lookat(lookfrom, lookat, vup, vfov, aspect) -> np.ndarray:
    """PinholeCamera::from_lookat (cameras.rs:91-123) in fp32.

    Returns the 16-float camera state: origin, lower_left, horizontal,
    vertical, forward, area.
    """
    rads = f32(f32(math.pi) / f32(180.0))
    theta = f32(f32(vfov) * rads)
    h = f32(math.tan(float(f32(theta / f32(2.0)))))
    vh = f32(f32(2.0) * h)
    vw = f32(f32(aspect) * vh)
    frm = np.asarray(lookfrom, np.float32)
    at = np.asarray(lookat, np.float32)
    up = np.asarray(vup, np.float32)
    w = normalize_f32(frm - at)
    u = normalize_f32(cross_f32(up, w))
    v = cross_f32(w, u)
    horiz = (u * vw).astype(np.float32)
    vert = (v * vh).astype(np.float32)
    llc = ((frm - horiz / f32(2.0)) - vert / f32(2.0)) - w
    fwd = -w
    return np.concatenate([frm, llc, horiz, vert, fwd, [f32(vw * vh)]]).astype(np.float32)


def camera_rays(cam: np.ndarray, u: np.ndarray, v: np.ndarray) -> np.ndarray:
    """PinholeCamera::generate_ray (cameras.rs:146-149): (n, 8) float32 rays."""
    u = np.asarray(u, np.float32)[:, None]
    v = np.asarray(v, np.float32)[:, None]
    o, llc, hz, vt = cam[0:3], cam[3:6], cam[6:9], cam[9:12]
    d = ((llc + u * hz) + v * vt) - o
    d = normalize_f32(d)
    n = d.shape[0]
    out = np.empty((n, 8), np.float32)
    out[:, 0:3] = o
    out[:, 3] = f32(0.001)
    out[:, 4:7] = d
    out[:, 7] = F32_MAX
    return out


def mat_translate_rotate_y_scale(t, deg, s) -> np.ndarray:
    """Column-major 4x4 = T(t) * Ry(deg) * S(s), composed in float64 then rounded."""
    c, sn = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    R = np.array([[c, 0, sn], [0, 1, 0], [-sn, 0, c]], np.float64)
    S = np.diag(np.asarray(s, np.float64) * np.ones(3))
    M = np.eye(4)
    M[:3, :3] = R @ S
    M[:3, 3] = t
    return M.T.reshape(16).astype(np.float32)  # column-major


def unit_cube_mesh():
    """Axis-aligned cube [-0.5, 0.5]^3, 8 vertices, 12 triangles."""
    p = np.array([[x, y, z] for z in (-0.5, 0.5) for y in (-0.5, 0.5) for x in (-0.5, 0.5)], np.float32)

    def q(a, b, c, d):
        return [[a, b, c], [a, c, d]]

    idx = (q(0, 2, 3, 1) + q(4, 5, 7, 6) + q(0, 1, 5, 4) + q(2, 6, 7, 3) + q(0, 4, 6, 2) + q(1, 3, 7, 5))
    return p, np.array(idx, np.uint32)


# ---------------------------------------------------------------------------
# Cornell box
# ---------------------------------------------------------------------------
def textured_cornell_box(seed: int = 0, point_light_twin: bool = True) -> SceneDesc:
    """cornell_box with image textures (texture.rs:65-135): the floor and one
    wall take an 8x6 texture through per-vertex tex_coords that run past
    [0, 1] and below 0 (repeat wrap), and the tall box a 3x5 texture
    addressed by the hit's barycentrics (it has no tex_coords; scene.rs:55-64).
    Texels are random 8-bit sRGB values (seeded)."""
    desc = cornell_box(seed=seed, point_light_twin=point_light_twin)
    rng = np.random.default_rng(seed + 1000)
    desc.textures = [rng.integers(0, 256, (6, 8, 3), dtype=np.uint8), rng.integers(0, 256, (5, 3, 3), dtype=np.uint8)]
    room = desc.meshes[0]
    p = room.positions
    room.uvs = np.stack([p[:, 0] * 2.0 + p[:, 2] * 1.5 - 0.3, p[:, 1] * 2.0 - 0.25], -1).astype(np.float32)
    desc.materials[0].texture = 0
    desc.materials[2].texture = 0
    desc.materials[desc.meshes[1].materials[0]].texture = 1
    return desc


def cornell_box(seed: int = 0, width: int = 256, height: int = 256, spp: int = 16, max_depth: int = 8,
                point_light_twin: bool = False) -> SceneDesc:
    """Procedural Cornell box: unit-cube room (5 faces, 10 triangles, front
    open), a tall box 0.3x0.6x0.3 rotated +18 deg and a short box 0.3^3 rotated
    -18 deg about Y, and a quad light [0.4,0.6]x{0.999}x[0.4,0.6] facing -Y
    with radiance 15.  Lambertian albedos per object and channel U[0.2,0.8].

    ``point_light_twin`` replaces the quad by a point light at its centre with
    the same power (I = pi*A*L / 4pi), the reference's own light type.
    """
    rng = np.random.default_rng(seed)
    desc = SceneDesc(width=width, height=height, spp=spp, max_depth=max_depth, rr_depth=3)
    alb = rng.uniform(0.2, 0.8, size=(7, 3))
    for a in alb:
        desc.materials.append(MaterialDesc(MAT_LAMBERTIAN, tuple(float(x) for x in a)))
    P = np.array([[0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1],
                  [0, 1, 0], [1, 1, 0], [1, 1, 1], [0, 1, 1]], np.float32)
    faces = [(0, 1, 2, 3),  # floor y=0
             (4, 7, 6, 5),  # ceiling y=1
             (3, 2, 6, 7),  # back z=1
             (0, 3, 7, 4),  # left x=0
             (1, 5, 6, 2)]  # right x=1
    idx, slots = [], []
    for fi, (a, b, c, d) in enumerate(faces):
        idx += [[a, b, c], [a, c, d]]
        slots += [fi, fi]
    desc.meshes.append(MeshDesc(P, np.array(idx, np.uint32), materials=[0, 1, 2, 3, 4],
                                slots=np.array(slots, np.uint32)))
    cp, ci = unit_cube_mesh()
    desc.meshes.append(MeshDesc(cp, ci, materials=[5],
                                transform=mat_translate_rotate_y_scale((0.33, 0.3, 0.62), 18.0, (0.3, 0.6, 0.3))))
    desc.meshes.append(MeshDesc(cp, ci, materials=[6],
                                transform=mat_translate_rotate_y_scale((0.66, 0.15, 0.35), -18.0, (0.3, 0.3, 0.3))))
    L = 15.0
    if point_light_twin:
        area = 0.2 * 0.2
        I = math.pi * area * L / (4.0 * math.pi)
        desc.lights.append(LightDesc(LIGHT_POINT, (0.5, 0.99, 0.5), (I, I, I)))
    else:
        desc.lights.append(LightDesc(LIGHT_QUAD, (0.4, 0.999, 0.4), (L, L, L), (0.2, 0.0, 0.0), (0.0, 0.0, 0.2)))
    desc.camera = CameraDesc((0.5, 0.5, -1.37), (0.5, 0.5, 0.5), (0.0, 1.0, 0.0), 40.0)
    return desc


# ---------------------------------------------------------------------------
# Triangle soup + ray sets (traversal microbenchmark)
# ---------------------------------------------------------------------------
def triangle_soup(n_tris: int = 1 << 20, seed: int = 0, size_range=(1e-3, 2e-2)) -> SceneDesc:
    """n random triangles: centroid U[0,1]^3, size s = exp(U[ln lo, ln hi])
    for size_range (lo, hi), vertices c + s * (independent uniform unit
    vectors)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 1.0, size=(n_tris, 1, 3))
    s = np.exp(rng.uniform(math.log(size_range[0]), math.log(size_range[1]), size=(n_tris, 1, 1)))
    g = rng.standard_normal(size=(n_tris, 3, 3))
    g /= np.linalg.norm(g, axis=2, keepdims=True)
    v = (c + s * g).astype(np.float32).reshape(-1, 3)
    idx = np.arange(3 * n_tris, dtype=np.uint32).reshape(-1, 3)
    desc = SceneDesc(light_dist=0)
    desc.materials.append(MaterialDesc(MAT_LAMBERTIAN, (0.5, 0.5, 0.5)))
    desc.meshes.append(MeshDesc(v, idx, materials=[0]))
    return desc


def soup_rays(n_rays: int = 1 << 24, seed: int = 1, image: int = 4096):
    """Ray set for the soup: first half coherent pinhole primaries (pixel
    centres of the middle rows of an image x image view from (0.5,0.5,2.5)
    toward the cube centre, vfov 40), second half incoherent (origin
    U[0,1]^3, direction uniform on the sphere, exact-zero components
    resampled, t in [0, f32::MAX]).  Returns (n, 8) float32."""
    half = n_rays // 2
    cam = camera_lookat((0.5, 0.5, 2.5), (0.5, 0.5, 0.5), (0, 1, 0), 40.0, 1.0)
    rows = max(1, half // image)
    y0 = (image - rows) // 2
    k = np.arange(half, dtype=np.int64)
    y = y0 + k // image
    x = k % image
    u = (x.astype(np.float32) + f32(0.5)) / f32(image)
    v = f32(1.0) - (y.astype(np.float32) + f32(0.5)) / f32(image)
    coh = camera_rays(cam, u, v)
    rng = np.random.default_rng(seed)
    m = n_rays - half
    o = rng.uniform(0.0, 1.0, size=(m, 3)).astype(np.float32)
    d = np.zeros((m, 3), np.float32)
    todo = np.arange(m)
    while todo.size:
        z = 1.0 - 2.0 * rng.uniform(size=todo.size)
        phi = 2.0 * math.pi * rng.uniform(size=todo.size)
        r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
        dd = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1).astype(np.float32)
        d[todo] = dd
        bad = np.any(dd == 0.0, axis=1)
        todo = todo[bad]
    d = normalize_f32(d)
    inc = np.empty((m, 8), np.float32)
    inc[:, 0:3] = o
    inc[:, 3] = 0.0
    inc[:, 4:7] = d
    inc[:, 7] = F32_MAX
    return np.ascontiguousarray(np.concatenate([coh, inc], 0))


# ---------------------------------------------------------------------------
# Procedural room
# ---------------------------------------------------------------------------
def icosphere(level: int):
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}
        nf = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts, np.float64)
    return V.astype(np.float32), np.array(faces, np.uint32), V.astype(np.float32)


def displaced_grid(n: int, size: float, amp: float, rng):
    xs = np.linspace(0.0, size, n + 1)
    X, Z = np.meshgrid(xs, xs, indexing="xy")
    ph = rng.uniform(0, 2 * math.pi, size=4)
    Y = amp * (np.sin(X * 2.1 + ph[0]) * np.cos(Z * 1.7 + ph[1]) + 0.5 * np.sin(X * 5.3 + Z * 4.1 + ph[2]))
    P = np.stack([X, Y, Z], -1).reshape(-1, 3).astype(np.float32)
    idx = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            b, c, d = a + 1, a + n + 2, a + n + 1
            idx += [(a, d, c), (a, c, b)]
    return P, np.array(idx, np.uint32)


def _pick_material(rng, desc: SceneDesc) -> int:
    r = rng.uniform()
    if r < 0.60:
        m = MaterialDesc(MAT_LAMBERTIAN, tuple(float(x) for x in rng.uniform(0.2, 0.8, 3)))
    elif r < 0.85:
        m = MaterialDesc(MAT_CONDUCTOR, tuple(float(x) for x in rng.uniform(0.5, 1.0, 3)),
                         roughness=float(rng.uniform(0.05, 0.5)))
    else:
        m = MaterialDesc(MAT_DIELECTRIC, (1.0, 1.0, 1.0), roughness=0.1, eta=1.5)
    desc.materials.append(m)
    return len(desc.materials) - 1


def procedural_room(seed: int = 0, width: int = 1920, height: int = 1080, spp: int = 64, max_depth: int = 16,
                    n_boxes: int = 100) -> SceneDesc:
    """~280k-triangle room: walls/floor/ceiling (Lambertian), 12 level-5
    icospheres (smooth vertex normals), 4 displaced 64x64 grids, n_boxes
    seed-placed boxes; per-object material 60% Lambertian / 25% GGX
    conductor (alpha U[0.05,0.5], F0 U[0.5,1]^3) / 15% dielectric (eta 1.5,
    alpha 0.1); 3 quad lights under the ceiling."""
    rng = np.random.default_rng(seed)
    desc = SceneDesc(width=width, height=height, spp=spp, max_depth=max_depth, rr_depth=3)
    W, H, D = 8.0, 4.0, 8.0
    wall = MaterialDesc(MAT_LAMBERTIAN, (0.7, 0.7, 0.65))
    desc.materials.append(wall)
    P = np.array([[0, 0, 0], [W, 0, 0], [W, 0, D], [0, 0, D], [0, H, 0], [W, H, 0], [W, H, D], [0, H, D]],
                 np.float32)
    faces = [(0, 1, 2, 3), (4, 7, 6, 5), (3, 2, 6, 7), (0, 3, 7, 4), (1, 5, 6, 2), (0, 4, 5, 1)]
    idx = []
    for a, b, c, d in faces:
        idx += [[a, b, c], [a, c, d]]
    desc.meshes.append(MeshDesc(P, np.array(idx, np.uint32), materials=[0]))
    sv, sf, sn = icosphere(5)
    for i in range(12):
        r = float(rng.uniform(0.3, 0.7))
        c = (float(rng.uniform(1, W - 1)), r + float(rng.uniform(0.0, 1.5)), float(rng.uniform(1, D - 1)))
        m = _pick_material(rng, desc)
        desc.meshes.append(MeshDesc(sv, sf, materials=[m], normals=sn,
                                    transform=mat_translate_rotate_y_scale(c, float(rng.uniform(0, 360)), r)))
    for i in range(4):
        gp, gi = displaced_grid(64, 3.0, 0.15, rng)
        m = _pick_material(rng, desc)
        off = (0.5 + (i % 2) * 4.0, 0.2, 0.5 + (i // 2) * 4.0)
        gp = gp + np.asarray(off, np.float32)
        desc.meshes.append(MeshDesc(gp, gi, materials=[m]))
    cp, ci = unit_cube_mesh()
    for i in range(n_boxes):
        s = rng.uniform(0.1, 0.5, 3)
        c = (float(rng.uniform(0.5, W - 0.5)), float(s[1] / 2 + rng.uniform(0, 2.5)), float(rng.uniform(0.5, D - 0.5)))
        m = _pick_material(rng, desc)
        desc.meshes.append(MeshDesc(cp, ci, materials=[m],
                                    transform=mat_translate_rotate_y_scale(c, float(rng.uniform(0, 90)), tuple(s))))
    for k, (x, z) in enumerate([(2.0, 2.0), (5.5, 3.0), (3.5, 6.0)]):
        Lr = (12.0, 11.0, 9.5) if k != 1 else (8.0, 9.0, 12.0)
        desc.lights.append(LightDesc(LIGHT_QUAD, (x, H - 0.01, z), Lr, (0.8, 0.0, 0.0), (0.0, 0.0, 0.8)))
    desc.camera = CameraDesc((4.0, 2.0, 0.3), (4.0, 1.2, 5.0), (0.0, 1.0, 0.0), 60.0)
    return desc


# ---------------------------------------------------------------------------
# Serving job mix
# ---------------------------------------------------------------------------
@dataclass
class JobSpec:
    index: int
    arrival: float
    width: int
    height: int
    spp: int
    integrator: int  # 0 PT, 1 BDPT, 2 GDPT
    seed: int


def serving_jobs(rate: float = 20.0, duration: float = 60.0, seed: int = 0, min_side: int = 64,
                 max_side: int = 1024, max_spp: int = 256) -> List[JobSpec]:
    """Poisson arrivals at `rate` jobs/s for `duration` s; square side
    U{min_side..max_side step 64}; spp = round(exp(U[0, ln max_spp]));
    integrator PT 60% / BDPT 25% / GDPT 15%; seed = job index."""
    rng = np.random.default_rng(seed)
    jobs, t, i = [], 0.0, 0
    sides = np.arange(min_side, max_side + 1, 64)
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            break
        side = int(rng.choice(sides))
        spp = max(1, int(round(math.exp(rng.uniform(0.0, math.log(max_spp))))))
        r = rng.uniform()
        integ = 0 if r < 0.60 else (1 if r < 0.85 else 2)
        jobs.append(JobSpec(i, t, side, side, spp, integ, i))
        i += 1
    return jobs
