"""The library's traversal code (agent-ray_amd/csrc/ar_traverse.h, written
for host and device) compiled for the host with g++ and compared with the
oracle ray by ray, without a GPU.

It walks the same trees the device gets (build_host_accel: binary trees for
reference order, quantized 4-wide trees for near-first order, built from
the reference's SAH tree or from the spatial-split tree) and runs each
query two ways: one ray after another, and in an emulation of the kernels'
64-lane waves with wave-local refill and maximal leaf parking (three and
four park slots).  Checks:
* reference order reproduces the oracle's (t bits, prim, instance) on every
  ray, closest and any hit;
* near-first order reproduces them too, except near-ties (t within 4 ulps,
  never farther) and at most a stated count;
* both drivers agree bit for bit (the visiting order does not change the
  answer) and the wave emulation starts every ray exactly once.
Sets: random soups, Cornell instances (rotated / scaled), the room, rays
starting on triangle vertices with a zero direction component (the slab
test's 0 * inf cases), and duplicated geometry where every hit is a tie."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "host", "host_traverse.cpp")
CSRC = os.path.join(ROOT, "agent-ray_amd", "csrc")
OUT = os.path.join(ROOT, "tests", "host", "_build", "libhost_traverse.so")
LIB = os.path.join(ROOT, "agent-ray_amd", "libagentray.so")


def _helper(counts=False):
    """The host harness library; with counts=True the build whose walker
    counts the records it reads (host_counts)."""
    out = OUT.replace(".so", "_counts.so") if counts else OUT
    deps = [SRC, LIB] + [os.path.join(CSRC, f) for f in ("ar_traverse.h", "ar_device.h", "ar_internal.h")]
    if not os.path.exists(out) or os.path.getmtime(out) < max(os.path.getmtime(d) for d in deps):
        if not os.path.exists("/opt/rocm/include/hip/hip_runtime.h"):
            pytest.skip("HIP headers not installed")
        os.makedirs(os.path.dirname(out), exist_ok=True)
        subprocess.check_call(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                               "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include")]
                              + (["-DHOST_COUNTS"] if counts else [])
                              + [SRC, "-o", out + ".tmp", "-L" + os.path.dirname(LIB), "-lagentray",
                                 "-Wl,-rpath," + os.path.dirname(LIB)])
        os.replace(out + ".tmp", out)
    h = C.CDLL(out)
    if counts:
        h.host_counts.argtypes = [C.POINTER(C.c_ulonglong)]
        h.host_counts.restype = None
    h.host_traverse.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64] + [C.c_int] * 4 + [C.c_void_p] * 6
    h.host_traverse.restype = C.c_int
    h.host_chunk_schedule.argtypes = [C.c_uint64, C.c_uint32, C.c_int, C.c_uint32, C.c_uint64,
                                      C.POINTER(C.c_uint64)]
    h.host_chunk_schedule.restype = C.c_int
    h.host_max_leaf.argtypes = [C.c_void_p, C.c_void_p]
    h.host_max_leaf.restype = C.c_int
    h.host_record_counts.argtypes = [C.c_void_p, C.c_void_p]
    h.host_record_counts.restype = C.c_int
    return h


BUILDERS = {"sah": 0, "split": 2}


def _scene(ar, desc, builder="split"):
    """The scene as the library holds it before commit (no device needed)."""
    lib = ar.lib()
    s = C.c_void_p()
    assert lib.ar_scene_create(C.byref(s)) == 0
    assert lib.ar_scene_set_builder(s, BUILDERS[builder]) == 0
    mids = []
    for m in desc.materials:
        out = C.c_uint32()
        assert lib.ar_scene_add_material(s, C.byref(ar._mat_from_desc(m).to_c()), C.byref(out)) == 0
        mids.append(out.value)
    p = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
    for m in desc.meshes:
        pos = np.ascontiguousarray(m.positions, np.float32)
        idx = np.ascontiguousarray(m.indices, np.uint32)
        nrm = None if m.normals is None else np.ascontiguousarray(m.normals, np.float32)
        slots = np.zeros(1, np.uint32) if m.slots is None else np.ascontiguousarray(m.slots, np.uint32).reshape(-1)
        mats = np.asarray([mids[i] for i in m.materials], np.uint32)
        xf = None if m.transform is None else np.ascontiguousarray(m.transform, np.float32).reshape(16)
        assert lib.ar_scene_add_mesh(s, p(pos), pos.shape[0], p(idx), idx.shape[0], p(nrm), p(slots),
                                     slots.shape[0], p(mats), mats.shape[0], p(xf), None) == 0
    for l in desc.lights:
        lt = ar.PointLight(l.position, l.intensity) if l.type == 0 else \
            ar.QuadLight(l.position, l.edge_u, l.edge_v, l.intensity)
        assert lib.ar_scene_add_light(s, C.byref(lt.to_c()), None) == 0
    return s


def _run(h, s, rays, any_hit, near, lds_stack=4, slots=3):
    n = rays.shape[0]
    bufs = [np.zeros(n, np.float32) if k % 3 == 0 else np.zeros(n, np.uint32) for k in range(6)]
    rc = h.host_traverse(s, rays.ctypes.data, n, int(any_hit), int(near), lds_stack, slots,
                         *[b.ctypes.data for b in bufs])
    assert rc == 0, rc
    return bufs[:3], bufs[3:]


def _near_ties(a, b):
    ta, tb = a[0].view(np.uint32).astype(np.int64), b[0].view(np.uint32).astype(np.int64)
    exact = (ta == tb) & (a[1] == b[1]) & (a[2] == b[2])
    both = (a[1] != 0xFFFFFFFF) & (b[1] != 0xFFFFFFFF)
    ties = (~exact) & both & (np.abs(ta - tb) <= 4) & (a[0] <= b[0])
    return int((~exact).sum()), int(ties.sum())


def _check(ar, orc, desc, rays, max_near_ties=0):
    """Both tree builders' scenes against the oracle: the reference's SAH for
    both tree forms, and the spatial-split 4-wide trees (the default)."""
    osc = orc.OracleScene(desc)
    ot, op, oi, _, _ = osc.trace(rays)
    occ, _, _ = osc.occluded(rays)
    for builder in BUILDERS:
        _check_one(ar, desc, rays, (ot, op, oi), occ, max_near_ties, builder)


def _check_one(ar, desc, rays, ora, occ, max_near_ties, builder):
    h = _helper()
    rays = np.ascontiguousarray(rays, np.float32)
    s = _scene(ar, desc, builder)
    try:
        for near, slots in ((False, 3), (True, 3), (True, 4)):
            one, wave = _run(h, s, rays, False, near, slots=slots)
            for k in range(3):  # the two drivers agree bit for bit
                assert np.array_equal(one[k].view(np.uint32), wave[k].view(np.uint32)), (near, slots, k)
            bad, ties = _near_ties(one, ora)
            if near:
                assert bad == ties and ties <= max_near_ties, (builder, bad, ties)
            else:
                assert bad == 0, (builder, bad)
            a1, a2 = _run(h, s, rays, True, near, slots=slots)
            assert np.array_equal(a1[0].astype(bool), occ.astype(bool)), (builder, near, slots)
            assert np.array_equal(a2[0].astype(bool), occ.astype(bool)), (builder, near, slots)
    finally:
        ar.lib().ar_scene_destroy(s)


def _interior_rays(wl, lo, hi, n, seed, tmin=0.0):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = rng.uniform(lo, hi, (n, 3))
    r[:, 3] = tmin
    r[:, 4:7] = wl.normalize_f32(rng.standard_normal((n, 3)).astype(np.float32))
    r[:, 7] = wl.F32_MAX
    r[::5, 7] = rng.uniform(0.0, 0.5, r[::5].shape[0])  # short (shadow-like) rays
    return r


def test_soup(ar, orc, wl):
    _check(ar, orc, wl.triangle_soup(1 << 14, seed=3), wl.soup_rays(1 << 14, seed=4, image=128))


def test_cornell_instances(ar, orc, wl):
    _check(ar, orc, wl.cornell_box(), _interior_rays(wl, 0.02, 0.98, 20000, 5), max_near_ties=2)


def test_room(ar, orc, wl):
    desc = wl.procedural_room(n_boxes=30)
    _check(ar, orc, desc, _interior_rays(wl, (0.1, 0.1, 0.1), (7.9, 3.9, 7.9), 12000, 6, 1e-4), max_near_ties=2)


def test_rays_on_box_planes(ar, orc, wl):
    desc = wl.triangle_soup(1 << 12, seed=7)
    rng = np.random.default_rng(8)
    P = desc.meshes[0].positions.reshape(-1, 3).astype(np.float32)
    m = 6000
    o = P[rng.integers(0, P.shape[0], m)].copy()
    for ax in range(3):
        sel = rng.random(m) < 0.5
        o[sel, ax] = P[rng.integers(0, P.shape[0], int(sel.sum())), ax]
    d = np.zeros((m, 3), np.float32)
    d[np.arange(m), rng.integers(0, 3, m)] = np.where(rng.random(m) < 0.5, -1.0, 1.0)
    k = m // 3
    d2 = rng.standard_normal((k, 3)).astype(np.float32)
    d2[np.arange(k), rng.integers(0, 3, k)] = 0.0
    d[:k] = wl.normalize_f32(d2)
    r = np.zeros((m, 8), np.float32)
    r[:, 0:3], r[:, 4:7], r[:, 7] = o, d, wl.F32_MAX
    _check(ar, orc, desc, r)


def test_duplicate_geometry_ties(ar, orc, wl):
    base = wl.triangle_soup(1 << 11, seed=12)
    m = base.meshes[0]
    idx = m.indices.reshape(-1, 3)
    dup = wl.MeshDesc(positions=m.positions, indices=np.concatenate([idx, idx[::-1]]).astype(np.uint32),
                      materials=m.materials, slots=m.slots, normals=None, transform=m.transform)
    desc = wl.SceneDesc(materials=base.materials, meshes=[dup, dup], lights=base.lights, camera=base.camera)
    _check(ar, orc, desc, wl.soup_rays(1 << 14, seed=13, image=128))


def test_leaves_of_more_than_four_triangles(ar, orc, wl):
    """The SAH leaf rule keeps a leaf of n > 4 triangles when no split costs
    less than n (here: seven coincident copies of each triangle, where every
    split costs 1 + n) or the node's area is below f32::EPSILON (a cluster of
    point triangles).  The 4-wide collapse must keep such a leaf as a range
    of Tri records, not as a node with a slot per triangle."""
    base = wl.triangle_soup(1 << 9, seed=14)
    m = base.meshes[0]
    idx = m.indices.reshape(-1, 3)
    sev = np.repeat(idx, 7, axis=0).astype(np.uint32)
    pts = np.repeat(np.float32([[0.25, 0.5, 0.5], [0.75, 0.5, 0.5]]), 3, axis=0)
    pidx = np.tile(np.arange(6, dtype=np.uint32).reshape(2, 3), (6, 1))  # 6 point triangles at each of 2 points
    meshes = [wl.MeshDesc(positions=m.positions, indices=sev, materials=m.materials, slots=m.slots, normals=None,
                          transform=m.transform),
              wl.MeshDesc(positions=pts, indices=pidx, materials=m.materials, slots=None, normals=None,
                          transform=None)]
    desc = wl.SceneDesc(materials=base.materials, meshes=meshes, lights=base.lights, camera=base.camera)
    h = _helper()
    s = _scene(ar, desc)
    try:
        sizes = np.zeros(2, np.uint32)
        assert h.host_max_leaf(s, sizes.ctypes.data) == 0
        assert sizes[0] >= 7 and sizes[1] >= 5, sizes  # the scene really has such leaves
    finally:
        ar.lib().ar_scene_destroy(s)
    _check(ar, orc, desc, wl.soup_rays(1 << 13, seed=15, image=64))


def test_split_and_binary_objects_in_one_scene(ar, orc, wl):
    """The split builder keeps a spatial-split tree only where it lowers the
    SAH cost by 5%: here for a soup of large triangles (sizes 0.02-0.2 in the
    unit cube: 1.9 references per triangle, SAH cost 114 -> 92), not for the
    Cornell instances beside it.  The scene's 4-wide trees then read their own
    records, the Cornell objects' in their binary leaf order; answers equal
    the oracle's for both builders."""
    cb = wl.cornell_box()
    desc = wl.SceneDesc(materials=cb.materials, meshes=wl.triangle_soup(4096, seed=16, size_range=(0.02, 0.2)).meshes + list(cb.meshes), lights=cb.lights,
                        camera=cb.camera)
    h = _helper()
    counts = np.zeros(2, np.uint64)
    for builder, more in (("sah", False), ("split", True)):
        s = _scene(ar, desc, builder)
        try:
            assert h.host_record_counts(s, counts.ctypes.data) == 0
        finally:
            ar.lib().ar_scene_destroy(s)
        assert (counts[1] > counts[0]) == more, (builder, counts)
    s = _scene(ar, wl.SceneDesc(materials=cb.materials, meshes=list(cb.meshes), lights=cb.lights, camera=cb.camera))
    try:  # no object keeps a split tree: one set of records
        assert h.host_record_counts(s, counts.ctypes.data) == 0 and counts[1] == 0, counts
    finally:
        ar.lib().ar_scene_destroy(s)
    _check(ar, orc, desc, _interior_rays(wl, 0.02, 0.98, 12000, 17), max_near_ties=2)


@pytest.mark.parametrize("waves", [1, 3, 16, 96])
def test_chunk_schedule_every_item_once(waves):
    """The walker's chunk bookkeeping (ChunkCursor, the code walk_items runs)
    under random interleavings of many waves: every item is started exactly
    once, with static shares and with the shared counter, at sizes around the
    4-chunks-per-wave threshold and ragged ends, for refill thresholds from 1
    to 64; the counter ends at zero (the last wave clears it), and it is used
    exactly when the launch has counters and at least 4 chunks per wave."""
    h = _helper()
    c = 4 * 64 * waves
    sizes = [0, 1, 63, 64, 65, 64 * waves - 1, 64 * waves + 7, c - 1, c, c + 1, 3 * c + 29, 9 * c + 64]
    for n in sizes:
        for counters in (0, 1):
            for refill_min in (1, 8, 64):
                for seed in (1, 2):
                    takes = C.c_uint64()
                    rc = h.host_chunk_schedule(n, waves, counters, refill_min, seed, C.byref(takes))
                    assert rc == 0, (n, waves, counters, refill_min, seed, rc)
                    dyn = counters and (n + 63) // 64 >= 4 * waves
                    assert (takes.value > 0) == bool(dyn), (n, waves, counters, takes.value)

