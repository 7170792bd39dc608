"""GPU parity of the traversal kernels (ar_trace / ar_occluded) against the
oracle's restatement of TLAS::intersect / Scene::occluded on the same rays.

Bar: bit-exact.  Reference order (AR_REF_ORDER) must reproduce the oracle's
(t bits, prim, instance) on every ray.  Near-first order must too, except on
a near-tie: two surfaces within 4 ulps of each other along the ray (two
triangles at the same t, or an object resting on another so that its face
and the floor meet the ray 1-2 ulps apart).  There the reference's answer
depends on its own visiting order (which box its slab test culls by an
ulp), while near-first returns the float-closest hit with ties broken by
instance record and leaf position, whatever order the wave imposes.
Near-ties are counted and bounded per scene."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _compare(gpu, ora):
    t, prim, inst = gpu
    ot, oprim, oinst = ora[0], ora[1], ora[2]
    tb, otb = t.view(np.uint32), ot.view(np.uint32)
    exact = (prim == oprim) & (tb == otb) & ((inst == oinst) | (oprim == 0xFFFFFFFF))
    both = (prim != 0xFFFFFFFF) & (oprim != 0xFFFFFFFF)
    ulps = np.abs(tb.astype(np.int64) - otb.astype(np.int64))  # t >= 0: bit distance = ulps
    ties = (~exact) & both & (ulps <= 4) & (t <= ot)
    return int((~exact).sum()), int(ties.sum())


def _check_scene(ar, orc, desc, rays, max_tie_frac=1e-4):
    sc = ar.Scene.from_desc(desc)
    osc = orc.OracleScene(desc)
    ora = osc.trace(rays)
    print(f"  scene: {desc.n_triangles} tris, rays {rays.shape[0]}", flush=True)
    bad_ref, _ = _compare(sc.intersect(rays, ref_order=True), ora)
    assert bad_ref == 0, f"reference-order mismatches: {bad_ref}"
    bad_near, ties = _compare(sc.intersect(rays, ref_order=False), ora)
    assert bad_near == ties, f"near-order mismatches that are not ties: {bad_near - ties}"
    assert ties <= max_tie_frac * rays.shape[0] + 1
    occ_o, _, _ = osc.occluded(rays)
    assert np.array_equal(sc.occluded(rays), occ_o)
    return sc, osc


def test_soup_closest_and_any(ar, orc, wl):
    desc = wl.triangle_soup(1 << 14, seed=21)
    rays = wl.soup_rays(1 << 16, seed=22, image=256)
    _check_scene(ar, orc, desc, rays)


def test_soup_short_rays_any_hit(ar, orc, wl):
    desc = wl.triangle_soup(1 << 13, seed=23)
    rays = wl.soup_rays(1 << 15, seed=24, image=128).copy()
    rng = np.random.default_rng(25)
    rays[:, 7] = rng.uniform(0.0, 0.2, rays.shape[0]).astype(np.float32)  # shadow-like t_max
    rays[::7, 3] = rays[::7, 7] + np.float32(0.1)                         # t_min > t_max: always miss
    _check_scene(ar, orc, desc, rays)


def _cornell_rays(wl, desc, n_side=128, seed=31):
    cam = wl.camera_lookat(desc.camera.lookfrom, desc.camera.lookat, desc.camera.vup, desc.camera.vfov, 1.0)
    k = np.arange(n_side * n_side)
    u = ((k % n_side).astype(np.float32) + np.float32(0.5)) / np.float32(n_side)
    v = np.float32(1) - ((k // n_side).astype(np.float32) + np.float32(0.5)) / np.float32(n_side)
    cr = wl.camera_rays(cam, u, v)
    rng = np.random.default_rng(seed)
    m = 1 << 14
    o = rng.uniform(0.02, 0.98, (m, 3)).astype(np.float32)
    d = wl.normalize_f32(rng.standard_normal((m, 3)).astype(np.float32))
    rr = np.zeros((m, 8), np.float32)
    rr[:, 0:3], rr[:, 4:7], rr[:, 7] = o, d, wl.F32_MAX
    # axis-aligned rays from interior points (zero direction components)
    ax = np.zeros((6, 8), np.float32)
    for i, dd in enumerate(np.eye(3).tolist() + (-np.eye(3)).tolist()):
        ax[i, 0:3] = (0.3, 0.4, 0.5)
        ax[i, 4:7] = dd
        ax[i, 7] = wl.F32_MAX
    return np.concatenate([cr, rr, ax]).astype(np.float32)


def test_cornell_instances(ar, orc, wl):
    # rotated + scaled instances exercise the TLAS dir_scale / renormalize path
    desc = wl.cornell_box()
    _check_scene(ar, orc, desc, _cornell_rays(wl, desc))


def test_room_scene(ar, orc, wl):
    desc = wl.procedural_room(n_boxes=30)
    cam = wl.camera_lookat(desc.camera.lookfrom, desc.camera.lookat, desc.camera.vup, desc.camera.vfov, 16 / 9)
    rng = np.random.default_rng(41)
    n = 1 << 15
    cr = wl.camera_rays(cam, rng.uniform(0, 1, n), rng.uniform(0, 1, n))
    o = rng.uniform((0.1, 0.1, 0.1), (7.9, 3.9, 7.9), (n, 3)).astype(np.float32)
    d = wl.normalize_f32(rng.standard_normal((n, 3)).astype(np.float32))
    rr = np.zeros((n, 8), np.float32)
    rr[:, 0:3], rr[:, 3], rr[:, 4:7], rr[:, 7] = o, 1e-4, d, wl.F32_MAX
    _check_scene(ar, orc, desc, np.concatenate([cr, rr]))


def _plane_rays(wl, desc, m, seed):
    """Rays that start on triangle vertices (so on leaf-box planes) with a
    zero direction component: the slab test meets 0 * inf = NaN there and
    the reference's SSE min/max order decides the box."""
    rng = np.random.default_rng(seed)
    P = np.concatenate([mm.positions.reshape(-1, 3) for mm in desc.meshes]).astype(np.float32)
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
    return r


def test_axis_rays_on_box_planes(ar, orc, wl):
    desc = wl.triangle_soup(1 << 12, seed=7)
    rays = _plane_rays(wl, desc, 1 << 13, seed=8)
    _check_scene(ar, orc, desc, rays, max_tie_frac=0.0)
    desc = wl.cornell_box()
    _check_scene(ar, orc, desc, _plane_rays(wl, desc, 1 << 12, seed=9), max_tie_frac=0.0)


def test_duplicate_geometry_ties(ar, orc, wl):
    """Every hit is a tie: each triangle appears twice in its mesh and the
    mesh is instanced twice with the same transform.  The reference keeps the
    first one its DFS meets (lowest instance record, then lowest leaf
    position); near-first order must pick the same one although its visiting
    order depends on the wave a ray shares."""
    base = wl.triangle_soup(1 << 11, seed=12)
    m = base.meshes[0]
    idx = m.indices.reshape(-1, 3)
    dup = wl.MeshDesc(positions=m.positions, indices=np.concatenate([idx, idx[::-1]]).astype(np.uint32),
                      materials=m.materials, slots=m.slots, normals=None, transform=m.transform)
    desc = wl.SceneDesc(materials=base.materials, meshes=[dup, dup], lights=base.lights, camera=base.camera)
    rays = wl.soup_rays(1 << 15, seed=13, image=128)
    _check_scene(ar, orc, desc, rays, max_tie_frac=0.0)


def test_empty_batch_and_device_pointers(ar, wl):
    import torch
    desc = wl.triangle_soup(4096, seed=51)
    sc = ar.Scene.from_desc(desc)
    t, p, i = sc.intersect(np.zeros((0, 8), np.float32))
    assert t.shape == (0,)
    rays = wl.soup_rays(1 << 12, seed=52, image=64)
    th, ph, ih = sc.intersect(rays)
    dr = torch.from_numpy(rays).cuda()
    td, pd, idd = sc.intersect(dr)
    torch.cuda.synchronize()
    assert np.array_equal(td.cpu().numpy().view(np.uint32), th.view(np.uint32))
    assert np.array_equal(pd.cpu().numpy().view(np.uint32), ph)
    occ = sc.occluded(dr)
    torch.cuda.synchronize()
    assert np.array_equal(occ.cpu().numpy().astype(bool), sc.occluded(rays))


def test_room_launch_variants_same_answer(ar, wl):
    """Multi-instance scene: the traversal order of a ray does not depend on
    the launch (wave composition, refill threshold, how many lanes may still
    be searching when a wave switches to its leaf phase), so every variant
    returns the same hits, equal-t ties across instances included."""
    desc = wl.procedural_room(n_boxes=30)
    rng = np.random.default_rng(71)
    n = 1 << 16
    o = rng.uniform((0.1, 0.1, 0.1), (7.9, 3.9, 7.9), (n, 3)).astype(np.float32)
    d = wl.normalize_f32(rng.standard_normal((n, 3)).astype(np.float32))
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 3], rays[:, 4:7], rays[:, 7] = o, 1e-4, d, wl.F32_MAX
    sc = ar.Scene.from_desc(desc)
    ref_t, ref_p, ref_i = sc.intersect(rays)
    for cfg in [(64, 8, 4, 1, 0, 3), (256, 8, 8, 64, 63, 4), (128, 16, 16, 8, 8, 4), (128, 8, 8, 24, 32, 3)]:
        sc.set_trace_config(*cfg)
        t, p, i = sc.intersect(rays)
        assert np.array_equal(t.view(np.uint32), ref_t.view(np.uint32)), cfg
        assert np.array_equal(p, ref_p) and np.array_equal(i, ref_i), cfg


def test_every_launch_variant_same_answer(ar, wl):
    """All (workgroup, LDS stack, blocks/CU) variants, including stacks that
    spill to HBM, give identical hits; the autotuner picks one of them."""
    import torch
    desc = wl.triangle_soup(1 << 15, seed=61)
    rays = wl.soup_rays(1 << 15, seed=62, image=128)
    sc = ar.Scene.from_desc(desc)
    ref_t, ref_p, _ = sc.intersect(rays)
    ref_o = sc.occluded(rays)
    for (b, s) in [(64, 8), (64, 16), (128, 8), (128, 16), (256, 8), (256, 12)]:
        for pc, rf, sm, ps in ((4, 1, 0, 3), (16, 64, 63, 4), (8, 24, 8, 4)):
            sc.set_trace_config(b, s, pc, rf, sm, ps)
            t, p, _ = sc.intersect(rays)
            assert np.array_equal(t.view(np.uint32), ref_t.view(np.uint32)) and np.array_equal(p, ref_p), \
                (b, s, pc, rf, sm, ps)
            assert np.array_equal(sc.occluded(rays), ref_o)
    bal = sc.trace_config()["balance"]
    cfg = sc.autotune(torch.from_numpy(rays).cuda())
    assert cfg["ms"] > 0 and cfg["workgroup_size"] in (64, 128, 256)
    assert cfg["balance"] == bal and cfg["anyhit"]["balance"] == bal  # the ray sample does not tune it
    assert cfg["refill_min"] in (8, 24, 48) and cfg["park_slots"] in (3, 4)
    # any hit keeps its own (separately tuned) launch configuration
    assert cfg["anyhit"]["ms"] > 0 and cfg["anyhit"]["workgroup_size"] in (64, 128, 256)
    closest_cfg = sc.trace_config()
    sc.set_anyhit_config(64, 16, 4, 8, 0)
    assert sc.anyhit_config()["lds_stack"] == 16 and sc.trace_config() == closest_cfg
    assert np.array_equal(sc.occluded(rays), ref_o)
    t, p, _ = sc.intersect(rays)
    assert np.array_equal(t.view(np.uint32), ref_t.view(np.uint32)) and np.array_equal(p, ref_p)
    with pytest.raises(ar.ArError):
        sc.set_anyhit_config(64, 8, 4, 24, 64)
    with pytest.raises(ar.ArError):
        sc.set_trace_config(96, 8, 4)
    with pytest.raises(ar.ArError):
        sc.set_trace_config(64, 8, 4, 65)
    with pytest.raises(ar.ArError):
        sc.set_trace_config(64, 8, 4, 24, 64)
    for ps in (0, 2, 5):
        with pytest.raises(ar.ArError):
            sc.set_trace_config(64, 8, 4, 24, 8, ps)
    with pytest.raises(ar.ArError):
        sc.set_trace_config(64, 8, 4, 24, 8, 3, 2)


def test_balanced_launches_same_answer(ar, wl):
    """Launches with enough chunks of 64 rays per wave for the shared chunk
    counter (balance 1: 2^20 rays over 1024 or 2048 waves, 8-16 chunks per
    wave) return the hits of static shares (balance 0) bit for bit, for
    closest and any hit, and so do the launches after them: each launch's
    last wave leaves the counters at zero for the next one."""
    desc = wl.triangle_soup(1 << 16, seed=63)
    rays = wl.soup_rays(1 << 20, seed=64, image=1024)
    sc = ar.Scene.from_desc(desc)
    assert sc.trace_config()["balance"] == 1  # 2^16 triangles: on by default
    small = ar.Scene.from_desc(wl.cornell_box())  # 36 triangles: off by default
    assert small.trace_config()["balance"] == 0 and small.anyhit_config()["balance"] == 0
    sc.set_trace_config(64, 8, 4, 8, 8, 3, 0)
    ref_t, ref_p, ref_i = sc.intersect(rays)
    ref_o = sc.occluded(rays)
    assert (ref_p != 0xFFFFFFFF).mean() > 0.2 and ref_o.mean() > 0.2
    for cfg in [(64, 8, 4, 8, 8, 3, 1), (128, 16, 4, 8, 16, 4, 1), (64, 16, 4, 1, 0, 4, 1)]:
        sc.set_trace_config(*cfg)
        for rep in range(3):
            t, p, i = sc.intersect(rays)
            assert np.array_equal(t.view(np.uint32), ref_t.view(np.uint32)), (cfg, rep)
            assert np.array_equal(p, ref_p) and np.array_equal(i, ref_i), (cfg, rep)
            assert np.array_equal(sc.occluded(rays), ref_o), (cfg, rep)


@pytest.mark.slow
def test_full_size_soup(ar, orc, wl):
    """1M triangles, 2^24 rays: every ray agrees between near-first and
    reference order on the GPU (size-independent property), and a 2^16 ray
    sample of both halves matches the oracle bit for bit."""
    import torch
    desc = wl.triangle_soup(1 << 20, seed=0)
    rays = wl.soup_rays(1 << 24, seed=1)
    sc = ar.Scene.from_desc(desc)
    dr = torch.from_numpy(rays).cuda()
    tn, pn, _ = sc.intersect(dr, ref_order=False)
    tr, pr, _ = sc.intersect(dr, ref_order=True)
    torch.cuda.synchronize()
    same_t = (tn.view(torch.int32) == tr.view(torch.int32))
    same_p = (pn == pr)
    assert bool(same_t.all())
    n_swap = int((~same_p).sum())
    assert n_swap <= 16, n_swap
    idx = np.concatenate([np.arange(0, 1 << 15), (1 << 23) + np.arange(0, 1 << 15)])
    osc = orc.OracleScene(desc)
    ot, op, oi, _, _ = osc.trace(rays[idx])
    tr_h = tr.cpu().numpy()[idx]
    pr_h = pr.cpu().numpy().view(np.uint32)[idx]
    assert np.array_equal(tr_h.view(np.uint32), ot.view(np.uint32))
    assert np.array_equal(pr_h, op)
