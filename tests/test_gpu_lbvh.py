"""The device tree builder (AR_BUILDER_LBVH: Morton order, radix sort, Karras
hierarchy) against the reference's SAH build:
* tree invariants on random, duplicated and degenerate box sets: preorder
  (left child follows its parent), every primitive in exactly one leaf of at
  most four, each leaf box the exact union of its primitives' boxes and each
  internal box the exact union of its children's;
* scenes built with it answer every query like the SAH-built scene: closest
  hit (t bits, primitive, instance) and occlusion on the soup and the room,
  in near-first and reference order (bitwise-equal-t ties between different
  primitives may resolve either way: counted, below 1e-4 of rays);
* a path-traced film is the SAH scene's bit for bit when no tie occurs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check_tree(boxes, ob, om, op):
    n = boxes.shape[0]
    m = om.shape[0]
    assert sorted(op.tolist()) == list(range(n))
    seen = np.zeros(n, np.int32)
    for i in range(m):
        l, r, c = (int(x) for x in om[i])
        if c:
            assert 1 <= c <= 4
            prims = op[l:l + c]
            seen[prims] += 1
            assert np.array_equal(ob[i, :3], boxes[prims, :3].min(0)) and np.array_equal(ob[i, 3:], boxes[prims, 3:].max(0))
        else:
            assert l == i + 1 and i < r < m
            assert np.array_equal(ob[i, :3], np.minimum(ob[l, :3], ob[r, :3]))
            assert np.array_equal(ob[i, 3:], np.maximum(ob[l, 3:], ob[r, 3:]))
    assert np.all(seen == 1)


def test_tree_invariants(ar):
    rng = np.random.default_rng(11)
    cases = []
    for n in (1, 2, 3, 5, 17, 1000, 100000):
        c = rng.uniform(0, 1, (n, 3)).astype(np.float32)
        e = rng.uniform(0.001, 0.05, (n, 3)).astype(np.float32)
        cases.append(np.hstack([c - e, c + e]))
    same = np.tile(np.array([[0.2, 0.2, 0.2, 0.4, 0.4, 0.4]], np.float32), (3000, 1))  # one centroid
    flat = np.hstack([rng.uniform(0, 1, (5000, 2)), np.zeros((5000, 1)), rng.uniform(1, 2, (5000, 2)),
                      np.zeros((5000, 1))]).astype(np.float32)  # a plane: one axis without extent
    cases += [same, flat]
    for b in cases:
        ob, om, op, depth = ar.bvh_build_host(b, device=True)
        _check_tree(b, ob, om, op)
        assert depth >= 1


def _same_hits(a, b):
    ta, pa, ia = (np.asarray(x) for x in a)
    tb, pb, ib = (np.asarray(x) for x in b)
    same = (ta.view(np.uint32) == tb.view(np.uint32)) & (pa == pb) & (ia == ib)
    tie = (~same) & (ta.view(np.uint32) == tb.view(np.uint32))
    assert np.all(same | tie)
    return int(tie.sum())


@pytest.mark.parametrize("which", ["soup", "room"])
def test_lbvh_scene_answers_like_sah(ar, wl, which):
    if which == "soup":
        desc = wl.triangle_soup(1 << 16, seed=21)
        rays = wl.soup_rays(1 << 16, seed=22, image=256)
    else:
        desc = wl.procedural_room(n_boxes=30, width=64, height=36)
        rng = np.random.default_rng(23)
        n = 1 << 16
        rays = np.zeros((n, 8), np.float32)
        rays[:, 0:3] = rng.uniform((0.1, 0.1, 0.1), (7.9, 3.9, 7.9), (n, 3))
        rays[:, 3] = 1e-4
        rays[:, 4:7] = wl.normalize_f32(rng.standard_normal((n, 3)).astype(np.float32))
        rays[:, 7] = wl.F32_MAX
    sah = ar.Scene.from_desc(desc)
    lb = ar.Scene.from_desc(desc, builder="lbvh")
    ties = 0
    for ref in (False, True):
        ties += _same_hits(sah.intersect(rays, ref_order=ref), lb.intersect(rays, ref_order=ref))
        assert np.array_equal(sah.occluded(rays, ref_order=ref), lb.occluded(rays, ref_order=ref))
    assert ties <= 2e-4 * rays.shape[0]


def test_lbvh_film_equals_sah(ar, wl):
    desc = wl.cornell_box()
    cam = ar.PinholeCamera.from_array(wl.camera_lookat(desc.camera.lookfrom, desc.camera.lookat, desc.camera.vup,
                                                       desc.camera.vfov, 1.0))
    films = []
    for builder in ("sah", "lbvh"):
        sc = ar.Scene.from_desc(desc, builder=builder)
        pt = ar.PathTracer(ar.PathTracerConfig(spp=8, max_depth=8, rr_depth=3), seed=3)
        films.append(pt.render(sc, cam, 64, 64).pixels.copy())
    assert np.array_equal(films[0].view(np.uint32), films[1].view(np.uint32))
