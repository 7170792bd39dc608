"""ar_trace_count, the instrumented build of the traversal kernel whose
counts give the bench's algorithmic bytes per ray, against the library's
walker run on the host (tests/host/host_traverse.cpp built with counting
hooks).  For rays that hit nothing, the records a walk reads do not depend
on the order it reads them (the cull distance never shrinks), so the GPU's
node, triangle and instance counts must equal the host's exactly.  Checked
on a soup whose spatial-split tree is kept and on the Cornell instances,
closest hit and any hit; and for all rays, closest hit spills exactly as
many stack entries as it reads back."""
import ctypes as C

import numpy as np
import pytest

import test_host_traverse as host

pytestmark = pytest.mark.gpu


def _interior_rays(wl, n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = rng.uniform(0.05, 0.95, (n, 3))
    r[:, 4:7] = wl.normalize_f32(rng.standard_normal((n, 3)).astype(np.float32))
    r[:, 7] = rng.uniform(0.002, 0.2, n)  # short: many miss everything
    return r


@pytest.mark.parametrize("which", ["large_soup", "cornell"])
def test_trace_count_equals_host_walk_on_misses(ar, wl, which):
    import torch
    print(f"trace_count vs host walk: {which}", flush=True)
    if which == "large_soup":
        desc = wl.triangle_soup(1 << 14, seed=51, size_range=(0.02, 0.2))
    else:
        desc = wl.cornell_box()
    rays = _interior_rays(wl, 1 << 15, 52)
    sc = ar.Scene.from_desc(desc)
    _, prim, _ = sc.intersect(rays)
    miss = np.ascontiguousarray(rays[prim == 0xFFFFFFFF])
    assert miss.shape[0] > 1000, miss.shape
    h = host._helper(counts=True)
    s = host._scene(ar, desc, "split")
    try:
        for anyhit in (False, True):
            gpu = sc.trace_counts(torch.from_numpy(miss).cuda(), anyhit=anyhit)
            c = (C.c_ulonglong * 3)()
            h.host_counts(c)
            host._run(h, s, miss, anyhit, True)
            h.host_counts(c)  # both host drivers walk every ray: twice the records
            assert gpu["rays"] == miss.shape[0]
            assert (2 * gpu["node_visits"], 2 * gpu["tri_tests"], 2 * gpu["inst_tests"]) == (c[0], c[1], c[2]), \
                (which, anyhit, gpu, list(c))
            assert gpu["node_visits"] > miss.shape[0]
    finally:
        ar.lib().ar_scene_destroy(s)


def test_closest_hit_reads_back_every_spilled_entry(ar, wl):
    import torch
    sc = ar.Scene.from_desc(wl.triangle_soup(1 << 16, seed=53))
    rays = torch.from_numpy(wl.soup_rays(1 << 18, seed=54, image=512)).cuda()
    sc.set_trace_config(256, 8, 8, 8, 16, 3, 1)  # the 8-entry stack spills
    c = sc.trace_counts(rays)
    assert c["spill_writes"] > 0 and c["spill_writes"] == c["spill_reads"], c
