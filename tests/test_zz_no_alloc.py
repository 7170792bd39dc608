"""ar_trace_noalloc, the allocation-free variant of ar_trace (its traversal
stack spill lives in a fixed per-GPU array instead of memory the library
allocates), against ar_trace on the same rays, bit for bit: two calls in a
row on one scene, the second with fewer rays, then launches large enough
for the walker's shared chunk counters (kept in the same fixed array), a
second scene, and the argument errors (host pointers, which would need
staging)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _trace_both(ar, torch, sc, rays, ref_order):
    lib = ar.lib()
    n = rays.shape[0]
    dr = torch.from_numpy(rays).cuda()
    flags = ar.AR_REF_ORDER if ref_order else 0
    h0 = torch.empty((n, 2), dtype=torch.float32, device="cuda")
    i0 = torch.empty((n,), dtype=torch.int32, device="cuda")
    h1 = torch.full((n, 2), -1.0, dtype=torch.float32, device="cuda")
    i1 = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    ar._check(lib.ar_trace(sc.handle, C.c_void_p(dr.data_ptr()), n, C.c_void_p(h0.data_ptr()),
                           C.c_void_p(i0.data_ptr()), flags, None))
    ar._check(lib.ar_trace_noalloc(sc.handle, C.c_void_p(dr.data_ptr()), n, C.c_void_p(h1.data_ptr()),
                                   C.c_void_p(i1.data_ptr()), flags, None))
    torch.cuda.synchronize()
    a, b = h0.cpu().numpy(), h1.cpu().numpy()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (n, ref_order)
    assert np.array_equal(i0.cpu().numpy(), i1.cpu().numpy()), (n, ref_order)
    return a


def test_noalloc_equals_ar_trace(ar, wl):
    import torch
    print("noalloc: soup 2^16 tris, 2^18 then 2^12 rays", flush=True)
    sc = ar.Scene.from_desc(wl.triangle_soup(1 << 16, seed=41))
    big = wl.soup_rays(1 << 18, seed=42, image=512)
    hits = _trace_both(ar, torch, sc, big, ref_order=False)
    assert (hits[:, 1].view(np.uint32) != 0xFFFFFFFF).sum() > big.shape[0] // 4  # the rays do hit
    small = wl.soup_rays(1 << 12, seed=43, image=64)
    _trace_both(ar, torch, sc, small, ref_order=False)
    _trace_both(ar, torch, sc, small, ref_order=True)
    # 1024-wave launches over 2^18 rays (4 chunks of 64 per wave): the walker
    # counters in the fixed array are used, twice in a row, then a launch too
    # small for them
    print("noalloc: balanced launches, 2^18 rays over 1024 waves, then 2^12", flush=True)
    sc.set_trace_config(64, 8, 4, 8, 8, 3, 1)
    for _ in range(2):
        _trace_both(ar, torch, sc, big, ref_order=False)
    _trace_both(ar, torch, sc, small, ref_order=False)
    print("noalloc: Cornell instances (rotated, scaled)", flush=True)
    desc = wl.cornell_box()
    sc2 = ar.Scene.from_desc(desc)
    cam = wl.camera_lookat(desc.camera.lookfrom, desc.camera.lookat, desc.camera.vup, desc.camera.vfov, 1.0)
    rng = np.random.default_rng(44)
    rays = wl.camera_rays(cam, rng.uniform(0, 1, 1 << 14), rng.uniform(0, 1, 1 << 14)).astype(np.float32)
    _trace_both(ar, torch, sc2, rays, ref_order=False)


def test_noalloc_rejects_host_pointers(ar, wl):
    sc = ar.Scene.from_desc(wl.triangle_soup(1 << 10, seed=45))
    rays = wl.soup_rays(1 << 8, seed=46, image=16)
    hits = np.empty((rays.shape[0], 2), np.float32)
    lib = ar.lib()
    code = lib.ar_trace_noalloc(sc.handle, rays.ctypes.data_as(C.c_void_p), rays.shape[0],
                                hits.ctypes.data_as(C.c_void_p), None, ar.AR_HOST_PTRS, None)
    assert code == -1
    assert b"device pointers only" in lib.ar_last_error()
