"""Render server (ar_server_*, continuous batching over a resident scene).

* every job's film in a concurrent mix of PT / BDPT / GDPT jobs of different
  sizes equals, bit for bit, the same job rendered alone with ar_render;
* a short job submitted while a long one runs is admitted and finishes first;
* queued and running jobs can be cancelled; the server stays usable;
* a job larger than the HBM budget is rejected at submit.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def room(ar, wl):
    desc = wl.procedural_room(n_boxes=20, width=64, height=64)
    return desc, ar.Scene.from_desc(desc)


def _cam(ar, wl, desc, w, h):
    cd = desc.camera
    return ar.PinholeCamera.from_array(wl.camera_lookat(cd.lookfrom, cd.lookat, cd.vup, cd.vfov, w / h))


def _integ(ar, kind, spp, seed):
    if kind == 0:
        return ar.PathTracer(ar.PathTracerConfig(spp=spp, max_depth=8), seed=seed)
    if kind == 1:
        return ar.BidirectionalPathTracer(ar.BdptConfig(spp=spp, max_depth=6), seed=seed)
    return ar.GradientDomainPathTracer(ar.GdptConfig(spp=spp, max_depth=8, poisson_iterations=30), seed=seed)


def test_mix_equals_alone(ar, wl, room):
    desc, sc = room
    specs = [(0, 64, 48, 8, 1), (1, 32, 32, 4, 2), (2, 48, 40, 4, 3), (0, 96, 64, 16, 4), (1, 40, 24, 2, 5),
             (2, 16, 16, 8, 6), (0, 128, 128, 4, 7), (0, 8, 8, 1, 8)]
    srv = ar.RenderServer(sc, max_active_jobs=4)
    ids = []
    for kind, w, h, spp, seed in specs:
        ids.append(srv.submit(_integ(ar, kind, spp, seed), _cam(ar, wl, desc, w, h), w, h))
    mixed = [srv.wait(i) for i in ids]
    srv.close()
    for (kind, w, h, spp, seed), (film, st) in zip(specs, mixed):
        alone = _integ(ar, kind, spp, seed).render(sc, _cam(ar, wl, desc, w, h), w, h).pixels
        assert np.array_equal(film.pixels.view(np.uint32), alone.view(np.uint32)), (kind, w, h, spp)
        assert st["finish"] >= st["start"] >= st["submit"]


def test_short_job_overtakes_long_job(ar, wl, room):
    desc, sc = room
    srv = ar.RenderServer(sc, max_active_jobs=4)
    # 512 x 512 x 256 = 2^26 paths: eight of the server's 2^23-path chunks
    long_id = srv.submit(_integ(ar, 0, 256, 1), _cam(ar, wl, desc, 512, 512), 512, 512)
    time.sleep(0.02)
    short_id = srv.submit(_integ(ar, 0, 1, 2), _cam(ar, wl, desc, 16, 16), 16, 16)
    _, s_short = srv.wait(short_id)
    _, s_long = srv.wait(long_id)
    srv.close()
    assert s_short["start"] < s_long["finish"]
    assert s_short["finish"] < s_long["finish"]


def test_cancel_and_budget(ar, wl, room):
    desc, sc = room
    srv = ar.RenderServer(sc, max_active_jobs=1)
    a = srv.submit(_integ(ar, 0, 512, 1), _cam(ar, wl, desc, 256, 256), 256, 256)
    b = srv.submit(_integ(ar, 0, 4, 1), _cam(ar, wl, desc, 32, 32), 32, 32)
    srv.cancel(b)   # still queued behind a
    srv.cancel(a)   # running
    for j in (a, b):
        with pytest.raises(ar.ArError) as e:
            srv.wait(j)
        assert e.value.code == -7
    c = srv.submit(_integ(ar, 0, 2, 3), _cam(ar, wl, desc, 16, 16), 16, 16)
    film, _ = srv.wait(c)
    assert film.pixels.shape == (16, 16, 3)
    srv.close()
    tiny = ar.RenderServer(sc, hbm_budget_bytes=1 << 20, max_active_jobs=2)
    with pytest.raises(ar.ArError) as e:
        tiny.submit(_integ(ar, 0, 4, 1), _cam(ar, wl, desc, 512, 512), 512, 512)
    assert e.value.code == -3
    tiny.close()
