"""Render server (ar_server_*, continuous batching over a resident scene).

* every job's film in a concurrent mix of PT / BDPT / GDPT jobs of different
  sizes equals, bit for bit, the same job rendered alone with ar_render:
  with every PT / GDPT job in the shared pool (paths of all jobs in one set
  of queues, refilled every bounce), with the server's default split, and
  with the pool off (every job in chunks of its own);
* a pooled job cancelled while its paths are in flight drains and ends as
  cancelled while the other pooled jobs still equal their solo renders;
  under churn (random submissions and cancellations into a small pool)
  every completed job still equals its solo render;
* the pool serves the job with the least remaining work first, and jobs
  spanning several accumulation windows equal their solo renders;
* a sample of the specified serving mix (workloads.serving_jobs: square
  sides up to 1024, spp log-uniform up to 256, PT / BDPT / GDPT) on the full
  room scene, submitted at Poisson arrivals: every job completes and equals
  its solo render bit for bit, with one lane and with several;
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


POOL_MODES = {"all_pooled": dict(pool_min_paths=1), "default": {}, "pool_off": dict(pool_paths=0xFFFFFFFF),
              "small_pool": dict(pool_paths=4096, pool_min_paths=1)}


@pytest.mark.parametrize("mode", sorted(POOL_MODES))
def test_mix_equals_alone(ar, wl, room, mode):
    desc, sc = room
    specs = [(0, 64, 48, 8, 1), (1, 32, 32, 4, 2), (2, 48, 40, 4, 3), (0, 96, 64, 16, 4), (1, 40, 24, 2, 5),
             (2, 16, 16, 8, 6), (0, 128, 128, 4, 7), (0, 8, 8, 1, 8)]
    print(f"  mode {mode}", flush=True)
    srv = ar.RenderServer(sc, max_active_jobs=4, **POOL_MODES[mode])
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


def test_pool_serves_least_remaining_work_first(ar, wl, room):
    """Inside the pool the job with the fewest paths left to generate takes
    the freed places first: a 2^23-path job submitted while a 2^26-path job
    renders finishes first, and both equal their solo renders."""
    desc, sc = room
    srv = ar.RenderServer(sc, max_active_jobs=4, pool_min_paths=1)
    big = (0, 1024, 64, 31)
    small = (0, 512, 32, 32)
    ids = []
    for kind, side, spp, seed in (big, small):
        ids.append(srv.submit(_integ(ar, kind, spp, seed), _cam(ar, wl, desc, side, side), side, side))
        time.sleep(0.03)
    (fb, sb), (fs, ss) = [srv.wait(i) for i in ids]
    srv.close()
    assert ss["finish"] < sb["finish"]
    for (kind, side, spp, seed), film in ((big, fb), (small, fs)):
        alone = _integ(ar, kind, spp, seed).render(sc, _cam(ar, wl, desc, side, side), side, side).pixels
        assert np.array_equal(film.pixels.view(np.uint32), alone.view(np.uint32)), (side, spp)


def test_pooled_jobs_over_several_windows_equal_alone(ar, wl, room):
    """Pooled jobs larger than one accumulation window (2^24 paths): their
    windows reuse the two ring slots, each accumulated in sample order.
    PT 1024^2 x 40 spp = 2.5 windows; GDPT 512^2 x 50 spp x 3 variants =
    2.4 windows (21 spp per window), with a small PT job arriving between."""
    desc, sc = room
    srv = ar.RenderServer(sc, max_active_jobs=4, pool_min_paths=1)
    specs = [(0, 1024, 40, 21), (2, 512, 50, 22), (0, 64, 3, 23)]
    ids = []
    for kind, side, spp, seed in specs:
        ids.append(srv.submit(_integ(ar, kind, spp, seed), _cam(ar, wl, desc, side, side), side, side))
        time.sleep(0.01)
    served = [srv.wait(i) for i in ids]
    srv.close()
    for (kind, side, spp, seed), (film, st) in zip(specs, served):
        alone = _integ(ar, kind, spp, seed).render(sc, _cam(ar, wl, desc, side, side), side, side).pixels
        assert np.array_equal(film.pixels.view(np.uint32), alone.view(np.uint32)), (kind, side, spp)
        assert st["paths"] == side * side * spp * (3 if kind == 2 else 1)


def test_pool_with_balanced_launches_equals_alone(ar, wl):
    """Pool iterations whose extension and shadow launches use the shared
    chunk counter (up to 2^22 live paths over 1024-wave launches) serve films
    equal to solo renders with static shares, bit for bit."""
    desc = wl.procedural_room(n_boxes=20, width=64, height=64)
    sc = ar.Scene.from_desc(desc)
    specs = [(0, 1024, 8, 31), (2, 256, 8, 32), (0, 96, 4, 33)]
    sc.set_trace_config(64, 8, 4, 8, 8, 3, 0)
    alone = [_integ(ar, kind, spp, seed).render(sc, _cam(ar, wl, desc, side, side), side, side).pixels.copy()
             for kind, side, spp, seed in specs]
    sc.set_trace_config(64, 8, 4, 8, 8, 3, 1)
    srv = ar.RenderServer(sc, max_active_jobs=4, pool_min_paths=1)
    ids = []
    for kind, side, spp, seed in specs:
        ids.append(srv.submit(_integ(ar, kind, spp, seed), _cam(ar, wl, desc, side, side), side, side))
        time.sleep(0.005)
    served = [srv.wait(i) for i in ids]
    srv.close()
    for (kind, side, spp, seed), (film, st), ref in zip(specs, served, alone):
        assert np.array_equal(film.pixels.view(np.uint32), ref.view(np.uint32)), (kind, side, spp)


def test_pooled_cancel_drains(ar, wl, room):
    desc, sc = room
    srv = ar.RenderServer(sc, max_active_jobs=4, pool_min_paths=1)
    keep = [(0, 96, 96, 64, 11), (2, 64, 64, 16, 12)]
    a = srv.submit(_integ(ar, 0, 1024, 10), _cam(ar, wl, desc, 256, 256), 256, 256)
    ids = [srv.submit(_integ(ar, k, spp, seed), _cam(ar, wl, desc, w, h), w, h) for k, w, h, spp, seed in keep]
    time.sleep(0.05)
    srv.cancel(a)
    with pytest.raises(ar.ArError) as e:
        srv.wait(a)
    assert e.value.code == -7
    for (k, w, h, spp, seed), i in zip(keep, ids):
        film, st = srv.wait(i)
        alone = _integ(ar, k, spp, seed).render(sc, _cam(ar, wl, desc, w, h), w, h).pixels
        assert np.array_equal(film.pixels.view(np.uint32), alone.view(np.uint32)), (k, w, h)
        assert st["paths"] == w * h * spp * (3 if k == 2 else 1)
    # the server keeps serving after the cancellation
    c = srv.submit(_integ(ar, 0, 8, 13), _cam(ar, wl, desc, 32, 32), 32, 32)
    film, _ = srv.wait(c)
    alone = _integ(ar, 0, 8, 13).render(sc, _cam(ar, wl, desc, 32, 32), 32, 32).pixels
    assert np.array_equal(film.pixels.view(np.uint32), alone.view(np.uint32))
    srv.close()


def test_churn_with_cancellations(ar, wl, room):
    """24 small jobs of all three integrators into a 2^16-place pool (every
    PT / GDPT job pooled), submitted at random moments, a third of them
    cancelled at random moments: every job ends (done or cancelled), every
    completed film equals its solo render bit for bit, and the server keeps
    serving afterwards."""
    desc, sc = room
    rng = np.random.default_rng(5)
    srv = ar.RenderServer(sc, max_active_jobs=6, pool_paths=1 << 16, pool_min_paths=1)
    specs, ids, cancelled = [], [], set()
    for k in range(24):
        kind = int(rng.integers(0, 3))
        w, h = int(rng.integers(8, 97)), int(rng.integers(8, 97))
        spp = int(rng.integers(1, 9))
        specs.append((kind, w, h, spp, 100 + k))
        ids.append(srv.submit(_integ(ar, kind, spp, 100 + k), _cam(ar, wl, desc, w, h), w, h))
        if rng.random() < 0.35:
            victim = int(rng.integers(0, len(ids)))
            srv.cancel(ids[victim])
            cancelled.add(victim)
        time.sleep(float(rng.exponential(0.004)))
    results = []
    for k, jid in enumerate(ids):
        try:
            results.append(srv.wait(jid))
        except ar.ArError as e:
            assert k in cancelled and e.code == -7, (k, e.code)
            results.append(None)
    c = srv.submit(_integ(ar, 0, 4, 7), _cam(ar, wl, desc, 32, 32), 32, 32)
    film_c, _ = srv.wait(c)
    srv.close()
    done = 0
    for (kind, w, h, spp, seed), res in zip(specs, results):
        if res is None:
            continue
        alone = _integ(ar, kind, spp, seed).render(sc, _cam(ar, wl, desc, w, h), w, h).pixels
        assert np.array_equal(res[0].pixels.view(np.uint32), alone.view(np.uint32)), (kind, w, h, spp)
        done += 1
    assert done >= 12
    alone = _integ(ar, 0, 4, 7).render(sc, _cam(ar, wl, desc, 32, 32), 32, 32).pixels
    assert np.array_equal(film_c.pixels.view(np.uint32), alone.view(np.uint32))


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


def _sample_mix(wl, budget=1.6e8):
    """Jobs of the specified mix, in arrival order, within a work budget (path
    equivalents: PT 1, GDPT 3, BDPT 5 per pixel-sample), keeping every
    integrator and the largest side (1024) represented."""
    jobs = wl.serving_jobs(rate=20.0, duration=60.0, seed=0)
    w = {0: 1, 1: 5, 2: 3}
    cost = lambda j: j.width * j.height * j.spp * w[j.integrator]
    pick, total = [], 0.0
    for j in jobs:
        if total + cost(j) <= budget and len(pick) < 12:
            pick.append(j)
            total += cost(j)
    for need in (lambda j: j.integrator == 1, lambda j: j.integrator == 2, lambda j: j.width == 1024):
        if not any(need(j) for j in pick):
            pick.append(min((j for j in jobs if need(j)), key=cost))
    pick.sort(key=lambda j: j.index)
    assert {j.integrator for j in pick} == {0, 1, 2} and any(j.width == 1024 for j in pick)
    return pick


@pytest.mark.parametrize("lanes", [1, 3])
def test_serving_mix_at_shape_equals_alone(ar, wl, lanes):
    desc = wl.procedural_room()
    sc = ar.Scene.from_desc(desc)
    cd = desc.camera
    jobs = _sample_mix(wl)
    print(f"  {len(jobs)} jobs: " + ", ".join(f"{['pt', 'bdpt', 'gdpt'][j.integrator]} {j.width}^2 x {j.spp}"
                                               for j in jobs), flush=True)

    def integ(j):
        if j.integrator == 0:
            return ar.PathTracer(ar.PathTracerConfig(spp=j.spp, max_depth=16, rr_depth=3), seed=j.seed)
        if j.integrator == 1:
            return ar.BidirectionalPathTracer(ar.BdptConfig(spp=j.spp, max_depth=10, rr_depth=3), seed=j.seed)
        return ar.GradientDomainPathTracer(ar.GdptConfig(spp=j.spp, max_depth=16, poisson_iterations=50),
                                           seed=j.seed)

    def cam(j):
        return ar.PinholeCamera.from_array(wl.camera_lookat(cd.lookfrom, cd.lookat, cd.vup, cd.vfov,
                                                            j.width / j.height))
    srv = ar.RenderServer(sc, max_active_jobs=8, max_concurrent_jobs=lanes)
    rng = np.random.default_rng(lanes)
    ids = []
    for j in jobs:
        time.sleep(float(rng.exponential(0.03)))
        ids.append(srv.submit(integ(j), cam(j), j.width, j.height))
    served = [srv.wait(i) for i in ids]
    srv.close()
    for j, (film, st) in zip(jobs, served):
        alone = integ(j).render(sc, cam(j), j.width, j.height).pixels
        assert np.array_equal(film.pixels.view(np.uint32), alone.view(np.uint32)), (j.index, j.integrator)
        assert st["finish"] >= st["start"] >= st["submit"] and st["paths"] > 0
