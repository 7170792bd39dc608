"""GPU path tracer (ar_render, PathTracer::render) against the oracle.

* schedules: megakernel and wavefront films are bitwise identical, and do
  not depend on the sample chunking (deterministic accumulation);
* shared counter RNG: GPU and oracle trace the same random numbers with the
  same operation order, so the films are bitwise equal (Cornell box and the
  280k-triangle room; conftest.film_mismatch explains the pixel kept in
  reserve);
* independent RNG: per-pixel means agree within 3 sigma (GPU replicates
  against the oracle's own per-pixel variance) up to the binomial quantile
  of 0.27 %; relMSE against a 4096-spp oracle render, averaged over 8 seeds,
  is within 5 % of the oracle's own at the same spp (check 2);
* closed form: direct lighting from a point light over a plane.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def render(ar, scene, cam, w, h, spp, depth=8, rr=3, seed=0, schedule=0, chunk=0):
    pt = ar.PathTracer(ar.PathTracerConfig(spp=spp, max_depth=depth, rr_depth=rr), seed=seed,
                       schedule=ar.Schedule(schedule), sample_chunk=chunk)
    film = pt.render(scene, cam, w, h)
    return film.pixels.copy(), pt.last_stats


@pytest.fixture(scope="module")
def cornell(ar, orc, wl):
    desc = wl.cornell_box()
    cam16 = wl.camera_lookat(desc.camera.lookfrom, desc.camera.lookat, desc.camera.vup, desc.camera.vfov, 1.0)
    return desc, ar.Scene.from_desc(desc), ar.PinholeCamera.from_array(cam16), cam16, orc.OracleScene(desc)


def test_launch_refinement_keeps_the_film(ar, wl):
    """PathTracer.tune_launch times launch candidates on the job and keeps the
    fastest; the film does not depend on which one is used."""
    desc = wl.cornell_box()
    sc = ar.Scene.from_desc(desc)
    cam = ar.PinholeCamera.from_array(
        wl.camera_lookat(desc.camera.lookfrom, desc.camera.lookat, desc.camera.vup, desc.camera.vfov, 1.0))
    before, _ = render(ar, sc, cam, 96, 64, 4)
    pt = ar.PathTracer(ar.PathTracerConfig(spp=4, max_depth=8, rr_depth=3))
    res = pt.tune_launch(sc, cam, 96, 64, spp=2)
    assert len(res["tried"]["closest"]) >= 6 and all(v > 0 for v in res["tried"]["closest"].values())
    assert len(res["tried"]["anyhit"]) >= 6
    c, a = sc.trace_config(), sc.anyhit_config()
    assert list(ar.config_key(c)) == res["closest"] and list(ar.config_key(a)) == res["anyhit"]
    # the winners were also timed with balance switched, and (before that
    # switch) with the other park slot count
    key = lambda c: "x".join(map(str, c))
    for stage, win in (("closest", res["closest"]), ("anyhit", res["anyhit"])):
        tried, win = res["tried"][stage], tuple(win)
        assert key(ar.other_balance(win)) in tried
        assert key(ar.other_park_slots(win)) in tried or key(ar.other_park_slots(ar.other_balance(win))) in tried
    assert pt.config.spp == 4
    after, _ = render(ar, sc, cam, 96, 64, 4)
    assert np.array_equal(before.view(np.uint32), after.view(np.uint32))


@pytest.mark.parametrize("integ", ["pt", "bdpt", "gdpt"])
def test_balanced_launches_keep_the_film(ar, wl, integ):
    """Extension, shadow and connection launches large enough for the shared
    chunk counter (256x256 on the 280k-triangle room with 1024-wave launches:
    8 or more chunks of 64 rays per wave) give the film of static shares bit
    for bit, render after render (each launch leaves the counters at zero)."""
    desc = wl.procedural_room()
    sc = ar.Scene.from_desc(desc)
    w = h = 256
    cd = desc.camera
    cam = ar.PinholeCamera.from_array(wl.camera_lookat(cd.lookfrom, cd.lookat, cd.vup, cd.vfov, w / h))

    def make():
        if integ == "pt":
            return ar.PathTracer(ar.PathTracerConfig(spp=8, max_depth=8, rr_depth=3), seed=5)
        if integ == "bdpt":
            return ar.BidirectionalPathTracer(ar.BdptConfig(spp=8, max_depth=6, rr_depth=3), seed=5)
        return ar.GradientDomainPathTracer(ar.GdptConfig(spp=4, max_depth=8, rr_depth=3, poisson_iterations=20),
                                           seed=5)

    films = {}
    for bal in (0, 1):
        sc.set_trace_config(64, 8, 4, 8, 8, 3, bal)
        films[bal] = [make().render(sc, cam, w, h).pixels.copy() for _ in range(2)]
    ref = films[0][0]
    assert ref.mean() > 0.01
    for bal in (0, 1):
        for f in films[bal]:
            assert np.array_equal(f.view(np.uint32), ref.view(np.uint32)), bal


@pytest.mark.parametrize("integ", ["pt", "bdpt", "gdpt"])
def test_chunk_tuning_keeps_the_film(ar, wl, integ):
    """tune_chunk times queue chunks of the job and keeps one on the scene; the
    chosen chunk is a whole number of samples per pixel within the job, and
    the film rendered with it equals the one-sample-per-chunk film bit for bit."""
    desc = wl.cornell_box()
    sc = ar.Scene.from_desc(desc)
    w, h, spp = 512, 256, 32
    cam = ar.PinholeCamera.from_array(
        wl.camera_lookat(desc.camera.lookfrom, desc.camera.lookat, desc.camera.vup, desc.camera.vfov, w / h))

    def make(chunk):
        if integ == "pt":
            return ar.PathTracer(ar.PathTracerConfig(spp=spp, max_depth=8, rr_depth=3), sample_chunk=chunk)
        if integ == "bdpt":
            return ar.BidirectionalPathTracer(ar.BdptConfig(spp=spp, max_depth=6, rr_depth=3), sample_chunk=chunk)
        return ar.GradientDomainPathTracer(ar.GdptConfig(spp=spp, max_depth=6, rr_depth=3), sample_chunk=chunk)

    per_sample = w * h * (3 if integ == "gdpt" else 1)
    target = make(0).tune_chunk(sc, cam, w, h)
    assert target % per_sample == 0 and per_sample <= target <= spp * per_sample
    tuned = make(0).render(sc, cam, w, h).pixels.copy()
    single = make(1).render(sc, cam, w, h).pixels.copy()
    assert np.array_equal(tuned.view(np.uint32), single.view(np.uint32))
    assert tuned.mean() > 0.01


def test_schedules_and_chunking_bitwise_identical(ar, cornell):
    desc, sc, cam, _, _ = cornell
    a, sa = render(ar, sc, cam, 64, 48, 8, schedule=1)
    b, sb = render(ar, sc, cam, 64, 48, 8, schedule=0)
    c, _ = render(ar, sc, cam, 64, 48, 8, schedule=0, chunk=3)
    d, _ = render(ar, sc, cam, 64, 48, 8, schedule=1, chunk=1)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(a.view(np.uint32), c.view(np.uint32))
    assert np.array_equal(a.view(np.uint32), d.view(np.uint32))
    assert sa.rays == sb.rays and sb.paths == 64 * 48 * 8
    assert a.mean() > 0.01


def test_counter_rng_matches_oracle(ar, cornell):
    from conftest import film_mismatch
    desc, sc, cam, cam16, osc = cornell
    w, h, spp = 64, 64, 16
    g, _ = render(ar, sc, cam, w, h, spp, seed=7)
    o, _, _, _ = osc.render_pt(cam16, w, h, spp, 8, 3, rng_mode=1, seed=7)
    n, rel = film_mismatch(g, o)
    print(f"  pixels not bitwise equal: {n} of {w * h}, max rel diff {rel:.3g}")
    assert n <= 1


def test_independent_rng_three_sigma(ar, cornell):
    from conftest import outlier_bound
    desc, sc, cam, cam16, osc = cornell
    w, h, reps, spp = 48, 48, 32, 16
    gs = np.stack([render(ar, sc, cam, w, h, spp, seed=300 + r)[0] for r in range(reps)])
    gm, gv = gs.mean(0), gs.var(0, ddof=1) / reps
    o, m2, _, _ = osc.render_pt(cam16, w, h, 512, 8, 3, rng_mode=0, want_m2=True)
    ov = np.maximum(m2 - o * o, 0.0) / 512
    z = (gm - o) / np.sqrt(gv + ov + 1e-20)
    n3 = int((np.abs(z) > 3).sum())
    bound = outlier_bound(z.size, dof=reps - 1)  # the GPU variance from reps replicates
    print(f"  |z| > 3: {n3} of {z.size} (bound {bound}), max |z| {np.abs(z).max():.2f}")
    assert n3 <= bound
    zg = (gm - o).sum() / math.sqrt((gv + ov).sum())
    assert abs(zg) < 4.0


def test_relmse_vs_4096spp_reference(ar, cornell):
    """Check (2): relMSE = mean over pixel-channels of (I - R)^2 / (R^2 + 1e-2)
    against a 4096-spp oracle render R; the GPU's at 16 spp, averaged over 8
    seeds, is within 5 % of the oracle's at 16 spp over 8 other seeds."""
    desc, sc, cam, cam16, osc = cornell
    w, h = 64, 64
    ref, _, _, _ = osc.render_pt(cam16, w, h, 4096, 8, 3, rng_mode=1, seed=12345)

    def relmse(x):
        return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))
    rg = [relmse(render(ar, sc, cam, w, h, 16, seed=100 + s)[0]) for s in range(8)]
    ro = [relmse(osc.render_pt(cam16, w, h, 16, 8, 3, rng_mode=1, seed=200 + s)[0]) for s in range(8)]
    ratio = float(np.mean(rg) / np.mean(ro))
    print(f"  relMSE gpu {np.mean(rg):.5g} cpu {np.mean(ro):.5g} ratio {ratio:.4f}")
    assert ratio <= 1.05


def test_edge_cases(ar, cornell):
    desc, sc, cam, _, _ = cornell
    z, _ = render(ar, sc, cam, 7, 5, 3, depth=0)
    assert np.all(z == 0)
    one, st = render(ar, sc, cam, 1, 1, 1, depth=1)
    assert one.shape == (1, 1, 3) and st.paths == 1
    odd_a, _ = render(ar, sc, cam, 37, 23, 5, schedule=0)
    odd_b, _ = render(ar, sc, cam, 37, 23, 5, schedule=1)
    assert np.array_equal(odd_a, odd_b)
    with pytest.raises(ar.ArError):
        render(ar, sc, cam, 0, 4, 1)
    # a chunk of 2^31 paths or more (32-bit path ids) is rejected before any
    # allocation or launch, by ar_render and by the server's submit
    for integ, w, h, spp, chunk in ((0, 2048, 2048, 1024, 1024), (2, 1920, 1080, 700, 700)):
        cfg = ar.GdptConfig(spp=spp) if integ == 2 else ar.PathTracerConfig(spp=spp)
        it = (ar.GradientDomainPathTracer if integ == 2 else ar.PathTracer)(cfg, sample_chunk=chunk)
        with pytest.raises(ar.ArError) as e:
            it.render(sc, cam, w, h)
        assert e.value.code == -1 and "sample_chunk" in str(e.value)
        srv = ar.RenderServer(sc, max_active_jobs=1)
        with pytest.raises(ar.ArError) as e:
            srv.submit(it, cam, w, h)
        assert e.value.code == -1
        srv.close()


def test_point_light_plane_closed_form(ar, wl):
    desc = wl.SceneDesc(light_dist=2)
    alb, I = 0.5, 4.0
    desc.materials.append(wl.MaterialDesc(0, (alb, alb, alb)))
    P = np.array([[-50, 0, -50], [50, 0, -50], [50, 0, 50], [-50, 0, 50]], np.float32)
    desc.meshes.append(wl.MeshDesc(P, np.array([[0, 1, 2], [0, 2, 3]], np.uint32), materials=[0]))
    desc.lights.append(wl.LightDesc(0, (0.0, 1.0, 0.0), (I, I, I)))
    sc = ar.Scene.from_desc(desc)
    cam16 = wl.camera_lookat((0, 3, 0.001), (0, 0, 0), (0, 1, 0), 30.0, 1.0)
    img, _ = render(ar, sc, ar.PinholeCamera.from_array(cam16), 32, 32, 4, depth=1)
    errs = []
    for y in range(0, 32, 5):
        for x in range(0, 32, 5):
            r = wl.camera_rays(cam16, np.array([(x + 0.5) / 32]), np.array([1 - (y + 0.5) / 32]))[0]
            t = -r[1] / r[5]
            p = r[0:3] + r[4:7] * t
            d = np.array([0, 1, 0]) - p
            r2 = float(d @ d)
            expect = alb / math.pi * I * (1.0 / math.sqrt(r2)) / r2
            errs.append(abs(img[y, x, 0] - expect) / expect)
    assert max(errs) < 0.02


def test_room_scene_schedules_and_oracle(ar, orc, wl):
    desc = wl.procedural_room(n_boxes=40, width=96, height=54)
    sc = ar.Scene.from_desc(desc)
    cam16 = wl.camera_lookat(desc.camera.lookfrom, desc.camera.lookat, desc.camera.vup, desc.camera.vfov, 96 / 54)
    cam = ar.PinholeCamera.from_array(cam16)
    a, sa = render(ar, sc, cam, 96, 54, 4, depth=16, schedule=0, seed=5)
    b, sb = render(ar, sc, cam, 96, 54, 4, depth=16, schedule=1, seed=5)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    osc = orc.OracleScene(desc)
    o, _, _, _ = osc.render_pt(cam16, 96, 54, 4, 16, 3, rng_mode=1, seed=5)
    from conftest import film_mismatch
    n, rel = film_mismatch(a, o)
    print(f"  room: pixels not bitwise equal {n} of {96 * 54}, max rel diff {rel:.3g}")
    assert n <= 1
    # larger render: 2M paths, millions of extension and shadow rays through
    # the wave-refill traversal vs one ray per lane; films bitwise equal
    cam2 = ar.PinholeCamera.from_array(wl.camera_lookat(desc.camera.lookfrom, desc.camera.lookat, desc.camera.vup,
                                                        desc.camera.vfov, 480 / 270))
    c, sc_ = render(ar, sc, cam2, 480, 270, 16, depth=16, schedule=0, seed=9)
    d, sd_ = render(ar, sc, cam2, 480, 270, 16, depth=16, schedule=1, seed=9)
    assert sc_.rays == sd_.rays
    assert np.array_equal(c.view(np.uint32), d.view(np.uint32))


def test_room_relmse_vs_4096spp_reference(ar, orc, wl):
    """Check (2) on the room (~280k triangles, mixed materials, three quad
    lights, depth 16, Russian roulette from bounce 3) at 48 x 27.  With the
    shared counter RNG the GPU's films are the oracle's bit for bit, so its
    relMSE against the 4096-spp oracle render equals the oracle's at the same
    spp and seeds.  Independent seeds are compared on the Cornell box only:
    on the room, caustic paths through the dielectrics make one seed's
    relMSE at 16 spp vary by a factor of several, far above the 5 % to
    resolve."""
    desc = wl.procedural_room(width=48, height=27)
    sc = ar.Scene.from_desc(desc)
    cam16 = wl.camera_lookat(desc.camera.lookfrom, desc.camera.lookat, desc.camera.vup, desc.camera.vfov, 48 / 27)
    cam = ar.PinholeCamera.from_array(cam16)
    osc = orc.OracleScene(desc)
    ref, _, _, _ = osc.render_pt(cam16, 48, 27, 4096, 16, 3, rng_mode=1, seed=54321)

    def relmse(x):
        return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))
    rg = [relmse(render(ar, sc, cam, 48, 27, 16, depth=16, seed=100 + s)[0]) for s in range(8)]
    ro = [relmse(osc.render_pt(cam16, 48, 27, 16, 16, 3, rng_mode=1, seed=100 + s)[0]) for s in range(8)]
    ratio = float(np.mean(rg) / np.mean(ro))
    print(f"  room relMSE gpu {np.mean(rg):.5g} cpu {np.mean(ro):.5g} ratio {ratio:.6f}, "
          f"per seed {min(rg):.3g} .. {max(rg):.3g}")
    assert abs(ratio - 1.0) < 1e-6
