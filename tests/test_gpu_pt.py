"""GPU path tracer (ar_render, PathTracer::render) against the oracle.

* schedules: megakernel and wavefront films are bitwise identical, and do
  not depend on the sample chunking (deterministic accumulation);
* shared counter RNG: GPU and oracle trace the same random numbers, so the
  films agree to rounding except on the rare path whose branch flips on a
  last-bit difference of sin/cos: at most 0.5% of pixel-channels off by more
  than 1e-4 relative, total |diff| below 1e-4 of the total film;
* independent RNG: per-pixel means agree within 3 sigma except for a
  binomial fraction of pixels; relMSE against a high-spp oracle render is
  within 5% of the oracle's own at equal spp;
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
    # the winners were also timed with the other park slot count
    for stage, win in (("closest", res["closest"]), ("anyhit", res["anyhit"])):
        assert "x".join(map(str, ar.other_park_slots(tuple(win)))) in res["tried"][stage]
    assert pt.config.spp == 4
    after, _ = render(ar, sc, cam, 96, 64, 4)
    assert np.array_equal(before.view(np.uint32), after.view(np.uint32))


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
    desc, sc, cam, cam16, osc = cornell
    w, h, spp = 64, 64, 16
    g, _ = render(ar, sc, cam, w, h, spp, seed=7)
    o, _, _, _ = osc.render_pt(cam16, w, h, spp, 8, 3, rng_mode=1, seed=7)
    diff = np.abs(g - o)
    tol = 1e-4 * np.maximum(1.0, np.abs(o))
    frac = float((diff > tol).mean())
    l1 = float(diff.sum() / o.sum())
    print(f"  pixel-channels off by > 1e-4 rel: {frac:.4%}, max diff {diff.max():.3g}, l1 rel {l1:.3g}")
    assert frac <= 0.005
    assert l1 < 1e-4
    assert abs(g.mean() - o.mean()) < 1e-4 * o.mean()


def test_independent_rng_three_sigma(ar, cornell):
    desc, sc, cam, cam16, osc = cornell
    w, h, spp = 48, 48, 256
    g, _ = render(ar, sc, cam, w, h, spp, seed=1)
    o, m2, _, _ = osc.render_pt(cam16, w, h, spp, 8, 3, rng_mode=0, want_m2=True)
    var = np.maximum(m2 - o * o, 0.0)
    sigma = np.sqrt(2.0 * var / spp) + 1e-6
    out = np.abs(g - o) > 3.0 * sigma
    frac = float(out.mean())
    print(f"  3-sigma outliers: {frac:.4%}")
    # 0.27% expected for Gaussian errors; heavy-tailed light paths allow more
    assert frac < 0.015


def test_relmse_vs_high_spp_reference(ar, cornell):
    desc, sc, cam, cam16, osc = cornell
    w, h = 32, 32
    ref, _, _, _ = osc.render_pt(cam16, w, h, 2048, 8, 3, rng_mode=1, seed=99)
    def relmse(x):
        return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))
    g, _ = render(ar, sc, cam, w, h, 16, seed=3)
    o, _, _, _ = osc.render_pt(cam16, w, h, 16, 8, 3, rng_mode=0)
    rg, ro = relmse(g), relmse(o)
    print(f"  relMSE gpu {rg:.4g} cpu {ro:.4g} ratio {rg / ro:.3f}")
    assert rg <= 1.25 * ro  # single seeds; the bench averages several


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
    rel = np.abs(a - o) / np.maximum(1.0, np.abs(o))
    print(f"  room: frac(rel>1e-3) {(rel > 1e-3).mean():.4%}  gpu mean {a.mean():.5f} cpu mean {o.mean():.5f}")
    assert float((rel > 1e-3).mean()) < 0.03
    # larger render: 2M paths, millions of extension and shadow rays through
    # the wave-refill traversal vs one ray per lane; films bitwise equal
    cam2 = ar.PinholeCamera.from_array(wl.camera_lookat(desc.camera.lookfrom, desc.camera.lookat, desc.camera.vup,
                                                        desc.camera.vfov, 480 / 270))
    c, sc_ = render(ar, sc, cam2, 480, 270, 16, depth=16, schedule=0, seed=9)
    d, sd_ = render(ar, sc, cam2, 480, 270, 16, depth=16, schedule=1, seed=9)
    assert sc_.rays == sd_.rays
    assert np.array_equal(c.view(np.uint32), d.view(np.uint32))
