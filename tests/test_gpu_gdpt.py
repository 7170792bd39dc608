"""GPU gradient-domain path tracer (GradientDomainPathTracer::render,
gradient_domain_pt.rs:346-415) against the oracle restatement.

* shared counter RNG: primal/dx/dy films bitwise equal to the oracle's, on
  the Cornell box and on the room at the benchmark's GDPT settings (depth
  16, alpha 0.2, 50 CG iterations, both right-hand-side signs);
* schedules give bitwise-identical films;
* the base paths are the path tracer's paths: primal == PT film up to the
  reference's different normalisation (accum * (1/spp) vs accum / spp);
* E[dx] is the finite difference of E[primal], pixel by pixel (independent
  replicates, binomial 3-sigma bound);
* the reconstruction equals the reference SOR (200 sweeps) on the same films
  within 1e-5 relative L2 (check 3).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(ar, orc, wl):
    desc = wl.cornell_box()
    cam16 = wl.camera_lookat(desc.camera.lookfrom, desc.camera.lookat, desc.camera.vup, desc.camera.vfov, 1.0)
    return desc, ar.Scene.from_desc(desc), ar.PinholeCamera.from_array(cam16), cam16, orc.OracleScene(desc)


def gdpt(ar, sc, cam, w, h, spp, seed=0, schedule=0, alpha=0.2, iters=50, sign=1, chunk=0):
    g = ar.GradientDomainPathTracer(ar.GdptConfig(spp=spp, max_depth=8, rr_depth=3, alpha=alpha,
                                                  poisson_iterations=iters, poisson_sign=sign),
                                    seed=seed, schedule=ar.Schedule(schedule), sample_chunk=chunk)
    film = g.render(sc, cam, w, h)
    return film.pixels.copy(), g.aux.copy(), g.last_stats


def test_films_match_oracle(ar, orc, setup):
    desc, sc, cam, cam16, osc = setup
    w, h, spp = 48, 40, 8
    img, aux, st = gdpt(ar, sc, cam, w, h, spp, seed=11)
    op, ox, oy, _ = orc.render_gdpt(osc, cam16, w, h, spp, 8, 3, rng_mode=1, seed=11)
    from conftest import film_mismatch
    for g, o, name in ((aux[0], op, "primal"), (aux[1], ox, "dx"), (aux[2], oy, "dy")):
        n, rel = film_mismatch(g, o)
        print(f"  {name}: pixels not bitwise equal {n}, max rel {rel:.3g}")
        assert n <= 1
    assert st.paths == 3 * w * h * spp


def test_schedules_identical_and_primal_is_pt(ar, setup):
    desc, sc, cam, cam16, _ = setup
    w, h, spp = 40, 32, 6
    a, aa, _ = gdpt(ar, sc, cam, w, h, spp, seed=3, schedule=0)
    b, ab, _ = gdpt(ar, sc, cam, w, h, spp, seed=3, schedule=1, chunk=2)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(aa.view(np.uint32), ab.view(np.uint32))
    pt = ar.PathTracer(ar.PathTracerConfig(spp=spp, max_depth=8, rr_depth=3), seed=3)
    film = pt.render(sc, cam, w, h).pixels
    np.testing.assert_allclose(aa[0], film, rtol=1e-6, atol=1e-7)
    # boundary gradients are zero (offset == base)
    assert np.all(aa[1][:, -1] == 0) and np.all(aa[2][-1, :] == 0)


def test_gradient_expectation(ar, setup):
    """E[dx(x)] = E[I(x+1)] - E[I(x)]: GDPT's dx film against finite
    differences of an independent path-tracer render, per pixel, with
    variances from independent replicates of both."""
    from conftest import outlier_bound
    desc, sc, cam, cam16, _ = setup
    w, h, spp, reps = 32, 32, 64, 32
    dxs, fds = [], []
    for r in range(reps):
        _, aux, _ = gdpt(ar, sc, cam, w, h, spp, seed=500 + r)
        dxs.append(aux[1][:, :-1])
        pt = ar.PathTracer(ar.PathTracerConfig(spp=spp, max_depth=8, rr_depth=3), seed=900 + r)
        img = pt.render(sc, cam, w, h).pixels
        fds.append(img[:, 1:] - img[:, :-1])
    dxs, fds = np.stack(dxs), np.stack(fds)
    z = (dxs.mean(0) - fds.mean(0)) / np.sqrt(dxs.var(0, ddof=1) / reps + fds.var(0, ddof=1) / reps + 1e-20)
    n3 = int((np.abs(z) > 3).sum())
    bound = outlier_bound(z.size, dof=reps - 1)  # variances from reps replicates each
    print(f"  |z| > 3: {n3} of {z.size} (bound {bound}), max |z| {np.abs(z).max():.2f}")
    assert n3 <= bound


def test_reconstruction_equals_reference_sor(ar, orc, setup):
    desc, sc, cam, cam16, _ = setup
    w, h, spp = 64, 64, 16
    for sign in (1, -1):
        img, aux, _ = gdpt(ar, sc, cam, w, h, spp, seed=7, sign=sign)
        sor, _ = orc.poisson_sor(aux[0], aux[1], aux[2], 0.2, iterations=200, omega=1.6, sign=sign)
        rel = np.linalg.norm(img - sor) / np.linalg.norm(sor)
        print(f"  sign {sign}: rel L2 vs SOR {rel:.3g}")
        assert rel < 1e-5


def test_room_gdpt_matches_oracle(ar, orc, wl):
    """The benchmark's GDPT configuration on the room (depth 16, alpha 0.2,
    CG 50 iterations, both signs) at 96 x 54 x 8 spp: films bitwise equal to
    the oracle's; the reconstruction equals the oracle's own pipeline (its
    films, reference SOR with 200 sweeps) within 1e-5 relative L2."""
    from conftest import film_mismatch
    desc = wl.procedural_room(width=96, height=54)
    sc = ar.Scene.from_desc(desc)
    cam16 = wl.camera_lookat(desc.camera.lookfrom, desc.camera.lookat, desc.camera.vup, desc.camera.vfov, 96 / 54)
    cam = ar.PinholeCamera.from_array(cam16)
    osc = orc.OracleScene(desc)
    op, ox, oy, _ = orc.render_gdpt(osc, cam16, 96, 54, 8, 16, 3, rng_mode=1, seed=13)
    for sign in (1, -1):
        g = ar.GradientDomainPathTracer(ar.GdptConfig(spp=8, max_depth=16, rr_depth=3, alpha=0.2,
                                                      poisson_iterations=50, poisson_sign=sign), seed=13)
        img = g.render(sc, cam, 96, 54).pixels
        for k, o in enumerate((op, ox, oy)):
            n, rel = film_mismatch(g.aux[k], o)
            assert n <= 1, (sign, k, n, rel)
        sor, _ = orc.poisson_sor(op, ox, oy, 0.2, iterations=200, omega=1.6, sign=sign)
        rel = np.linalg.norm(img - sor) / np.linalg.norm(sor)
        print(f"  room sign {sign}: rel L2 vs oracle SOR-200 {rel:.3g}, mean {img.mean():.5f}")
        assert rel < 1e-5 and img.mean() > 0
