"""Pins of the oracle's path-sampling functions that the reference cannot pin
itself (it ships no tests, fixtures or golden images): every check here is
fixed by the mathematics, not by the oracle's own formulas.

* Light selection by power (lights.rs:171-235): pmf = each light's share of
  the luminance of Light::power() (point: 4 pi I, quad: pi A L), CDF ending at
  exactly 1, sample_index = first CDF entry >= u, frequencies equal to the pmf,
  zero-power lights never picked, all-zero powers -> uniform.
* A quad area light at depth one (the extension completing
  path_tracer.rs:79-113): the MIS, light-sampling-only and BSDF-only
  estimates each equal the closed form rho L F of a Lambertian plane under a
  parallel rectangle (F: view factor from a point to the rectangle).
* Russian roulette (path_tracer.rs:139-145): rr_depth 1 (roulette from the
  first bounce) equals no roulette in expectation, pixel by pixel.
* Multi-bounce transport, path length by path length: the path tracer's
  length-k image equals the unweighted image of every BDPT strategy (s, t)
  with s + t = k (the per-strategy images of
  bidirectional_path_tracer.rs:1175-1200), including light tracing (s = 1)
  and, with the quad light, emitter hits (t = 0).  These are different
  estimators of one integral: the path tracer's bounce loop, the camera and
  light subpath loops, the connections and the splats must all agree.  The
  reference's light tracing leaves out the camera's cosine (its We = 1 / (A
  cos^4) is applied without the |cos theta| of the camera end,
  bidirectional_path_tracer.rs:1104-1108), so its image times cos(theta) of
  each pixel's direction is what equals the path tracer's; uncorrected it is
  measurably brighter, which the test also asserts.
* MIS weights: with the textbook walk (mis_mode 2, every vertex's reverse
  pdf taken once) the weights of a path sum to one, so the BDPT image equals
  the path tracer's; that fixes the vertex pdfs (pdf_fwd, pdf_rev, the
  connection pdfs) the reference's own walk (mode 0) multiplies.

Statistical bounds: per-pixel 3-sigma outlier counts against the binomial
quantile of 0.27 % (p = 1e-4), and whole-image / block z-scores from
independent replicates."""
import math

import numpy as np
import pytest
from scipy import stats


def _outlier_bound(n, alpha=1e-4, p=0.0027):
    """Largest count of |z| > 3 pixels a correct estimator exceeds with
    probability below alpha."""
    return int(stats.binom.ppf(1.0 - alpha, n, p))


def _cornell(wl, twin):
    desc = wl.cornell_box(point_light_twin=twin)
    cd = desc.camera
    return desc, wl.camera_lookat(cd.lookfrom, cd.lookat, cd.vup, cd.vfov, 1.0)


# --- light selection by power ------------------------------------------------
def _light_scene(wl, lights, dist=2):
    desc = wl.SceneDesc(light_dist=dist)
    desc.materials.append(wl.MaterialDesc(0, (0.5, 0.5, 0.5)))
    P = np.array([[0, 0, 0], [1, 0, 0], [1, 0, 1]], np.float32)
    desc.meshes.append(wl.MeshDesc(P, np.array([[0, 1, 2]], np.uint32), materials=[0]))
    desc.lights.extend(lights)
    return desc


def _power_lum(l):
    if l.type == 0:
        p = 4.0 * math.pi * np.asarray(l.intensity, np.float64)
    else:
        area = float(np.linalg.norm(np.cross(np.asarray(l.edge_u, np.float64), np.asarray(l.edge_v, np.float64))))
        p = math.pi * area * np.asarray(l.intensity, np.float64)
    return max(0.2126 * p[0] + 0.7152 * p[1] + 0.0722 * p[2], 0.0)


def test_power_light_distribution(orc, wl):
    L = wl.LightDesc
    sets = [
        [L(0, (0.5, 1, 0.5), (1.0, 1.0, 1.0))],
        [L(0, (0.5, 1, 0.5), (0.0, 2.0, 0.0)), L(1, (0, 2, 0), (5.0, 1.0, 0.5), (0.3, 0, 0), (0, 0, 0.2))],
        [L(0, (0.5, 1, 0.5), (3.0, 0.5, 0.0)), L(0, (0.2, 1, 0.5), (0.0, 0.0, 0.0)),
         L(1, (0, 2, 0), (1.0, 1.0, 4.0), (0.5, 0, 0), (0, 0, 0.5))],
    ]
    grid = (np.arange(100000, dtype=np.float64) + 0.5) / 100000
    for lights in sets:
        sc = orc.OracleScene(_light_scene(wl, lights))
        pmf, cdf = sc.light_dist()
        lum = np.array([_power_lum(l) for l in lights])
        want = lum / lum.sum()
        assert np.allclose(pmf, want, rtol=2e-6, atol=1e-7), (pmf, want)
        assert cdf[-1] == np.float32(1.0) and np.all(np.diff(cdf) >= 0)
        counts = np.zeros(len(lights))
        for u in grid[::10]:
            i, p = sc.light_sample_index(u)
            assert p == pmf[i]
            counts[i] += 1
        assert np.allclose(counts / counts.sum(), want, atol=2e-4)
        assert counts[lum == 0].sum() == 0  # a zero-power light is never picked
        # first CDF entry >= u: u at an entry picks it, just above picks the next
        for i in range(len(lights) - 1):
            if pmf[i] > 0:
                assert sc.light_sample_index(cdf[i])[0] == i
            j = sc.light_sample_index(np.nextafter(cdf[i], np.float32(2)))[0]
            assert j == next(k for k in range(i + 1, len(lights)) if pmf[k] > 0)
        assert sc.light_sample_index(0.0)[0] == int(np.argmax(pmf > 0))
    # every power zero: uniform; the uniform distribution: floor(u n)
    zero = [L(0, (0.5, 1, 0.5), (0.0, 0.0, 0.0)), L(0, (0.2, 1, 0.5), (0.0, 0.0, 0.0))]
    pmf, cdf = orc.OracleScene(_light_scene(wl, zero)).light_dist()
    assert np.array_equal(pmf, np.float32([0.5, 0.5])) and np.array_equal(cdf, np.float32([0.5, 1.0]))
    three = [L(0, (0.5, 1, 0.5), (k + 1.0, 1.0, 1.0)) for k in range(3)]
    uni = orc.OracleScene(_light_scene(wl, three, dist=1))
    for u in (0.0, 0.3333, 0.3334, 0.9999999):
        i, p = uni.light_sample_index(u)
        assert i == min(int(np.float32(u) * np.float32(3)), 2) and p == np.float32(1.0) / np.float32(3)


# --- quad light over a plane: closed form -------------------------------------
def _view_factor_corner(a, b, h):
    """Point to a parallel rectangle a x b at height h with one corner above it."""
    A, B = a / h, b / h
    sa, sb = math.sqrt(1 + A * A), math.sqrt(1 + B * B)
    return (A / sa * math.atan(B / sa) + B / sb * math.atan(A / sb)) / (2 * math.pi)


def test_quad_light_direct_closed_form(orc, wl):
    rho, Le, h = 0.5, 2.0, 1.0
    x0, x1, z0, z1 = -0.5, 0.5, -0.4, 0.6
    desc = wl.SceneDesc(light_dist=2)
    desc.materials.append(wl.MaterialDesc(0, (rho, rho, rho)))
    P = np.array([[-50, 0, -50], [50, 0, -50], [50, 0, 50], [-50, 0, 50]], np.float32)
    desc.meshes.append(wl.MeshDesc(P, np.array([[0, 1, 2], [0, 2, 3]], np.uint32), materials=[0]))
    # normal cross(edge_u, edge_v) = -y: the light faces the floor
    desc.lights.append(wl.LightDesc(1, (x0, h, z0), (Le, Le, Le), (x1 - x0, 0, 0), (0, 0, z1 - z0)))
    sc = orc.OracleScene(desc)
    W = 16
    cam = wl.camera_lookat((0.0, 0.6, 0.1), (0.0, 0.0, 0.1), (0, 0, 1), 60.0, 1.0)
    # expectation per pixel: the closed form averaged over the pixel (8 x 8 grid)
    sub = (np.arange(8) + 0.5) / 8
    expect = np.zeros((W, W))
    for y in range(W):
        for x in range(W):
            us, vs = np.meshgrid((x + sub) / W, 1 - (y + sub) / W)
            r = wl.camera_rays(cam, us.ravel(), vs.ravel()).astype(np.float64)
            t = -r[:, 1] / r[:, 5]
            px, pz = r[:, 0] + r[:, 4] * t, r[:, 2] + r[:, 6] * t
            F = [_view_factor_corner(x1 - a, z1 - b, h) + _view_factor_corner(a - x0, z1 - b, h) +
                 _view_factor_corner(x1 - a, b - z0, h) + _view_factor_corner(a - x0, b - z0, h)
                 for a, b in zip(px, pz)]
            expect[y, x] = rho * Le * float(np.mean(F))
    spp = 512
    for mode in (0, 1, 2):  # MIS, light sampling only, BSDF hits only
        mean, m2, _, _ = sc.render_pt(cam, W, W, spp, max_depth=2, rr_depth=100, seed=11 + mode, want_m2=True,
                                      pt_mode=mode)
        m, v = mean[:, :, 0].astype(np.float64), np.maximum(m2[:, :, 0] - mean[:, :, 0] ** 2, 1e-12)
        z = (m - expect) / np.sqrt(v / spp)
        zg = (m - expect).sum() / math.sqrt((v / spp).sum())
        print(f"  mode {mode}: mean {m.mean():.5f} vs {expect.mean():.5f}, |z|>3 {int((abs(z) > 3).sum())}, "
              f"global z {zg:.2f}")
        assert int((np.abs(z) > 3).sum()) <= _outlier_bound(z.size)
        assert abs(zg) < 4.0


# --- Russian roulette ---------------------------------------------------------
@pytest.mark.parametrize("twin", [True, False])
def test_russian_roulette_unbiased(orc, wl, twin):
    desc, cam = _cornell(wl, twin)
    sc = orc.OracleScene(desc)
    W, spp = 16, 512
    a, a2, _, _ = sc.render_pt(cam, W, W, spp, max_depth=8, rr_depth=1, seed=21, want_m2=True)
    b, b2, _, _ = sc.render_pt(cam, W, W, spp, max_depth=8, rr_depth=100, seed=22, want_m2=True)
    va, vb = np.maximum(a2 - a * a, 1e-12), np.maximum(b2 - b * b, 1e-12)
    z = (a - b) / np.sqrt(va / spp + vb / spp)
    zg = (a - b).sum() / math.sqrt(((va + vb) / spp).sum())
    print(f"  twin={twin}: means {a.mean():.5f} / {b.mean():.5f}, |z|>3 {int((abs(z) > 3).sum())}, global z {zg:.2f}")
    assert int((np.abs(z) > 3).sum()) <= _outlier_bound(z.size)
    assert abs(zg) < 4.0


# --- per path length: PT against every BDPT strategy, unweighted -------------
def _replicates(fn, reps, blocks=2, weight=None):
    """Block means (blocks x blocks quadrants, summed over channels, each pixel
    times `weight`) of `reps` independent renders fn(seed) -> (H, W, 3)."""
    out = []
    for r in range(reps):
        img = fn(1000 + 7 * r).sum(axis=2)
        if weight is not None:
            img = img * weight
        H, W = img.shape
        out.append(img.reshape(blocks, H // blocks, blocks, W // blocks).mean(axis=(1, 3)).ravel())
    return np.asarray(out, np.float64)


def _z(a, b):
    ra, rb = a.shape[0], b.shape[0]
    se = np.sqrt(a.var(axis=0, ddof=1) / ra + b.var(axis=0, ddof=1) / rb)
    return (a.mean(axis=0) - b.mean(axis=0)) / np.maximum(se, 1e-30)


def _pixel_cos(wl, cam, W):
    """cos(theta) between each pixel centre's camera ray and the optical axis."""
    k = np.arange(W * W)
    r = wl.camera_rays(cam, ((k % W) + 0.5) / W, 1 - ((k // W) + 0.5) / W)
    fwd = np.asarray(cam[12:15], np.float64)
    return (r[:, 4:7].astype(np.float64) @ fwd).reshape(W, W)


@pytest.mark.parametrize("twin", [True, False])
def test_strategies_equal_path_tracer_by_length(orc, wl, twin):
    desc, cam = _cornell(wl, twin)
    sc = orc.OracleScene(desc)
    W, spp, reps, md = 16, 64, 12, 5
    cos = _pixel_cos(wl, cam, W)
    for k in (3, 4, 5):
        pt = _replicates(lambda seed: sc.render_pt(cam, W, W, spp, max_depth=md, rr_depth=100, seed=seed,
                                                   only_len=k)[0], reps)
        strategies = [(k - t, t) for t in range(0, k) if k - t >= 1 and not (k - t == 1 and t < 2)]
        if twin:
            strategies = [st for st in strategies if st[1] >= 1]  # a point light cannot be hit (no t = 0)
        for st in strategies:
            render = lambda seed: orc.render_bdpt(sc, cam, W, W, spp, max_depth=md, rr_depth=100, seed=seed,
                                                  strategy=st, unweighted=True)[0]
            b = _replicates(render, reps, weight=cos if st[0] == 1 else None)
            z = _z(b, pt)
            print(f"  twin={twin} k={k} {st}: {b.mean():.5f} vs PT {pt.mean():.5f}, z {np.round(z, 2)}")
            assert np.all(np.abs(z) < 4.5), (k, st, z)
            if st[0] == 1 and k == 3:  # without the cosine: bright by ~1 / cos
                raw = _replicates(render, reps)
                ratio = raw.sum() / b.sum()
                assert abs(ratio - float((1 / cos).mean())) < 0.01 and ratio > 1.02, ratio


# --- MIS: the textbook walk makes BDPT equal PT --------------------------------
@pytest.mark.parametrize("twin", [True, False])
def test_textbook_mis_bdpt_equals_path_tracer(orc, wl, twin):
    desc, cam = _cornell(wl, twin)
    sc = orc.OracleScene(desc)
    W, spp, reps, md = 16, 128, 12, 5
    pt = _replicates(lambda seed: sc.render_pt(cam, W, W, spp, max_depth=md, rr_depth=3, seed=seed)[0], reps)
    bd = _replicates(lambda seed: orc.render_bdpt(sc, cam, W, W, spp, max_depth=md, rr_depth=3, mis_mode=2,
                                                  seed=seed + 1)[0], reps)
    ref = _replicates(lambda seed: orc.render_bdpt(sc, cam, W, W, spp, max_depth=md, rr_depth=3, mis_mode=0,
                                                   seed=seed + 2)[0], 4)
    z = _z(bd, pt)
    print(f"  twin={twin}: textbook / PT {bd.mean() / pt.mean():.4f} (z {np.round(z, 2)}); "
          f"reference walk / PT {ref.mean() / pt.mean():.4f}")
    assert np.all(np.abs(z) < 4.5)
    # with the emitter's own strategies the textbook walk's advantage over
    # light tracing is visible: it must not be the reference walk's estimate
    assert abs(ref.mean() / pt.mean() - 1.0) > 0.05
