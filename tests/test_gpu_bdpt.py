"""GPU bidirectional path tracer (BidirectionalPathTracer::render,
bidirectional_path_tracer.rs:1221-1357) against the oracle restatement.

The reference BDPT is reproduced bug-compatibly (its MIS walk counts the
connection vertex twice, and its light tracing leaves out the camera's
cosine, so its image is biased against the path tracer: +21 % with the
Cornell box's point light, +10 % with its quad light, measured with the
oracle; tests/test_oracle_pins.py pins the parts and shows a textbook walk
equal to the path tracer).  These tests therefore check parity with the
oracle, not unbiasedness:
* shared counter RNG: films bitwise equal to the oracle's (light-tracing
  splats summed in the same 2^-24 fixed point on both sides), including the
  t = 0 strategy (the camera sees the quad light);
* both light types, power, uniform and textbook MIS modes, beta = 1 and 2; the room
  at the benchmark's BDPT settings (depth 10, power heuristic beta = 2,
  three quad lights, power light selection);
* the textbook walk (mis_mode 2, no reference counterpart) makes the GPU's
  BDPT image equal its path tracer's, block by block over replicates;
* determinism: identical films across runs and sample chunkings;
* the wavefront schedule (subpaths as bounce loops of extend and shade
  stages; connection enumerated, every shadow ray traced in one any-hit
  stage, gathered in enumeration order) gives the film of the megakernel
  schedule (one lane per path, rays traced as they come) bit for bit, with
  one and with several connection passes per chunk and for any launch.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bdpt(ar, sc, cam, w, h, spp, seed=0, depth=8, mode=0, beta=2.0, chunk=0, schedule=None):
    b = ar.BidirectionalPathTracer(ar.BdptConfig(spp=spp, max_depth=depth, rr_depth=3, mis_mode=ar.MisMode(mode),
                                                 mis_beta=beta), seed=seed, sample_chunk=chunk,
                                   schedule=ar.Schedule.WAVEFRONT if schedule is None else schedule)
    film = b.render(sc, cam, w, h)
    return film.pixels.copy(), b.last_stats


@pytest.mark.parametrize("twin,mode,beta", [(True, 0, 2.0), (False, 0, 2.0), (True, 1, 2.0), (False, 0, 1.0),
                                             (True, 2, 2.0), (False, 2, 2.0)])
def test_bdpt_matches_oracle(ar, orc, wl, twin, mode, beta):
    desc = wl.cornell_box(point_light_twin=twin)
    sc = ar.Scene.from_desc(desc)
    cam16 = wl.camera_lookat(desc.camera.lookfrom, desc.camera.lookat, desc.camera.vup, desc.camera.vfov, 1.0)
    cam = ar.PinholeCamera.from_array(cam16)
    w, h, spp = 40, 40, 8
    g, st = bdpt(ar, sc, cam, w, h, spp, seed=4, mode=mode, beta=beta)
    osc = orc.OracleScene(desc)
    o, _, _ = orc.render_bdpt(osc, cam16, w, h, spp, 8, 3, mode, beta, rng_mode=1, seed=4)
    from conftest import film_mismatch
    n, rel = film_mismatch(g, o)
    print(f"  twin={twin} mode={mode} beta={beta}: pixels not bitwise equal {n} of {w * h}, max rel {rel:.3g}")
    assert n <= 1
    assert st.paths == w * h * spp and st.rays > 0


def test_bdpt_deterministic(ar, wl):
    desc = wl.cornell_box()
    sc = ar.Scene.from_desc(desc)
    cam16 = wl.camera_lookat(desc.camera.lookfrom, desc.camera.lookat, desc.camera.vup, desc.camera.vfov, 1.0)
    cam = ar.PinholeCamera.from_array(cam16)
    a, _ = bdpt(ar, sc, cam, 32, 24, 6, seed=9)
    b, _ = bdpt(ar, sc, cam, 32, 24, 6, seed=9, chunk=2)
    c, _ = bdpt(ar, sc, cam, 32, 24, 6, seed=9)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(a.view(np.uint32), c.view(np.uint32))


def test_bdpt_room_matches_oracle(ar, orc, wl):
    """The benchmark's BDPT configuration (the room: ~280k triangles, three
    quad lights, mixed Lambertian / conductor / dielectric; depth 10, power
    heuristic beta 2) at 96 x 54 x 4 spp, against the oracle."""
    from conftest import film_mismatch
    desc = wl.procedural_room(width=96, height=54)
    sc = ar.Scene.from_desc(desc)
    cam16 = wl.camera_lookat(desc.camera.lookfrom, desc.camera.lookat, desc.camera.vup, desc.camera.vfov, 96 / 54)
    g, st = bdpt(ar, sc, ar.PinholeCamera.from_array(cam16), 96, 54, 4, seed=6, depth=10)
    o, _, _ = orc.render_bdpt(orc.OracleScene(desc), cam16, 96, 54, 4, 10, 3, 0, 2.0, rng_mode=1, seed=6)
    n, rel = film_mismatch(g, o)
    print(f"  room: pixels not bitwise equal {n} of {96 * 54}, max rel {rel:.3g}, mean {g.mean():.5f}")
    assert n <= 1 and g.mean() > 0
    assert st.paths == 96 * 54 * 4


@pytest.mark.parametrize("twin,mode", [(True, 0), (False, 0), (False, 1), (False, 2)])
def test_bdpt_wavefront_connection_bitwise(ar, wl, twin, mode):
    desc = wl.cornell_box(point_light_twin=twin)
    sc = ar.Scene.from_desc(desc)
    cam16 = wl.camera_lookat(desc.camera.lookfrom, desc.camera.lookat, desc.camera.vup, desc.camera.vfov, 1.0)
    cam = ar.PinholeCamera.from_array(cam16)
    # 256 x 256 x 32 at depth 16: 2^21 paths in one chunk, more than one
    # 12 GiB connection pass holds (153 slots of 53 B per path), so two passes
    mega, sm = bdpt(ar, sc, cam, 256, 256, 32, seed=3, depth=16, mode=mode, schedule=ar.Schedule.MEGAKERNEL)
    wave, sw = bdpt(ar, sc, cam, 256, 256, 32, seed=3, depth=16, mode=mode)
    assert np.array_equal(mega.view(np.uint32), wave.view(np.uint32))
    assert sm.shadow_rays == sw.shadow_rays > 0
    chunked, _ = bdpt(ar, sc, cam, 256, 256, 32, seed=3, depth=16, mode=mode, chunk=5)
    assert np.array_equal(chunked.view(np.uint32), wave.view(np.uint32))


@pytest.mark.parametrize("twin", [True, False])
def test_textbook_bdpt_equals_path_tracer(ar, wl, twin):
    """With the textbook walk every path's weights sum to one, so BDPT and the
    path tracer estimate the same image: 4 x 4 block means (channels summed)
    over 16 independent replicates each agree within 4.5 standard errors, and
    the reference's own walk does not (it is biased by +21 % / +10 %)."""
    desc = wl.cornell_box(point_light_twin=twin)
    sc = ar.Scene.from_desc(desc)
    W = 64
    cam = ar.PinholeCamera.from_array(
        wl.camera_lookat(desc.camera.lookfrom, desc.camera.lookat, desc.camera.vup, desc.camera.vfov, 1.0))

    def blocks(img):
        return img.sum(axis=2).reshape(4, W // 4, 4, W // 4).mean(axis=(1, 3)).ravel()

    reps = 16
    pt = np.asarray([blocks(ar.PathTracer(ar.PathTracerConfig(spp=128, max_depth=5, rr_depth=3), seed=500 + r)
                            .render(sc, cam, W, W).pixels) for r in range(reps)], np.float64)
    bd = np.asarray([blocks(bdpt(ar, sc, cam, W, W, 32, seed=700 + r, depth=5, mode=2)[0]) for r in range(reps)],
                    np.float64)
    ref = np.asarray([blocks(bdpt(ar, sc, cam, W, W, 32, seed=900 + r, depth=5, mode=0)[0]) for r in range(4)],
                     np.float64)
    se = np.sqrt(bd.var(axis=0, ddof=1) / reps + pt.var(axis=0, ddof=1) / reps)
    z = (bd.mean(axis=0) - pt.mean(axis=0)) / np.maximum(se, 1e-30)
    print(f"  twin={twin}: textbook / PT {bd.mean() / pt.mean():.4f}, max |z| {np.abs(z).max():.2f}; "
          f"reference walk / PT {ref.mean() / pt.mean():.4f}")
    assert np.all(np.abs(z) < 4.5)
    assert abs(ref.mean() / pt.mean() - 1.0) > 0.05


def test_bdpt_wavefront_connection_launches(ar, wl):
    desc = wl.procedural_room(n_boxes=20, width=96, height=54)
    sc = ar.Scene.from_desc(desc)
    cam16 = wl.camera_lookat(desc.camera.lookfrom, desc.camera.lookat, desc.camera.vup, desc.camera.vfov, 96 / 54)
    cam = ar.PinholeCamera.from_array(cam16)
    ref, _ = bdpt(ar, sc, cam, 96, 54, 4, depth=10, schedule=ar.Schedule.MEGAKERNEL)
    for cfg in [(64, 16, 16, 8, 16, 3), (256, 8, 8, 24, 8, 4), (128, 8, 4, 1, 0, 3)]:
        sc.set_anyhit_config(*cfg)
        g, _ = bdpt(ar, sc, cam, 96, 54, 4, depth=10)
        assert np.array_equal(g.view(np.uint32), ref.view(np.uint32)), cfg
