"""Image textures and imported scenes on the GPU against the oracle
(texture.rs:65-151, lambertian.rs:35-56, scene.rs:55-64, importer.rs):
* a Cornell box whose floor and wall take a texture through tex_coords that
  wrap (repeat, negative) and whose tall box takes one through the hit's
  barycentrics: PT with the shared counter RNG equals the oracle's film bit
  for bit (as the untextured Cornell test), both schedules bitwise equal;
* BDPT on the same scene equals the oracle's film bit for bit (textured
  albedo at subpath vertices and at connections);
* the OBJ/MTL fixture loaded by the library renders like the oracle scene
  built from what the importer read (meshes, materials, the decoded PNG)."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def _cam(wl, desc, aspect=1.0):
    return wl.camera_lookat(desc.camera.lookfrom, desc.camera.lookat, desc.camera.vup, desc.camera.vfov, aspect)


def _pt(ar, sc, cam, w, h, spp, seed, schedule=0, depth=8):
    pt = ar.PathTracer(ar.PathTracerConfig(spp=spp, max_depth=depth, rr_depth=3), seed=seed,
                       schedule=ar.Schedule(schedule))
    return pt.render(sc, cam, w, h).pixels.copy()


def _close(g, o):
    from conftest import film_mismatch
    n, rel = film_mismatch(g, o)
    print(f"  pixels not bitwise equal: {n} of {g.shape[0] * g.shape[1]}, max rel diff {rel:.3g}")
    assert n <= 1


def test_textured_pt_matches_oracle(ar, orc, wl):
    desc = wl.textured_cornell_box()
    sc = ar.Scene.from_desc(desc)
    cam16 = _cam(wl, desc)
    cam = ar.PinholeCamera.from_array(cam16)
    g = _pt(ar, sc, cam, 64, 64, 16, seed=7)
    o, _, _, _ = orc.OracleScene(desc).render_pt(cam16, 64, 64, 16, 8, 3, rng_mode=1, seed=7)
    _close(g, o)
    m = _pt(ar, sc, cam, 64, 64, 16, seed=7, schedule=1)
    assert np.array_equal(g.view(np.uint32), m.view(np.uint32))
    # the textures matter: the same scene untextured renders differently
    plain = wl.cornell_box(point_light_twin=True)
    p = _pt(ar, ar.Scene.from_desc(plain), cam, 64, 64, 16, seed=7)
    assert np.abs(p - g).mean() > 10 * np.abs(g - o).mean()


def test_textured_bdpt_matches_oracle(ar, orc, wl):
    desc = wl.textured_cornell_box()
    sc = ar.Scene.from_desc(desc)
    cam16 = _cam(wl, desc)
    b = ar.BidirectionalPathTracer(ar.BdptConfig(spp=8, max_depth=8, rr_depth=3), seed=4)
    g = b.render(sc, ar.PinholeCamera.from_array(cam16), 40, 40).pixels.copy()
    o, _, _ = orc.render_bdpt(orc.OracleScene(desc), cam16, 40, 40, 8, 8, 3, 0, 2.0, rng_mode=1, seed=4)
    _close(g, o)


def test_obj_fixture_renders_like_oracle(ar, orc, wl):
    import test_importer as ti
    gold = os.path.join(ti.GOLD, "scene.obj")
    light = ar.PointLight((0.5, 0.5, 2.0), (4.0, 4.0, 4.0))
    sc = ar.Scene([], [light], obj_files=[(gold, None)])
    assert sc.obj_objects == 3
    # the oracle scene from what the importer read
    hs = ti.HostScene(ar)
    try:
        hs.load_obj(gold)
        nmesh, nmat, _ = hs.counts()
        mats = [hs.material(i) for i in range(nmat)]
        meshes = [hs.mesh(i) for i in range(nmesh)]
    finally:
        hs.close()
    tex = ar.load_png(os.path.join(ti.GOLD, "tex", "checker.png"))
    desc = wl.SceneDesc(textures=[tex], light_dist=2, camera=wl.CameraDesc((0.5, 0.5, 2.5), (0.5, 0.5, 0.0)))
    for m in mats:
        desc.materials.append(wl.MaterialDesc(m["type"], tuple(float(x) for x in m["albedo"]), m["roughness"],
                                              m["eta"], m["tex"]))
    for m in meshes:
        desc.meshes.append(wl.MeshDesc(m["pos"], m["idx"], [int(x) for x in m["mats"]], m["slots"], m["nrm"],
                                       None, m["uv"]))
    desc.lights.append(wl.LightDesc(0, (0.5, 0.5, 2.0), (4.0, 4.0, 4.0)))
    cam16 = _cam(wl, desc)
    g = _pt(ar, sc, ar.PinholeCamera.from_array(cam16), 48, 48, 8, seed=2)
    o, _, _, _ = orc.OracleScene(desc).render_pt(cam16, 48, 48, 8, 8, 3, rng_mode=1, seed=2)
    assert g.mean() > 0
    _close(g, o)
