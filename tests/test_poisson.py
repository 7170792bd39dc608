"""Screened Poisson reconstruction (gradient_domain_pt.rs:449-630).

CPU: the oracle's SOR restatement is pinned by a dense direct solve of the
same 5-point system on small grids (its fixed point), and the textbook sign
by a closed form (exact gradients of an image reconstruct that image).
GPU: the conjugate-gradient solver matches the reference SOR to 1e-5
relative L2 on identical buffers."""
import numpy as np
import pytest


def _system(P, gx, gy, alpha, sign):
    """Dense A and b of the reference system, one channel (tiny grids only)."""
    h, w = P.shape
    n = w * h
    A = np.zeros((n, n))
    b = np.zeros(n)
    for y in range(h):
        for x in range(w):
            i = y * w + x
            nb = [(x - 1, y), (x + 1, y), (x, y - 1), (x, y + 1)]
            nb = [(a, c) for a, c in nb if 0 <= a < w and 0 <= c < h]
            A[i, i] = alpha + len(nb)
            for a, c in nb:
                A[i, c * w + a] -= 1.0
            div = gx[y, x] - (gx[y, x - 1] if x > 0 else 0) + gy[y, x] - (gy[y - 1, x] if y > 0 else 0)
            b[i] = alpha * P[y, x] + sign * div
    return A, b


@pytest.mark.parametrize("sign", [1, -1])
def test_sor_fixed_point_is_dense_solution(orc, sign):
    rng = np.random.default_rng(0)
    h, w = 9, 13
    P = rng.uniform(0.5, 2.0, (h, w, 3)).astype(np.float32)
    gx = rng.normal(0, 0.1, (h, w, 3)).astype(np.float32)
    gy = rng.normal(0, 0.1, (h, w, 3)).astype(np.float32)
    out, _ = orc.poisson_sor(P, gx, gy, 0.2, iterations=2000, omega=1.6, sign=sign)
    for c in range(3):
        A, b = _system(P[..., c].astype(np.float64), gx[..., c], gy[..., c], 0.2, sign)
        x = np.maximum(np.linalg.solve(A, b), 0).reshape(h, w)
        assert np.abs(out[..., c] - x).max() < 2e-5 * max(1.0, np.abs(x).max())


def test_textbook_sign_reconstructs_exact_gradients(orc):
    rng = np.random.default_rng(1)
    h, w = 12, 17
    I = rng.uniform(0.1, 1.0, (h, w, 3)).astype(np.float32)
    gx = np.zeros_like(I)
    gy = np.zeros_like(I)
    gx[:, :-1] = I[:, 1:] - I[:, :-1]   # forward differences, 0 at the border
    gy[:-1, :] = I[1:, :] - I[:-1, :]
    out, _ = orc.poisson_sor(I, gx, gy, 0.2, iterations=500, sign=-1)
    assert np.abs(out - I).max() < 1e-4
    # the reference's sign does not reconstruct the image (documented bug)
    ref, _ = orc.poisson_sor(I, gx, gy, 0.2, iterations=500, sign=1)
    assert np.abs(ref - I).max() > 1e-2


@pytest.mark.gpu
@pytest.mark.parametrize("size,sign", [((720, 1280), 1), ((64, 48), 1), ((33, 17), -1), ((1, 1), 1)])
def test_cg_matches_reference_sor(ar, orc, size, sign):
    h, w = size
    rng = np.random.default_rng(2)
    # smooth image + noise, gradients of a nearby image + noise (GDPT-like)
    yy, xx = np.mgrid[0:h, 0:w] / max(h, w)
    base = (0.5 + 0.4 * np.sin(6 * xx)[..., None] * np.cos(4 * yy)[..., None] * np.array([1, 0.8, 0.6]))
    P = (base + rng.normal(0, 0.05, (h, w, 3))).astype(np.float32)
    tgt = base.astype(np.float32)
    gx = np.zeros_like(tgt)
    gy = np.zeros_like(tgt)
    gx[:, :-1] = tgt[:, 1:] - tgt[:, :-1]
    gy[:-1, :] = tgt[1:, :] - tgt[:-1, :]
    gx += rng.normal(0, 0.01, gx.shape).astype(np.float32)
    gy += rng.normal(0, 0.01, gy.shape).astype(np.float32)
    sor, secs = orc.poisson_sor(P, gx, gy, 0.2, iterations=200, omega=1.6, sign=sign)
    cg, res = ar.screened_poisson_reconstruct(P, gx, gy, 0.2, iterations=50, sign=sign)
    rel = np.linalg.norm(cg - sor) / max(np.linalg.norm(sor), 1e-30)
    print(f"  {w}x{h} sign {sign}: rel L2 {rel:.3g}, CG residual {res:.3g}, SOR {secs:.2f}s")
    assert rel < 1e-5
