"""The film output stage's oracle, pinned without a GPU (the library's kernels
are compared with it in test_gpu_film.py):
* to_rgb_image (film.rs:283-303) against what the mathematics fixes: 8-bit
  levels k/255 come back as k; gray pixels L = k / (255 - k) under Reinhard
  (L / (1 + L) = k / 255) come back as k; the extended white point maps to
  255; extended Reinhard with an unbounded white point is plain Reinhard;
  gamma 2.2 inverts (k/255)^2.2; Reinhard keeps channel ratios; clamping,
  NaN and infinities follow Rust's min/max and saturating casts.
* f32 -> f16 (half::f16::from_f32) against numpy's float16 conversion (IEEE
  round to nearest even) on normals, subnormals, ties, overflow and
  underflow, plus NaN handling.
* the EXR block bytes against the format: an independent reader
  (tests/_exr.py) inverts predictor and split and must return numpy's
  float16 of the film in B, G, R order."""
import os
import sys
import zlib

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import _exr  # noqa: E402


def _gray(v):
    v = np.asarray(v, np.float32)
    return np.repeat(v[:, None], 3, axis=1).reshape(1, -1, 3)


def test_rgb8_levels_round_trip(orc):
    k = np.arange(256)
    out = orc.film_rgb8(_gray(k / 255.0), tonemapper=0, gamma=1.0)
    assert np.array_equal(out[0, :, 0], k) and np.array_equal(out[0, :, 1], k)
    # out of range, NaN and infinities: clamp first (NaN -> 0), then quantise
    special = _gray([-1.0, -0.0, 2.0, np.inf, -np.inf, np.nan])
    assert out.dtype == np.uint8
    assert orc.film_rgb8(special, tonemapper=0, gamma=1.0)[0, :, 0].tolist() == [0, 0, 255, 255, 0, 0]


def test_rgb8_reinhard_gray_levels(orc):
    k = np.arange(255)
    L = k / (255.0 - k)  # Reinhard maps L to L / (1 + L) = k / 255
    out = orc.film_rgb8(_gray(L), tonemapper=1, gamma=1.0)
    assert np.array_equal(out[0, :, 2], k)
    # scale multiplies first: L * 4 with scale 1/4 is the same film
    assert np.array_equal(orc.film_rgb8(_gray(L * 4), tonemapper=1, gamma=1.0, scale=0.25), out)


def test_rgb8_extended_white_point(orc):
    for W in (1.0, 4.0, 16.0):
        assert orc.film_rgb8(_gray([W]), tonemapper=2, white_point=W, gamma=1.0)[0, 0].tolist() == [255] * 3
    rng = np.random.default_rng(5)
    film = rng.lognormal(0.0, 1.5, (32, 32, 3)).astype(np.float32)
    a = orc.film_rgb8(film, tonemapper=1, gamma=2.2)
    b = orc.film_rgb8(film, tonemapper=2, white_point=np.float32(1e30), gamma=2.2)  # lum / lw2 == 0
    assert np.array_equal(a, b)


def test_rgb8_gamma_inverts_power(orc):
    k = np.arange(256)
    x = (k / 255.0) ** 2.2
    out = orc.film_rgb8(_gray(x), tonemapper=0, gamma=2.2)
    assert np.array_equal(out[0, :, 1], k)


def test_rgb8_reinhard_keeps_hue(orc):
    rng = np.random.default_rng(7)
    film = rng.uniform(0.05, 3.0, (1, 4096, 3)).astype(np.float32)
    out = orc.film_rgb8(film, tonemapper=1, gamma=1.0).astype(np.float64)
    lum = film @ np.array([0.2126, 0.7152, 0.0722], np.float32)
    expect = film / (1.0 + lum[..., None]) * 255.0  # every channel scaled by 1 / (1 + lum)
    ok = (expect < 254.0).all(-1)  # no channel clipped
    assert ok.sum() > 1000
    assert np.abs(out[ok] - expect[ok]).max() <= 0.5 + 1e-3


def test_f16_matches_numpy(orc):
    rng = np.random.default_rng(11)
    parts = [
        (rng.standard_normal(200000) * 10.0 ** rng.integers(-9, 6, 200000)).astype(np.float32),
        np.float32(2.0) ** np.arange(-30, 17, dtype=np.float32),  # powers of two across the range
        np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e6, -1e6, np.inf, -np.inf, 5.96e-8, 2.98e-8,
                  2.9802322e-8, 2.98023259e-8, 6.1e-5, 6.1035156e-5], np.float32),
    ]
    # exact ties between neighbouring halves (normal and subnormal)
    h = np.arange(0, 0x7BFF, 7, dtype=np.uint16).view(np.float16).astype(np.float32)
    h2 = (np.arange(1, 0x7C00, 7, dtype=np.uint16)).view(np.float16).astype(np.float32)
    mid = ((h.astype(np.float64) + h2.astype(np.float64)) / 2).astype(np.float32)
    x = np.concatenate(parts + [mid, -mid]).astype(np.float32)
    with np.errstate(over="ignore"):
        expect = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(orc.f32_to_f16(x), expect)
    nan = np.array([np.nan, -np.nan], np.float32)
    nan = np.concatenate([nan, np.array([0x7F800001, 0xFFC12345], np.uint32).view(np.float32)])
    bits = orc.f32_to_f16(nan)
    assert ((bits & 0x7C00) == 0x7C00).all() and ((bits & 0x0200) != 0).all()  # quiet NaNs
    assert ((bits >> 15) == (nan.view(np.uint32) >> 31)).all()  # sign kept


@pytest.mark.parametrize("h,w", [(1, 1), (5, 17), (16, 3), (33, 64), (40, 129)])
def test_exr_blocks_decode_to_float16(orc, h, w):
    rng = np.random.default_rng(h * 1000 + w)
    film = (rng.lognormal(-1.0, 2.0, (h, w, 3)) * rng.choice([-1, 1], (h, w, 3))).astype(np.float32)
    film[0, 0] = (0.0, 1e-9, 2e5)  # zero, underflow to 0, overflow to inf (after the scale)
    scale = np.float32(0.5)
    blocks = orc.exr_blocks(film, scale=scale)
    # each block through zlib and the reader's inverse: B, G, R planes of f16(film * scale)
    expect = (film * scale).astype(np.float16)
    for y0 in range(0, h, 16):
        lines = min(16, h - y0)
        n = lines * w * 6
        chunk = blocks[y0 * w * 6:y0 * w * 6 + n].tobytes()
        raw = _exr.unpredict(zlib.decompress(zlib.compress(chunk)))
        a = np.frombuffer(raw, "<u2").reshape(lines, 3, w)
        for k, c in enumerate((2, 1, 0)):  # B, G, R
            assert np.array_equal(a[:, k], expect[y0:y0 + lines, :, c].view(np.uint16)), (y0, k)
