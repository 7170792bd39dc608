"""Film output stage on the GPU (ar_film_to_rgb8, ar_film_save_exr) against
the oracle (film.rs:153-180, 283-303, 310-346):
* 8-bit images of random HDR films (heavy-tailed, negative, NaN, infinite
  pixels; a pixel count that is not a multiple of four) for the three tone
  mappers, gamma 1 and 2.2, two scales, device, host and unaligned device
  buffers: bit-exact with gamma 1; with gamma 2.2 the device's and
  the host's powf may differ by an ulp, so a channel may differ by one level
  where x^(1/2.2) * 255 + 0.5 lies within an ulp of an integer (bounded at
  2e-5 of the channels);
* EXR files: the header (channels B, G, R as HALF, ZIP, data window, line
  order), the offset table and chunk positions, the pre-deflate bytes of
  every block bit-exact with the oracle's, and the decoded halves equal to
  numpy's float16 of film * scale;
* argument errors."""
import os
import struct
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import _exr  # noqa: E402

pytestmark = pytest.mark.gpu


def _film(h, w, seed):
    rng = np.random.default_rng(seed)
    f = rng.lognormal(-0.5, 2.0, (h, w, 3)).astype(np.float32)
    f[rng.random((h, w)) < 0.02] *= -1.0
    flat = f.reshape(-1)
    if flat.size >= 64:
        idx = rng.choice(flat.size, 64, replace=False)
        flat[idx[:16]] = np.nan
        flat[idx[16:32]] = np.inf
        flat[idx[32:48]] = 0.0
        flat[idx[48:]] = 1e-30
    return f


@pytest.mark.parametrize("tm", [0, 1, 2])
def test_rgb8_matches_oracle(ar, orc, tm):
    import torch
    f = _film(193, 257, tm)
    tms = {0: ar.ToneMapper.Clamp, 1: ar.ToneMapper.Reinhard, 2: ar.ToneMapper.ReinhardExtended(3.0)}
    for gamma in (1.0, 2.2):
        for scale in (1.0, 1.0 / 64):
            ref = orc.film_rgb8(f, tonemapper=tm, white_point=3.0, gamma=gamma, scale=scale)
            host = ar.film_to_rgb8(f, tms[tm], gamma, scale)
            dev = ar.film_to_rgb8(torch.from_numpy(f).cuda(), tms[tm], gamma, scale).cpu().numpy()
            assert np.array_equal(host, dev)
            # a film 4 bytes past a 16-byte boundary takes the one-pixel-per-thread kernel
            shifted = torch.empty(f.size + 1, dtype=torch.float32, device="cuda")
            shifted[1:] = torch.from_numpy(f.reshape(-1)).cuda()
            odd = ar.film_to_rgb8(shifted[1:].view(f.shape), tms[tm], gamma, scale).cpu().numpy()
            assert np.array_equal(host, odd)
            diff = host.astype(np.int32) - ref.astype(np.int32)
            if gamma == 1.0:
                assert not diff.any(), (tm, scale, int((diff != 0).sum()))
            else:
                assert np.abs(diff).max() <= 1 and (diff != 0).mean() <= 2e-5, (tm, scale, int((diff != 0).sum()))


def test_film_to_rgb_image_method(ar, orc):
    f = _film(64, 48, 3)
    film = ar.Film(48, 64, f)
    img = film.to_rgb_image(ar.ToneMapper.Reinhard, 1.0, 0.5)
    assert img.shape == (64, 48, 3) and img.dtype == np.uint8
    assert np.array_equal(img, orc.film_rgb8(f, tonemapper=1, gamma=1.0, scale=0.5))


@pytest.mark.parametrize("h,w", [(1, 1), (15, 7), (16, 16), (33, 129), (720, 1280)])
def test_exr_matches_oracle(ar, orc, tmp_path, h, w):
    import torch
    f = _film(h, w, h + w)
    scale = np.float32(1.0 / 16)
    expect = (f * scale).astype(np.float16)
    for where in ("host", "device"):
        path = tmp_path / f"{where}.exr"
        src = f if where == "host" else torch.from_numpy(f).cuda()
        ar.film_save_exr(src, path, float(scale))
        data = path.read_bytes()
        hdr, offsets, table_end = _exr.parse(data)
        assert [c[:2] for c in _exr.channels(hdr)] == [("B", 1), ("G", 1), ("R", 1)]
        assert hdr["compression"] == ("compression", bytes([3]))
        assert struct.unpack("<iiii", hdr["dataWindow"][1]) == (0, 0, w - 1, h - 1)
        assert hdr["displayWindow"] == hdr["dataWindow"] and hdr["lineOrder"][1] == bytes([0])
        assert len(offsets) == (h + 15) // 16 and offsets[0] == table_end
        assert all(struct.unpack_from("<i", data, o)[0] == 16 * b for b, o in enumerate(offsets))
        pre, _ = _exr.predicted_blocks(data)
        assert np.array_equal(np.frombuffer(pre, np.uint8), orc.exr_blocks(f, scale=scale)), where
        planes, _ = _exr.read_half_image(data)
        for name, c in (("R", 0), ("G", 1), ("B", 2)):
            assert np.array_equal(planes[name].view(np.uint16), expect[..., c].view(np.uint16)), (where, name)


def test_film_save_exr_method_and_errors(ar, tmp_path):
    f = _film(20, 30, 9)
    ar.Film(30, 20, f).save_exr(tmp_path / "a.exr", 1.0)
    planes, _ = _exr.read_half_image((tmp_path / "a.exr").read_bytes())
    assert np.array_equal(planes["G"].view(np.uint16), f[..., 1].astype(np.float16).view(np.uint16))
    with pytest.raises(ar.ArError):
        ar.film_to_rgb8(f, ar.ToneMapper.Reinhard, 0.0, 1.0)  # gamma must be > 0
    with pytest.raises(ar.ArError):
        ar.film_to_rgb8(f, ar.ToneMapper.ReinhardExtended(0.0), 2.2, 1.0)
    with pytest.raises(ar.ArError):
        ar.film_to_rgb8(f, ar.ToneMapper(7), 2.2, 1.0)
    with pytest.raises(ar.ArError):
        ar.film_save_exr(f, tmp_path / "missing_dir" / "b.exr", 1.0)
