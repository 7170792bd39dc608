"""A minimal OpenEXR reader for the film-output tests, written from the
OpenEXR 2.0 file layout (magic, version, attribute list, scan-line offset
table, chunks of (y, packed size, data)) and the ZIP compression's
preprocessing (even/odd byte split, delta predictor with +128 bias), so the
tests check the library's files against the format itself.  Handles what
the writers here produce: one scan-line part, HALF channels, NO or ZIP
compression."""
import struct
import zlib

import numpy as np


def parse(data: bytes):
    """Returns (header dict, offsets list, position after the offset table)."""
    magic, version = struct.unpack_from("<ii", data, 0)
    assert magic == 20000630, magic
    assert version & 0xFF == 2 and (version & 0x1E00) == 0, hex(version)  # single-part scan lines
    pos, hdr = 8, {}
    while data[pos] != 0:
        end = data.index(b"\0", pos)
        name = data[pos:end].decode()
        pos = end + 1
        end = data.index(b"\0", pos)
        typ = data[pos:end].decode()
        pos = end + 1
        (size,) = struct.unpack_from("<i", data, pos)
        pos += 4
        hdr[name] = (typ, data[pos:pos + size])
        pos += size
    pos += 1
    y0, y1 = struct.unpack_from("<iiii", hdr["dataWindow"][1])[1::2]
    nlines = y1 - y0 + 1
    comp = hdr["compression"][1][0]
    per = {0: 1, 2: 1, 3: 16}[comp]
    nchunks = (nlines + per - 1) // per
    offsets = list(struct.unpack_from(f"<{nchunks}Q", data, pos))
    return hdr, offsets, pos + 8 * nchunks


def channels(hdr):
    typ, raw = hdr["channels"]
    assert typ == "chlist"
    out, pos = [], 0
    while raw[pos] != 0:
        end = raw.index(b"\0", pos)
        name = raw[pos:end].decode()
        ptype, _plin, xs, ys = struct.unpack_from("<iiii", raw, end + 1)
        out.append((name, ptype, xs, ys))
        pos = end + 1 + 16
    return out


def unpredict(buf: bytes) -> bytes:
    """Inverse of the ZIP preprocessing: undo the delta predictor, then
    re-interleave the even-first split."""
    t = np.frombuffer(buf, np.uint8).astype(np.int64)
    t = np.cumsum(np.concatenate([t[:1], t[1:] - 128])) & 0xFF
    n = t.size
    even = (n + 1) // 2
    raw = np.empty(n, np.uint8)
    raw[0::2] = t[:even]
    raw[1::2] = t[even:]
    return raw.tobytes()


def read_chunks(data: bytes):
    """[(y, packed bytes)] in file order, checked against the offset table."""
    hdr, offsets, pos = parse(data)
    out = []
    for off in offsets:
        assert off >= pos, (off, pos)
        y, size = struct.unpack_from("<ii", data, off)
        out.append((y, data[off + 8:off + 8 + size]))
    return hdr, out


def read_half_image(data: bytes):
    """Returns ({channel name: (h, w) float16 array}, header)."""
    hdr, chunks = read_chunks(data)
    x0, y0, x1, y1 = struct.unpack_from("<iiii", hdr["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    chans = channels(hdr)
    assert all(c[1] == 1 and c[2] == 1 and c[3] == 1 for c in chans)  # HALF, no subsampling
    comp = hdr["compression"][1][0]
    per = {0: 1, 3: 16}[comp]
    planes = {c[0]: np.zeros((h, w), np.uint16) for c in chans}
    for y, packed in chunks:
        lines = min(per, h - (y - y0))
        n = lines * w * 2 * len(chans)
        raw = packed if len(packed) == n else unpredict(zlib.decompress(packed))
        assert len(raw) == n
        a = np.frombuffer(raw, "<u2").reshape(lines, len(chans), w)
        for k, c in enumerate(chans):
            planes[c[0]][y - y0:y - y0 + lines] = a[:, k]
    return {k: v.view(np.float16) for k, v in planes.items()}, hdr


def predicted_blocks(data: bytes):
    """The pre-deflate (split and predicted) bytes of every chunk, back to
    back, for comparison with the oracle's exr_blocks."""
    hdr, chunks = read_chunks(data)
    x0, _, x1, _ = struct.unpack_from("<iiii", hdr["dataWindow"][1])
    w = x1 - x0 + 1
    parts = []
    for _y, packed in chunks:
        try:
            parts.append(zlib.decompress(packed))
        except zlib.error:
            # stored raw: redo the preprocessing to compare like with like
            raw = np.frombuffer(packed, np.uint8)
            n = raw.size
            even = (n + 1) // 2
            t = np.concatenate([raw[0::2], raw[1::2]]).astype(np.int64)
            p = t.copy()
            p[1:] = (t[1:] - t[:-1] + 128) & 0xFF
            assert even == raw[0::2].size
            parts.append(p.astype(np.uint8).tobytes())
    return b"".join(parts), w
