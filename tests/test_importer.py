"""Scene ingestion without a GPU: the library's PNG decoder, its sRGB texel
conversion, the oracle's texture sampling, and the OBJ/MTL importer
(importer.rs:167-241 with tobj's GPU_LOAD_OPTIONS; texture.rs:65-151).

* PNG: decoded pixels equal PIL's (an independent decoder) for 8-bit RGB,
  RGBA, gray, gray + alpha, palette and 1-bit files written by PIL, for an
  Adam7-interlaced file written here from the PNG specification (read by
  PIL too), and for the reference's own texture files when present.
* sRGB -> linear: the texels the library stores equal the oracle's
  srgb_to_linear of every 8-bit level; that function's power is within 1 ulp
  of float64 on the same float32 steps, the whole within 4e-7 relative of
  the IEC 61966-2-1 formula, with its known points (0, 1, the linear segment
  below the 0.04045 knee).
* Texture sampling (oracle): texel centres return texels, halfway points the
  mean of two neighbours, the four-corner point the mean of four, repeat wrap
  (u and u + k sample alike; negative u wraps), a constant image is constant.
* OBJ/MTL: a hand-written fixture covering a polygon (fan triangulation),
  negative indices, v/vt/vn, v/vt and v forms, usemtl and g splits and each
  material rule (textured diffuse, dielectric for d < 0.5, conductor for
  luminance(Ks) > 0.04, plain diffuse) gives the meshes and materials worked
  out by hand below; the reference's fireplace_room.mtl converts by the same
  rules (its textures are not shipped here, so the constant Kd stands in, as
  the reference does when a texture cannot be opened)."""
import ctypes as C
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "obj")
SRC = os.path.join(ROOT, "tests", "host", "host_scene.cpp")
OUT = os.path.join(ROOT, "tests", "host", "_build", "libhost_scene.so")
LIB = os.path.join(ROOT, "agent-ray_amd", "libagentray.so")
REF_TEX = "/root/reference/assets/fireplace_room/textures"

PIL = pytest.importorskip("PIL.Image")


def _helper():
    deps = [SRC, LIB, os.path.join(ROOT, "agent-ray_amd", "csrc", "ar_internal.h")]
    if not os.path.exists(OUT) or os.path.getmtime(OUT) < max(os.path.getmtime(d) for d in deps):
        if not os.path.exists("/opt/rocm/include/hip/hip_runtime.h"):
            pytest.skip("HIP headers not installed")
        os.makedirs(os.path.dirname(OUT), exist_ok=True)
        subprocess.check_call(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-D__HIP_PLATFORM_AMD__",
                               "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"), SRC, "-o", OUT + ".tmp",
                               "-L" + os.path.dirname(LIB), "-lagentray", "-Wl,-rpath," + os.path.dirname(LIB)])
        os.replace(OUT + ".tmp", OUT)
    return C.CDLL(OUT)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class HostScene:
    """An uncommitted library scene and read-back of its host-side contents."""

    def __init__(self, ar):
        self.ar, self.lib, self.h = ar, ar.lib(), _helper()
        self.s = C.c_void_p()
        assert self.lib.ar_scene_create(C.byref(self.s)) == 0

    def load_obj(self, path, xf=None):
        n = C.c_uint32()
        xfa = None if xf is None else np.ascontiguousarray(xf, np.float32)
        rc = self.lib.ar_scene_load_obj(self.s, os.fsencode(path), None if xfa is None else _p(xfa), C.byref(n))
        assert rc == 0, self.lib.ar_last_error()
        return n.value

    def counts(self):
        c = np.zeros(3, np.uint32)
        self.h.host_scene_counts(self.s, _p(c))
        return [int(x) for x in c]

    def mesh(self, i):
        sz = np.zeros(6, np.uint32)
        self.h.host_scene_mesh_sizes(self.s, C.c_uint32(i), _p(sz))
        nv, nt, hn, hu, ns, nm = (int(x) for x in sz)
        pos, nrm, uv = np.zeros((nv, 3), np.float32), np.zeros((nv, 3), np.float32), np.zeros((nv, 2), np.float32)
        idx, slots, mats = np.zeros((nt, 3), np.uint32), np.zeros(ns, np.uint32), np.zeros(nm, np.uint32)
        xf = np.zeros(16, np.float32)
        self.h.host_scene_mesh(self.s, C.c_uint32(i), _p(pos), _p(nrm), _p(uv), _p(idx), _p(slots), _p(mats), _p(xf))
        return dict(pos=pos, nrm=nrm if hn else None, uv=uv if hu else None, idx=idx, slots=slots, mats=mats, xf=xf)

    def material(self, i):
        o = np.zeros(7, np.float32)
        self.h.host_scene_material(self.s, C.c_uint32(i), _p(o))
        return dict(type=int(o[0]), albedo=o[1:4].copy(), roughness=float(o[4]), eta=float(o[5]), tex=int(o[6]))

    def texture(self, i):
        wh = np.zeros(2, np.uint32)
        self.h.host_scene_texture(self.s, C.c_uint32(i), _p(wh), None)
        tx = np.zeros((int(wh[1]), int(wh[0]), 3), np.float32)
        self.h.host_scene_texture(self.s, C.c_uint32(i), _p(wh), _p(tx))
        return tx

    def add_texture(self, rgb):
        rgb = np.ascontiguousarray(rgb, np.uint8)
        out = C.c_uint32()
        assert self.lib.ar_scene_add_texture(self.s, _p(rgb), rgb.shape[1], rgb.shape[0], C.byref(out)) == 0
        return out.value

    def close(self):
        self.lib.ar_scene_destroy(self.s)


# ---- PNG ------------------------------------------------------------------------

def _png_chunk(t, d):
    return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)


def _adam7_png(rgb):
    """An 8-bit RGB PNG with Adam7 interlacing, every row filter 0, written
    from the PNG specification's pass table."""
    h, w = rgb.shape[:2]
    raw = b""
    for x0, y0, dx, dy in ((0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2),
                           (0, 1, 1, 2)):
        sub = rgb[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        for row in sub:
            raw += b"\0" + row.tobytes()
    ihdr = struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 1)
    return b"\x89PNG\r\n\x1a\n" + _png_chunk(b"IHDR", ihdr) + _png_chunk(b"IDAT", zlib.compress(raw)) + \
        _png_chunk(b"IEND", b"")


def test_png_decoder_matches_pil(ar, tmp_path):
    rng = np.random.default_rng(3)
    rgb = rng.integers(0, 256, (37, 53, 3), dtype=np.uint8)
    rgb[:, :20] = rgb[:, :1]  # flat runs so the encoder picks several filter types
    cases = {
        "rgb": PIL.fromarray(rgb, "RGB"),
        "rgba": PIL.fromarray(np.dstack([rgb, rng.integers(0, 256, rgb.shape[:2], dtype=np.uint8)]), "RGBA"),
        "gray": PIL.fromarray(rgb[..., 0], "L"),
        "gray_alpha": PIL.fromarray(np.dstack([rgb[..., 0], rgb[..., 1]]), "LA"),
        "palette": PIL.fromarray(rgb, "RGB").quantize(colors=37),
        "bits1": PIL.fromarray(rgb[..., 0] > 127),
    }
    for name, im in cases.items():
        path = tmp_path / f"{name}.png"
        im.save(path, optimize=False)
        got = ar.load_png(path)
        want = np.asarray(PIL.open(path).convert("RGB"))
        assert np.array_equal(got, want), name
    path = tmp_path / "adam7.png"
    path.write_bytes(_adam7_png(rgb))
    assert PIL.open(path).info.get("interlace") == 1
    assert np.array_equal(ar.load_png(path), np.asarray(PIL.open(path).convert("RGB")))
    assert np.array_equal(ar.load_png(path), rgb)
    with pytest.raises(ar.ArError):
        ar.load_png(tmp_path / "missing.png")
    (tmp_path / "bad.png").write_bytes(b"not a png")
    with pytest.raises(ar.ArError):
        ar.load_png(tmp_path / "bad.png")


@pytest.mark.skipif(not os.path.isdir(REF_TEX), reason="the reference's texture files are not present")
def test_png_decoder_on_reference_textures(ar):
    for f in sorted(os.listdir(REF_TEX)):
        if f.endswith(".png"):
            got = ar.load_png(os.path.join(REF_TEX, f))
            want = np.asarray(PIL.open(os.path.join(REF_TEX, f)).convert("RGB"))
            assert np.array_equal(got, want), f


# ---- sRGB and texture sampling ----------------------------------------------------

def test_srgb_levels(ar, orc):
    hs = HostScene(ar)
    try:
        levels = np.arange(256, dtype=np.uint8)
        hs.add_texture(np.repeat(levels[None, :, None], 3, axis=2))
        got = hs.texture(0)[0, :, 0]
    finally:
        hs.close()
    L = orc.lib()
    ref = np.array([L.orc_srgb_to_linear(np.float32(k) / np.float32(255.0)) for k in range(256)], np.float32)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    # the f32 steps (k / 255, + 0.055, / 1.055) in numpy's IEEE float32, then
    # the power in float64: powf within 1 ulp of it
    x32 = np.arange(256, dtype=np.float32) / np.float32(255.0)
    y32 = (x32 + np.float32(0.055)) / np.float32(1.055)
    lin = np.where(x32 <= np.float32(0.04045), (x32 / np.float32(12.92)).astype(np.float64),
                   y32.astype(np.float64) ** np.float64(np.float32(2.4)))
    ulp = np.spacing(np.abs(lin).astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(ref - lin) <= ulp)
    # and within 4e-7 relative of the IEC 61966-2-1 formula evaluated exactly
    # (the float32 steps before the power, amplified 2.4x by it)
    x = np.arange(256, dtype=np.float64) / 255.0
    exact = np.where(x <= 0.04045, x / 12.92, ((x + 0.055) / 1.055) ** 2.4)
    assert np.all(np.abs(ref - exact) <= 4e-7 * exact)
    assert ref[0] == 0.0 and ref[255] == 1.0
    assert ref[10] == np.float32(np.float32(10 / 255.0) / np.float32(12.92))  # below the knee: linear segment


def _tex_scene(orc, wl, rgb):
    desc = wl.SceneDesc(textures=[rgb])
    return orc.OracleScene(desc)


def test_texture_sampling_closed_forms(orc, wl):
    rng = np.random.default_rng(5)
    rgb = rng.integers(0, 256, (3, 4, 3), dtype=np.uint8)
    sc = _tex_scene(orc, wl, rgb)
    L = orc.lib()
    lin = np.vectorize(lambda v: L.orc_srgb_to_linear(np.float32(v) / np.float32(255.0)), otypes=[np.float32])(rgb)
    h, w = rgb.shape[:2]
    # texel centres (u = (x + 0.5) / w) return the texel
    xs, ys = np.meshgrid(np.arange(w), np.arange(h))
    uv = np.stack([(xs + 0.5) / w, (ys + 0.5) / h], -1).reshape(-1, 2)
    assert np.allclose(sc.texture_sample(0, uv), lin.reshape(-1, 3), rtol=0, atol=1e-7)
    # halfway between two horizontal neighbours: their mean; four corners: mean of four
    uv = np.array([[1.0 / w, 0.5 / h], [2.0 / w, 1.0 / h]], np.float32)
    got = sc.texture_sample(0, uv)
    assert np.allclose(got[0], (lin[0, 0] + lin[0, 1]) / 2, atol=1e-6)
    assert np.allclose(got[1], (lin[0, 1] + lin[0, 2] + lin[1, 1] + lin[1, 2]) / 4, atol=1e-6)
    # repeat wrap: the left edge blends the last column; u + k and u - k sample alike
    got = sc.texture_sample(0, np.array([[0.0, 0.5 / h]], np.float32))
    assert np.allclose(got[0], (lin[0, 0] + lin[0, w - 1]) / 2, atol=1e-6)
    u = rng.uniform(0, 1, (200, 2)).astype(np.float32)
    base = sc.texture_sample(0, u)
    for k in (1.0, 3.0, -2.0):
        assert np.allclose(sc.texture_sample(0, u + np.float32(k)), base, atol=2e-5)
    flat = _tex_scene(orc, wl, np.full((5, 7, 3), 77, np.uint8))
    c = np.float32(L.orc_srgb_to_linear(np.float32(77) / np.float32(255.0)))
    assert np.allclose(flat.texture_sample(0, rng.uniform(-3, 3, (100, 2))), c, rtol=1e-6, atol=0)


# ---- OBJ / MTL ---------------------------------------------------------------------

def test_obj_fixture(ar):
    hs = HostScene(ar)
    try:
        xf = np.eye(4, dtype=np.float32).reshape(16)
        xf[12:15] = (1.0, 2.0, 3.0)  # translation (column-major)
        assert hs.load_obj(os.path.join(GOLD, "scene.obj"), xf) == 3
        nmesh, nmat, ntex = hs.counts()
        assert (nmesh, nmat, ntex) == (3, 4, 1)
        m = [hs.material(i) for i in range(4)]
        # plain: Lambertian Kd
        assert m[0]["type"] == 0 and np.allclose(m[0]["albedo"], (0.1, 0.2, 0.3)) and m[0]["tex"] == -1
        # wood: Lambertian with the map_Kd texture (backslash path), luminance(Ks) = 0.01 <= 0.04
        assert m[1]["type"] == 0 and m[1]["tex"] == 0
        # glass: d 0.2 < 0.5 -> dielectric, eta = Ni, roughness = sqrt(2 / (Ns + 2))
        assert m[2]["type"] == 2 and m[2]["eta"] == np.float32(1.33)
        assert m[2]["roughness"] == np.float32(np.sqrt(np.float32(2.0) / np.float32(102.0)))
        # metal: luminance(Ks) = 0.9 > 0.04 -> conductor F0 = Ks, Ns 0 -> roughness 1
        assert m[3]["type"] == 1 and np.allclose(m[3]["albedo"], 0.9) and m[3]["roughness"] == 1.0
        tex = hs.texture(0)
        assert tex.shape == (2, 4, 3) and tex[0, 0, 0] == 1.0 and tex[0, 0, 1] == 0.0
        quad, tri, other = (hs.mesh(i) for i in range(3))
        for mm in (quad, tri, other):
            assert np.array_equal(mm["mats"], np.arange(4)) and np.array_equal(mm["xf"], xf)
        # quad: v1..v4 with vt and vn, fan (0 1 2) (0 2 3), material wood
        assert np.array_equal(quad["pos"], [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]])
        assert np.array_equal(quad["uv"], [[0, 0], [1, 0], [1, 1], [0, 1]])
        assert np.array_equal(quad["nrm"], [[0, 0, 1]] * 4)
        assert np.array_equal(quad["idx"], [[0, 1, 2], [0, 2, 3]]) and quad["slots"].tolist() == [1]
        # f -5/1 -4/2 -1/4: v1 v2 v5, uvs, no normals; usemtl glass starts a new model
        assert np.array_equal(tri["pos"], [[0, 0, 0], [1, 0, 0], [0, 0, 1]])
        assert np.array_equal(tri["uv"], [[0, 0], [1, 0], [0, 1]]) and tri["nrm"] is None
        assert np.array_equal(tri["idx"], [[0, 1, 2]]) and tri["slots"].tolist() == [2]
        # g other + usemtl metal: v2 v3 v5, positions only
        assert np.array_equal(other["pos"], [[1, 0, 0], [1, 1, 0], [0, 0, 1]])
        assert other["uv"] is None and other["nrm"] is None and other["slots"].tolist() == [3]
    finally:
        hs.close()


def test_obj_without_mtl_and_errors(ar, tmp_path):
    (tmp_path / "a.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    hs = HostScene(ar)
    try:
        assert hs.load_obj(str(tmp_path / "a.obj")) == 1
        assert hs.counts()[1] == 1  # the default mid-grey Lambertian (importer.rs:139-141)
        mat = hs.material(0)
        assert mat["type"] == 0 and np.allclose(mat["albedo"], 0.5)
        (tmp_path / "b.obj").write_text("v 0 0 0\nf 1 2 3\n")
        n = C.c_uint32()
        assert hs.lib.ar_scene_load_obj(hs.s, os.fsencode(str(tmp_path / "b.obj")), None, C.byref(n)) == -1
        assert hs.lib.ar_scene_load_obj(hs.s, os.fsencode(str(tmp_path / "none.obj")), None, C.byref(n)) == -8
        # a directory opens on Linux and fails on the first read: a status, not a C++ exception
        (tmp_path / "sub").mkdir()
        assert hs.lib.ar_scene_load_obj(hs.s, os.fsencode(str(tmp_path / "sub")), None, C.byref(n)) == -8
        (tmp_path / "c.obj").write_text("mtllib sub\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
        assert hs.lib.ar_scene_load_obj(hs.s, os.fsencode(str(tmp_path / "c.obj")), None, C.byref(n)) in (0, -8)
        with pytest.raises(ar.ArError):
            ar.load_png(tmp_path / "sub")
    finally:
        hs.close()


def test_fireplace_mtl_conversion(ar, tmp_path):
    import shutil
    shutil.copy(os.path.join(GOLD, "fireplace_room.mtl"), tmp_path / "fireplace_room.mtl")
    (tmp_path / "room.obj").write_text("mtllib fireplace_room.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    mtl = {}
    cur = None
    for line in open(tmp_path / "fireplace_room.mtl"):
        t = line.split()
        if not t or t[0].startswith("#"):
            continue
        if t[0] == "newmtl":
            cur = mtl.setdefault(" ".join(t[1:]), {})
        elif cur is not None and t[0] in ("Kd", "Ks"):
            cur[t[0]] = np.array([float(x) for x in t[1:4]], np.float32)
        elif cur is not None and t[0] in ("Ns", "Ni", "d"):
            cur[t[0]] = float(t[1])
    hs = HostScene(ar)
    try:
        hs.load_obj(str(tmp_path / "room.obj"))
        assert hs.counts()[1] == len(mtl) == 22
        for i, (name, p) in enumerate(mtl.items()):
            got = hs.material(i)
            ks = p.get("Ks", np.zeros(3, np.float32))
            lum = np.float32(0.2126) * ks[0] + np.float32(0.7152) * ks[1] + np.float32(0.0722) * ks[2]
            if p.get("d", 1.0) < 0.5:
                assert got["type"] == 2, name
            elif lum > np.float32(0.04):
                assert got["type"] == 1 and np.allclose(got["albedo"], ks), name
            else:
                assert got["type"] == 0 and got["tex"] == -1, name  # textures absent here: constant Kd
                assert np.allclose(got["albedo"], p.get("Kd", np.full(3, 0.5, np.float32))), name
    finally:
        hs.close()


def test_emitter_material_is_untextured(ar):
    """A quad light's emissive geometry gets a material of its own; it must
    carry no texture (index -1), whatever textures the scene holds."""
    hs = HostScene(ar)
    try:
        hs.add_texture(np.zeros((2, 2, 3), np.uint8))
        lt = ar.QuadLight((0.4, 0.999, 0.4), (0.2, 0.0, 0.0), (0.0, 0.0, 0.2), (15.0, 15.0, 15.0)).to_c()
        import ctypes as C
        assert hs.lib.ar_scene_add_light(hs.s, C.byref(lt), None) == 0
        nmesh, nmat, ntex = hs.counts()
        assert (nmesh, nmat, ntex) == (1, 1, 1)
        assert hs.material(0)["tex"] == -1
    finally:
        hs.close()


def _mutate_bytes(rng, b):
    b = bytearray(b)
    for _ in range(int(rng.integers(1, 7))):
        if not b:
            break
        at = int(rng.integers(0, len(b)))
        op = int(rng.integers(0, 5))
        if op == 0:
            b[at] ^= 1 << int(rng.integers(0, 8))
        elif op == 1:
            b[at] = int(rng.integers(0, 256))
        elif op == 2:
            del b[at:]
        elif op == 3:
            b[at:at + 4] = b"\x7f\xff\xff\xff"  # a huge chunk length or image size
        else:
            seg = b[at:at + int(rng.integers(1, 65))]
            k = int(rng.integers(0, len(b)))
            b[k:k] = seg
    return bytes(b)


_HOSTILE = [b"0", b"-1", b"-2147483648", b"4294967297", b"nan", b"inf", b"1e39", b"", b"/", b"1//1", b"3/3/3", b"x"]


def _mutate_text(rng, t):
    lines = t.split(b"\n")
    for _ in range(int(rng.integers(1, 5))):
        i = int(rng.integers(0, len(lines)))
        toks = lines[i].split(b" ")
        op = int(rng.integers(0, 3))
        if op == 0:
            toks[int(rng.integers(0, len(toks)))] = _HOSTILE[int(rng.integers(0, len(_HOSTILE)))]
            lines[i] = b" ".join(toks)
        elif op == 1:
            del lines[i]
            if not lines:
                lines = [b""]
        else:
            lines.insert(int(rng.integers(0, len(lines))), lines[i])
    return b"\n".join(lines)


def test_readers_survive_mutated_files(ar, tmp_path):
    """Seeded mutations of the PNG, OBJ and MTL fixtures (bit flips, truncation,
    huge lengths, hostile index and number tokens, repeated or missing lines):
    every call returns a status -- decoded images have the size they claim --
    and none aborts the process."""
    rng = np.random.default_rng(7)
    png = open(os.path.join(GOLD, "tex", "checker.png"), "rb").read()
    obj = open(os.path.join(GOLD, "scene.obj"), "rb").read()
    mtl = open(os.path.join(GOLD, "scene.mtl"), "rb").read()
    (tmp_path / "tex").mkdir()
    lib = ar.lib()
    for _ in range(400):
        (tmp_path / "tex" / "checker.png").write_bytes(_mutate_bytes(rng, png))
        try:
            img = ar.load_png(tmp_path / "tex" / "checker.png")
            assert img.ndim == 3 and img.shape[2] == 3
        except ar.ArError:
            pass
        if rng.integers(0, 2):
            (tmp_path / "scene.obj").write_bytes(_mutate_text(rng, obj))
            (tmp_path / "scene.mtl").write_bytes(mtl)
        else:
            (tmp_path / "scene.obj").write_bytes(obj)
            (tmp_path / "scene.mtl").write_bytes(_mutate_text(rng, mtl))
        s = C.c_void_p()
        assert lib.ar_scene_create(C.byref(s)) == 0
        n = C.c_uint32()
        rc = lib.ar_scene_load_obj(s, os.fsencode(str(tmp_path / "scene.obj")), None, C.byref(n))
        assert rc in (0, -1, -2, -3, -8), rc
        lib.ar_scene_destroy(s)
