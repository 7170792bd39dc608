"""CPU checks of the HIP library boundary: it loads, exports every entry point
declared in include/agentray.h, validates arguments on the host, and its host
SAH builder produces the oracle's tree.  No GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def declared_functions():
    src = open(os.path.join(ROOT, "include", "agentray.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ar_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol(ar):
    lib = ar.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_header_structs_match_ctypes(ar, tmp_path):
    """Struct layouts of include/agentray.h (compiled by the C compiler) equal
    the ctypes mirrors used by the Python host layer."""
    names = ["ar_material", "ar_light", "ar_camera", "ar_render_config", "ar_render_stats", "ar_scene_info",
             "ar_job_desc", "ar_job_stats", "ar_server_config", "ar_trace_config"]
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include "agentray.h"\nint main(){' +
                   "".join(f'printf("%zu\\n", sizeof({n}));' for n in names) + "return 0;}\n")
    exe = tmp_path / "sz"
    import subprocess
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    sizes = [int(x) for x in subprocess.check_output([str(exe)]).split()]
    mirrors = [ar.ArMaterial, ar.ArLight, ar.ArCamera, ar.ArRenderConfig, ar.ArRenderStats, ar.ArSceneInfo,
               ar.ArJobDesc, ar.ArJobStats, ar.ArServerConfig, ar.ArTraceConfig]
    assert sizes == [C.sizeof(m) for m in mirrors]


def test_host_validation_errors(ar):
    lib = ar.lib()
    s = C.c_void_p()
    assert lib.ar_scene_create(C.byref(s)) == 0
    mid = C.c_uint32()
    m = ar.ArMaterial(0, (C.c_float * 3)(0.5, 0.5, 0.5), 0.5, 1.5)
    assert lib.ar_scene_add_material(s, C.byref(m), C.byref(mid)) == 0
    bad = ar.ArMaterial(7, (C.c_float * 3)(0.5, 0.5, 0.5), 0.5, 1.5)
    assert lib.ar_scene_add_material(s, C.byref(bad), None) == -2   # UNSUPPORTED_TYPE
    pos = np.zeros((3, 3), np.float32)
    idx = np.array([[0, 1, 3]], np.uint32)                            # index out of range
    slots = np.zeros(1, np.uint32)
    mats = np.zeros(1, np.uint32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    assert lib.ar_scene_add_mesh(s, p(pos), 3, p(idx), 1, None, p(slots), 1, p(mats), 1, None, None) == -1
    assert b"out of range" in lib.ar_last_error()
    idx = np.array([[0, 1, 2]], np.uint32)
    two = np.zeros(2, np.uint32)
    assert lib.ar_scene_add_mesh(s, p(pos), 3, p(idx), 1, None, p(two), 2, p(mats), 1, None, None) == -1
    assert lib.ar_scene_add_mesh(s, p(pos), 3, p(idx), 0, None, p(slots), 1, p(mats), 1, None, None) == -1
    pos[0, 0] = np.nan
    assert lib.ar_scene_add_mesh(s, p(pos), 3, p(idx), 1, None, p(slots), 1, p(mats), 1, None, None) == -1
    lt = ar.ArLight(9, (C.c_float * 3)(), (C.c_float * 3)(), (C.c_float * 3)(), (C.c_float * 3)())
    assert lib.ar_scene_add_light(s, C.byref(lt), None) == -2
    assert lib.ar_scene_set_light_distribution(s, 5) == -2
    assert lib.ar_scene_destroy(s) == 0


def test_camera_matches_oracle(ar, orc):
    for (frm, at, fov, asp) in [((0.5, 0.5, -1.37), (0.5, 0.5, 0.5), 40.0, 1.0),
                                ((4, 2, 0.3), (4, 1.2, 5), 60.0, 16 / 9)]:
        a = ar.PinholeCamera.from_lookat(frm, at, (0, 1, 0), fov, asp).to_array()
        b = orc.camera_lookat(frm, at, (0, 1, 0), fov, asp)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("case", ["soup", "dups", "flat", "tiny"])
def test_builder_tree_equals_oracle(ar, orc, wl, case):
    if case == "soup":
        desc = wl.triangle_soup(20000, seed=11)
    elif case == "dups":       # many identical triangles: leaves above 4, SAH refuses splits
        desc = wl.triangle_soup(64, seed=12)
        P = desc.meshes[0].positions.reshape(-1, 3, 3)
        P[32:] = P[0]
        desc.meshes[0].positions = P.reshape(-1, 3)
    elif case == "flat":       # coplanar, axis-aligned triangles (zero-extent boxes, centroid ties)
        rng = np.random.default_rng(13)
        g = np.floor(rng.uniform(0, 8, size=(500, 1, 2))) / 8
        P = np.zeros((500, 3, 3), np.float32)
        P[:, :, 0] = g[:, :, 0] + np.array([0, 0.1, 0])
        P[:, :, 2] = g[:, :, 1] + np.array([0, 0, 0.1])
        desc = wl.triangle_soup(500, seed=14)
        desc.meshes[0].positions = P.reshape(-1, 3)
    else:
        desc = wl.triangle_soup(3, seed=15)
    sc = orc.OracleScene(desc)
    ob, om, op = sc.bvh(0)
    P = desc.meshes[0].positions.reshape(-1, 3, 3)
    boxes = np.concatenate([P.min(1), P.max(1)], 1)
    lb, lm, lp, depth = ar.bvh_build_host(boxes)
    assert np.array_equal(lb.view(np.uint32), ob.view(np.uint32))
    assert np.array_equal(lm, om)
    assert np.array_equal(lp, op)
    assert depth >= 1


def test_rust_binding_declares_the_header():
    """INTEGRATION.md's Rust extern block declares every function of
    include/agentray.h, with the same number of arguments, and nothing else."""
    src = open(os.path.join(ROOT, "include", "agentray.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    header = {}
    for name, args in re.findall(r"\b(ar_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", src, flags=re.S):
        a = " ".join(args.split())
        header[name] = 0 if a in ("", "void") else a.count(",") + 1
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    block = doc[doc.index('extern "C" {'):]
    block = block[:block.index("\n}")]
    block = re.sub(r"//[^\n]*", "", block)
    rust = {}
    for name, args in re.findall(r"pub fn (ar_[a-z0-9_]+)\s*\(([^)]*)\)", block, flags=re.S):
        a = " ".join(args.split()).rstrip(",")
        rust[name] = 0 if not a else a.count(",") + 1
    assert set(rust) == set(header), (sorted(set(header) - set(rust)), sorted(set(rust) - set(header)))
    assert {k: v for k, v in rust.items() if v != header[k]} == {}


def test_rust_binding_constants_match_the_header():
    """Every constant INTEGRATION.md's Rust binding declares has the header's
    value."""
    src = open(os.path.join(ROOT, "include", "agentray.h")).read()
    defs = {k: int(v, 0) for k, v in re.findall(r"#define\s+(AR_[A-Z0-9_]+)\s+\(?(0x[0-9a-fA-F]+|\d+)u?\)?", src)}
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    consts = {k: int(v) for k, v in re.findall(r"pub const (AR_[A-Z0-9_]+): u32 = (\d+);", doc)}
    assert consts and all(k in defs for k in consts), sorted(set(consts) - set(defs))
    assert {k: v for k, v in consts.items() if defs[k] != v} == {}
