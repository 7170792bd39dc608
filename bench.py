"""Benchmark driver.  Prints ONE JSON line.

Default workload (N = 1): BVH traversal over 2^20 random triangles (SAH BVH)
with 2^24 rays (half coherent 4096^2 pinhole primaries, half incoherent), all
resident in HBM.  One step = one closest-hit pass + one any-hit pass over all
2^24 rays; value = rays traced per second (both passes) in Mrays/s.

Under "extra" (unless --no-extra): path-tracer throughput in Mpaths/s on the
Cornell box (256^2, 16 spp) and the 280k-triangle room (1920x1080, 64 spp),
megakernel and wavefront schedules.  --workload serving: Poisson job arrivals
against the resident room scene (p50/p99 latency).

Roofline: the dominant kernel is the closest-hit traversal.  Its algorithmic
bytes per ray are B_ray = 32 (ray) + 8 (hit) + 64 N_node + 48 N_tri + 96 N_inst
+ 8 (N_spill_w + N_spill_r), the records the timed kernel reads and writes,
counted per ray by the counting build of the same kernel (ar_trace_count,
same grid and launch configuration, same rays); achieved = B_ray * rays /
average kernel duration (HIP events on the stream the kernel runs on).
`traffic` is the HBM traffic of one launch from two rocprofv3 --pmc child
runs (L2-to-memory read and write requests by size).  CPU baseline: the
oracle (C++ restatement of the reference CPU path) on the host cores, on a
bounded sample of the rays.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
METRIC = "Mpaths/s & Mrays/s at equal relMSE vs CPU ref; p50/p99 job latency (serving)"


def load_pkg():
    if "agent_ray_amd" in sys.modules:
        return sys.modules["agent_ray_amd"]
    pkg = os.path.join(ROOT, "agent-ray_amd")
    spec = importlib.util.spec_from_file_location("agent_ray_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["agent_ray_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def traversal_cpu_baseline(desc, rays, target_s=12.0):
    """Oracle closest-hit on a sample of the same rays (both halves), all host
    threads.  Returns (Mrays/s, N_node, N_tri, sample description)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    threads = cpu_threads()
    t0 = time.time()
    osc = orc.OracleScene(desc)
    log(f"oracle BVH built in {time.time() - t0:.1f}s")
    n = rays.shape[0]
    half = n // 2

    def sample(m):
        a = np.linspace(0, half - 1, m // 2).astype(np.int64)
        return np.concatenate([a, half + a])

    idx = sample(1 << 16)
    _, _, _, cnt, secs = osc.trace(rays[idx], nthreads=threads)
    rate = len(idx) / max(secs, 1e-9)
    m = int(min(n, max(1 << 16, rate * target_s)))
    m -= m % 2
    idx = sample(m)
    _, _, _, cnt, secs = osc.trace(rays[idx], nthreads=threads)
    return (len(idx) / secs / 1e6, float(cnt[0]) / len(idx), float(cnt[1]) / len(idx), threads,
            f"{len(idx)} rays (evenly spaced over both halves of the 2^24 set), closest-hit, {threads} threads, "
            f"{secs:.1f}s, {cpu_model()}")


def bench_traversal(args, ar, torch):
    wl = ar.workloads
    n_tris, n_rays = 1 << 20, 1 << 24
    log("generating 2^20-triangle soup and 2^24 rays")
    desc = wl.triangle_soup(n_tris, seed=0)
    rays = wl.soup_rays(n_rays, seed=1)
    t0 = time.time()
    sc = ar.Scene.from_desc(desc)
    build_s = time.time() - t0
    log(f"scene built+uploaded in {build_s:.2f}s, {sc.info.n_nodes} wide nodes, depth {sc.info.max_depth}")
    dr = torch.from_numpy(rays).cuda()
    hits = torch.empty((n_rays, 2), dtype=torch.float32, device="cuda")
    occ = torch.empty((n_rays,), dtype=torch.int32, device="cuda")
    lib = ar.lib()
    # A stream of our own (default priority): torch's default stream has the
    # NULL handle, which the library reads as "the scene's own stream", and
    # the events below must sit on the stream the kernels run on.
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sh = C_void(stream.cuda_stream)
    assert stream.cuda_stream != 0
    tuned = None
    if args.trace_config:
        sc.set_trace_config(*[int(v) for v in args.trace_config.split(",")])
        if args.anyhit_config:
            sc.set_anyhit_config(*[int(v) for v in args.anyhit_config.split(",")])
    elif not args.no_autotune:
        # per-scene autotuning of the traversal launch on 2^20 rays from both halves
        idx = torch.cat([torch.arange(0, 1 << 19), (n_rays // 2) + torch.arange(0, 1 << 19)]).cuda()
        sample = dr.index_select(0, idx).contiguous()
        tuned = sc.autotune(sample, stream)
        log(f"autotuned traversal: {tuned}")
        del sample

    def closest():
        ar._check(lib.ar_trace(sc.handle, C_void(dr.data_ptr()), n_rays, C_void(hits.data_ptr()), None,
                               ar.AR_ASYNC, sh))

    def anyhit():
        ar._check(lib.ar_occluded(sc.handle, C_void(dr.data_ptr()), n_rays, C_void(occ.data_ptr()), ar.AR_ASYNC, sh))

    if tuned is not None:
        # refinement on the workload itself (untimed): the candidates, best of
        # three launches each, then the three fastest again; closest and any
        # hit
        def launch_ms(fn, reps=3):
            best = float("inf")
            for _ in range(reps):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                fn()
                e1.record(stream)
                torch.cuda.synchronize()
                best = min(best, e0.elapsed_time(e1))
            return best
        cands = ar.TUNE_CANDIDATES
        best_c, best_a = ar.config_key(sc.trace_config()), ar.config_key(sc.anyhit_config())
        times_c = {}

        def try_c(c):
            sc.set_trace_config(*c)
            sc.set_anyhit_config(*best_a)
            closest()
            times_c[c] = launch_ms(closest)
        for c in [best_c] + [ar.with_balance(c, best_c[6]) for c in cands if ar.with_balance(c, best_c[6]) != best_c]:
            try_c(c)
        for other in (ar.other_park_slots, ar.other_balance):
            flip = other(min(times_c, key=times_c.get))
            if flip not in times_c:
                try_c(flip)

        def settle(times, run):
            # the three fastest again, best of five each, twice in turn: one
            # best-of-three differs by 1-2% between runs, which reordered
            # candidates that close
            top = sorted(times, key=times.get)[:3]
            again = {c: float("inf") for c in top}
            for _ in range(2):
                for c in top:
                    again[c] = min(again[c], run(c))
            best = min(again, key=again.get)
            return best, again[best]

        def rerun_c(c):
            sc.set_trace_config(*c)
            sc.set_anyhit_config(*best_a)
            return launch_ms(closest, 5)
        best_c, best_t = settle(times_c, rerun_c)
        sc.set_trace_config(*best_c)
        times_a = {}

        def try_a(a):
            sc.set_anyhit_config(*a)
            anyhit()
            times_a[a] = launch_ms(anyhit)
        for a in [best_a] + [ar.with_balance(a, best_a[6]) for a in cands if ar.with_balance(a, best_a[6]) != best_a]:
            try_a(a)
        for other in (ar.other_park_slots, ar.other_balance):
            flip = other(min(times_a, key=times_a.get))
            if flip not in times_a:
                try_a(flip)

        def rerun_a(a):
            sc.set_anyhit_config(*a)
            return launch_ms(anyhit, 5)
        best_a, best_ta = settle(times_a, rerun_a)
        sc.set_trace_config(*best_c)
        sc.set_anyhit_config(*best_a)
        log(f"refined on the workload: closest {best_c} ({best_t:.2f} ms), any hit {best_a} ({best_ta:.2f} ms)")

    for _ in range(args.warmup):
        closest()
        anyhit()
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        e = ev[k]
        e[0].record(stream)
        closest()
        e[1].record(stream)
        e[2].record(stream)
        anyhit()
        e[3].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_closest = [ev[k][0].elapsed_time(ev[k][1]) for k in range(args.steps)]
    t_any = [ev[k][2].elapsed_time(ev[k][3]) for k in range(args.steps)]
    ms_step = wall * 1e3 / args.steps
    value = 2 * n_rays / (wall / args.steps) / 1e6
    # quick self-check: the GPU answer on a few rays stays consistent
    t_h = hits[:4096, 0].cpu().numpy()
    assert np.all(np.isfinite(t_h) | np.isinf(t_h))
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "bvh_traversal_soup_1M_tris_16M_rays", "triangles": n_tris, "rays": n_rays,
                   "rays_coherent": n_rays // 2, "rays_incoherent": n_rays // 2, "step": "closest-hit + any-hit",
                   "bvh": "SAH full sweep, max leaf 4", "parallelism": "single-gpu"},
        "closest_mrays_s": round(n_rays / (np.mean(t_closest) / 1e3) / 1e6, 1),
        "anyhit_mrays_s": round(n_rays / (np.mean(t_any) / 1e3) / 1e6, 1),
        "closest_kernel_ms": round(float(np.mean(t_closest)), 3),
        "anyhit_kernel_ms": round(float(np.mean(t_any)), 3),
        "scene_build_s": round(build_s, 2),
        "launch_config": dict(sc.trace_config(), anyhit=sc.anyhit_config()),
    }
    # per-half rates (coherent / incoherent), one extra timed call each
    for name, sl in (("coherent", slice(0, n_rays // 2)), ("incoherent", slice(n_rays // 2, n_rays))):
        sub = dr[sl]
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ar._check(lib.ar_trace(sc.handle, C_void(sub.data_ptr()), sub.shape[0], C_void(hits.data_ptr()), None,
                               ar.AR_ASYNC, sh))
        e1.record(stream)
        torch.cuda.synchronize()
        out[f"closest_{name}_mrays_s"] = round(sub.shape[0] / (e0.elapsed_time(e1) / 1e3) / 1e6, 1)
    # Roofline of the dominant kernel (closest hit).  Algorithmic bytes per
    # ray from the timed kernel's own work: the counting build of the same
    # kernel, same grid and launch configuration, over the same rays.
    cnt = sc.trace_counts(dr)
    per = {k: cnt[k] / n_rays for k in ("node_visits", "tri_tests", "inst_tests", "spill_writes", "spill_reads")}
    b_ray = (32 + 8 + 64 * per["node_visits"] + 48 * per["tri_tests"] + 96 * per["inst_tests"]
             + 8 * (per["spill_writes"] + per["spill_reads"]))
    kernel_s = float(np.mean(t_closest)) / 1e3
    achieved = b_ray * n_rays / kernel_s / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                       "kernel": "k_trace<closest, near-first>", "kernel_ms": round(float(np.mean(t_closest)), 3),
                       "bytes_per_ray": round(b_ray, 1),
                       "per_ray": {k: round(v, 3) for k, v in per.items()},
                       "basis": "32 B ray + 8 B hit + 64 B per node record + 48 B per triangle record + 96 B per "
                                "instance record + 8 B per spilled stack entry written or read, counts of the "
                                "timed launch (ar_trace_count)"}
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle) on a ray sample")
        rate, nn, nt, threads, sample = traversal_cpu_baseline(desc, rays)
        out["cpu_baseline"] = {"value": round(rate, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
                               "sample": sample}
        # the reference's own traversal (binary tree, left child first, box
        # test at pop) on the same sample, in the reference's record sizes
        out["roofline"]["reference_order"] = {
            "bytes_per_ray": round(32 + 8 + 32 * nn + 36 * nt, 1), "n_node": round(nn, 2), "n_tri": round(nt, 2),
            "note": "oracle's reference-order visits; not the timed kernel's work"}
    if not args.no_traffic:
        c = sc.trace_config()
        cfg = ",".join(str(v) for v in ar.config_key(c))
        traffic, detail = measure_traffic(cfg)
        out["roofline"]["traffic"] = None if traffic is None else round(traffic)
        out["roofline"]["traffic_detail"] = detail
        if traffic is not None:
            dram = traffic / kernel_s / 1e9
            out["roofline"]["traffic_gbs"] = round(dram, 1)
            out["roofline"]["traffic_frac"] = round(dram / HBM_PEAK_GBS, 4)
    return out


def bench_copy(args, torch):
    """1 GiB fp32 device-to-device copy (torch's elementwise copy kernel):
    reads and writes exactly 2^30 bytes per launch."""
    n = 1 << 28
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(max(args.warmup, 1)):
        y.copy_(x)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        y.copy_(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    return {"workload": "copy_1GiB_f32", "bytes_read": 4 * n, "bytes_written": 4 * n, "ms_per_copy": round(ms, 4),
            "gbs": round(8 * n / (ms / 1e3) / 1e9, 1)}


def pmc_probe(args, ar, torch):
    """Child of measure_traffic: one closest-hit launch over the full ray set
    with the given launch configuration, then one 1 GiB device copy (the
    calibration kernel), each followed by a synchronize."""
    wl = ar.workloads
    desc = wl.triangle_soup(1 << 20, seed=0)
    rays = wl.soup_rays(1 << 24, seed=1)
    sc = ar.Scene.from_desc(desc)
    if args.trace_config:
        sc.set_trace_config(*[int(v) for v in args.trace_config.split(",")])
    dr = torch.from_numpy(rays).cuda()
    hits = torch.empty((rays.shape[0], 2), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ar._check(ar.lib().ar_trace(sc.handle, C_void(dr.data_ptr()), rays.shape[0], C_void(hits.data_ptr()), None,
                                ar.AR_ASYNC, None))
    torch.cuda.synchronize()
    n = 1 << 28
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "ok", "copy_bytes": 4 * n}), flush=True)


READ_COUNTERS = ("TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum", "TCC_EA0_RDREQ_sum")
WRITE_COUNTERS = ("TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum")


def _dram_bytes(v):
    """Bytes moved between L2 and memory from request counts by size: reads
    32 / 64 / 128 B (a request of none of those sizes, if any, as 64 B),
    writes 64 B or else 32 B."""
    r32, r64, r128, rall = (v.get(k, 0.0) for k in READ_COUNTERS)
    rest = max(rall - r32 - r64 - r128, 0.0)
    w, w64 = (v.get(k, 0.0) for k in WRITE_COUNTERS)
    return 32 * r32 + 64 * r64 + 128 * r128 + 64 * rest, 32 * (w - w64) + 64 * w64


def measure_traffic(cfg: str, kernel_key: str = "k_trace<false, true"):
    """HBM bytes of one closest-hit launch from two rocprofv3 --pmc child runs
    of the probe (read requests by size, then write requests by size).  The
    same counters on the probe's 1 GiB copy (2^30 bytes read, 2^30 written)
    check the conversion; no scale is applied.  Returns (bytes, detail) or
    (None, reason)."""
    import csv
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None, "rocprofv3 not found"
    env = dict(os.environ, TMPDIR="/tmp")
    trav, copy = {}, {}
    for counters in (READ_COUNTERS, WRITE_COUNTERS):
        d = tempfile.mkdtemp(prefix="ar_pmc_", dir="/tmp")
        cmd = [prof, "--pmc", *counters, "--output-format", "csv", "-d", d, "-o", "probe", "--",
               os.path.realpath(sys.executable), os.path.join(ROOT, "bench.py"), "--workload", "pmc-probe",
               "--trace-config", cfg]
        log(f"traffic: rocprofv3 --pmc {' '.join(counters)}")
        try:
            r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                               timeout=240)
        except subprocess.TimeoutExpired:
            return None, f"rocprofv3 --pmc {counters[0]} timed out"
        files = [os.path.join(dp, f) for dp, _, fs in os.walk(d) for f in fs if f.endswith("counter_collection.csv")]
        if not files:
            return None, f"rocprofv3 --pmc {counters[0]}: no counter file (rc {r.returncode})"
        for row in csv.DictReader(open(files[0])):
            v, name, cn = float(row["Counter_Value"]), row["Kernel_Name"], row["Counter_Name"]
            if kernel_key in name:
                trav[cn] = trav.get(cn, 0.0) + v
            elif "copyBuffer" in name or "copy_kernel" in name:
                copy[cn] = max(copy.get(cn, 0.0), v)
        if not any(k in trav for k in counters) or not any(k in copy for k in counters):
            return None, f"rocprofv3 --pmc {counters[0]}: kernels not found"
    rd, wr = _dram_bytes(trav)
    crd, cwr = _dram_bytes(copy)
    return rd + wr, {"read_bytes": round(rd), "write_bytes": round(wr),
                     "copy_check": {"read_bytes": round(crd), "write_bytes": round(cwr), "true_bytes": 1 << 30},
                     "counters": {k: trav.get(k) for k in READ_COUNTERS + WRITE_COUNTERS},
                     "source": "rocprofv3 --pmc TCC_EA0_RDREQ_{32B,64B,128B} / TCC_EA0_WRREQ{,_64B} on one "
                               "closest-hit launch, bytes by request size"}


def C_void(p):
    import ctypes
    return ctypes.c_void_p(p)


def autotune_scene(ar, torch, sc, desc, aspect):
    """Per-scene traversal autotuning on ray samples shaped like what the path
    tracer's stages trace.  Closest hit (extension stage): 2^18 camera rays
    and ~2^19 bounce rays leaving the surface points that camera rays hit, in
    uniform directions.  Any hit (shadow stage): segments from those points to
    random points on the lights (quad lights) or to the light position."""
    wl = ar.workloads
    cd = desc.camera
    rng = np.random.default_rng(3)
    m = 1 << 18
    cam16 = wl.camera_lookat(cd.lookfrom, cd.lookat, cd.vup, cd.vfov, aspect)
    cr = wl.camera_rays(cam16, rng.uniform(0, 1, 2 * m), rng.uniform(0, 1, 2 * m)).astype(np.float32)
    t, prim, _ = sc.intersect(cr)
    hit = prim != 0xFFFFFFFF
    crh = cr[hit]
    # surface points, pulled back a little along the camera ray
    p = crh[:, 0:3] + crh[:, 4:7] * (t[hit][:, None] - 1e-3).astype(np.float32)
    k = p.shape[0]
    d = wl.normalize_f32(rng.standard_normal((k, 3)).astype(np.float32))
    br = np.zeros((k, 8), np.float32)
    br[:, 0:3], br[:, 3], br[:, 4:7], br[:, 7] = p, 1e-4, d, wl.F32_MAX
    parts = [cr[:m], br]
    shadow = None
    if desc.lights and k:
        sel = rng.integers(0, len(desc.lights), k)
        tgt = np.zeros((k, 3), np.float32)
        for li, L in enumerate(desc.lights):
            mask = sel == li
            u, v = rng.uniform(0, 1, (2, int(mask.sum()), 1))
            tgt[mask] = (np.asarray(L.position, np.float32) + u * np.asarray(L.edge_u, np.float32)
                         + v * np.asarray(L.edge_v, np.float32))
        dv = tgt - p
        dist = np.linalg.norm(dv, axis=1).astype(np.float32)
        sr = np.zeros((k, 8), np.float32)
        sr[:, 0:3], sr[:, 3], sr[:, 4:7] = p, 1e-4, wl.normalize_f32(dv.astype(np.float32))
        sr[:, 7] = dist * np.float32(1 - 1e-4)
        shadow = sr
    # closest hit on camera + bounce rays (the extension stage), any hit on
    # the shadow segments (the shadow stage) when the scene has lights
    sample = torch.from_numpy(np.concatenate(parts).astype(np.float32)).cuda()
    if shadow is None:
        cfg = sc.autotune(sample)
    else:
        cfg = sc.autotune(sample, which="closest")
        ssample = torch.from_numpy(shadow).cuda()
        sc.autotune(ssample, which="anyhit")
        cfg["anyhit"] = sc.anyhit_config()
        del ssample
    del sample
    return cfg


def bench_pt(ar, torch, which, schedules=("megakernel", "wavefront")):
    """Path tracer throughput on the Cornell box (256^2, 16 spp, depth 8) or
    the procedural room (1920x1080, 64 spp, depth 16).  Mpaths/s = W*H*spp /
    render wall time (scene upload excluded, film readback included)."""
    wl = ar.workloads
    desc = wl.cornell_box() if which == "cornell" else wl.procedural_room()
    t0 = time.time()
    sc = ar.Scene.from_desc(desc)
    build = time.time() - t0
    cd = desc.camera
    cam = ar.PinholeCamera.from_array(wl.camera_lookat(cd.lookfrom, cd.lookat, cd.vup, cd.vfov,
                                                       desc.width / desc.height))
    out = {"width": desc.width, "height": desc.height, "spp": desc.spp, "max_depth": desc.max_depth,
           "triangles": desc.n_triangles, "scene_build_s": round(build, 2)}
    out["launch_config"] = autotune_scene(ar, torch, sc, desc, desc.width / desc.height)
    # refine the closest-hit launch on one chunk of the real job (16 spp)
    tuner = ar.PathTracer(ar.PathTracerConfig(spp=desc.spp, max_depth=desc.max_depth, rr_depth=3), seed=0)
    out["launch_refined"] = tuner.tune_launch(sc, cam, desc.width, desc.height, spp=min(16, desc.spp))
    out["launch_config"] = dict(sc.trace_config(), anyhit=sc.anyhit_config())
    # and the queue chunk (paths per wavefront pass), kept on the scene
    out["chunk_paths"] = tuner.tune_chunk(sc, cam, desc.width, desc.height)
    films = {}
    for sched in schedules:
        pt = ar.PathTracer(ar.PathTracerConfig(spp=desc.spp, max_depth=desc.max_depth, rr_depth=3), seed=0,
                           schedule=ar.Schedule.MEGAKERNEL if sched == "megakernel" else ar.Schedule.WAVEFRONT)
        pt.render(sc, cam, desc.width, desc.height)  # warm-up (allocations)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        film = pt.render(sc, cam, desc.width, desc.height)
        dt = time.perf_counter() - t0
        st = pt.last_stats
        films[sched] = film.pixels
        out[sched] = {"mpaths_s": round(st.paths / dt / 1e6, 2), "mrays_s": round(st.rays / dt / 1e6, 1),
                      "seconds": round(dt, 4), "rays": int(st.rays)}
        log(f"pt {which} {sched}: {out[sched]}")
    if len(films) == 2:
        a, b = films.values()
        out["schedules_bitwise_equal"] = bool(np.array_equal(a.view(np.uint32), b.view(np.uint32)))
        out["wavefront_speedup"] = round(out["megakernel"]["seconds"] / out["wavefront"]["seconds"], 3)
    if which == "room" and films:
        out["film_output"] = bench_film_output(ar, torch, next(iter(films.values())), desc.spp)
    return out


def bench_builders(ar, torch):
    """The 2^20-triangle soup committed with the reference's host SAH, the
    device LBVH builder and the spatial-split builder (commit wall time: tree
    build, flattening, upload), then the 2^24 bench rays traced through each
    with the same launches (closest and any hit 256 x 12 x 8 x 8 x 16 x 3,
    balance 1), best of three launches each."""
    wl = ar.workloads
    desc = wl.triangle_soup(1 << 20, seed=0)
    rays = wl.soup_rays(1 << 24, seed=1)
    dr = torch.from_numpy(rays).cuda()
    n = rays.shape[0]
    hits = torch.empty((n, 2), dtype=torch.float32, device="cuda")
    occ = torch.empty((n,), dtype=torch.int32, device="cuda")
    stream = torch.cuda.Stream()
    sh = C_void(stream.cuda_stream)
    lib = ar.lib()
    out = {}
    for builder in ("sah", "lbvh", "split"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sc = ar.Scene.from_desc(desc, builder=builder)
        build_s = time.perf_counter() - t0
        sc.set_trace_config(256, 12, 8, 8, 16, 3, 1)
        sc.set_anyhit_config(256, 12, 8, 8, 16, 3, 1)
        res = {"commit_s": round(build_s, 3), "binary_nodes": int(sc.info.n_nodes), "max_depth": int(sc.info.max_depth)}
        for name, fn in (("closest", lambda: ar._check(lib.ar_trace(sc.handle, C_void(dr.data_ptr()), n,
                                                                      C_void(hits.data_ptr()), None, ar.AR_ASYNC, sh))),
                         ("anyhit", lambda: ar._check(lib.ar_occluded(sc.handle, C_void(dr.data_ptr()), n,
                                                                       C_void(occ.data_ptr()), ar.AR_ASYNC, sh)))):
            fn()
            best = float("inf")
            for _ in range(3):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                fn()
                e1.record(stream)
                torch.cuda.synchronize()
                best = min(best, e0.elapsed_time(e1))
            res[f"{name}_mrays_s"] = round(n / best / 1e3, 1)
        out[builder] = res
        del sc
    log(f"tree builders (soup 1M): {out}")
    return out


def bench_film_output(ar, torch, pixels, spp):
    """Film output stage on the rendered film (device resident): Reinhard +
    gamma 2.2 + 8-bit encode (Film::to_rgb_image; 12 B read + 3 B written
    per pixel), timed with HIP events over 20 launches, and the half-float
    EXR file (Film::save_exr) timed on the wall clock, device film to file."""
    import tempfile
    h, w = pixels.shape[:2]
    dev = torch.from_numpy(np.ascontiguousarray(pixels)).cuda()
    st = torch.cuda.current_stream()
    ar.film_to_rgb8(dev, ar.ToneMapper.Reinhard, 2.2, 1.0 / spp)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    e0.record(st)
    for _ in range(reps):
        ar.film_to_rgb8(dev, ar.ToneMapper.Reinhard, 2.2, 1.0 / spp)
    e1.record(st)
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) / reps * 1e3
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "film.exr")
        ar.film_save_exr(dev, path, 1.0 / spp)
        t0 = time.perf_counter()
        ar.film_save_exr(dev, path, 1.0 / spp)
        exr_ms = (time.perf_counter() - t0) * 1e3
        exr_bytes = os.path.getsize(path)
    res = {"rgb8_us": round(us, 1), "rgb8_gbs": round(15.0 * w * h / (us * 1e-6) / 1e9, 1),
           "exr_ms": round(exr_ms, 2), "exr_bytes": exr_bytes, "d2h_bytes_rgb8": 3 * w * h,
           "d2h_bytes_exr": 6 * w * h, "d2h_bytes_f32": 12 * w * h}
    log(f"film output {w}x{h}: {res}")
    return res


ROOM_JOBS = {  # integrator: (width, height, spp, max_depth), the bench's room configurations
    "pt": (1920, 1080, 64, 16),
    "bdpt": (1280, 720, 16, 10),
    "gdpt": (1280, 720, 32, 16),
}


def bench_render_room(args, ar, torch):
    """One integrator alone on the room at its bench configuration (ROOM_JOBS;
    BDPT with the power heuristic beta 2, GDPT with 50 CG iterations): the
    launches given by --trace-config / --anyhit-config (else autotuned on ray
    samples), the queue chunk given by --chunk (samples per pixel; else the
    default), the wavefront schedule, one warm-up render, then --steps timed
    renders.  For per-stage profiles."""
    wl = ar.workloads
    integ = args.integrator
    w, h, spp, depth = ROOM_JOBS[integ]
    desc = wl.procedural_room(width=w, height=h)
    sc = ar.Scene.from_desc(desc)
    cd = desc.camera
    cam = ar.PinholeCamera.from_array(wl.camera_lookat(cd.lookfrom, cd.lookat, cd.vup, cd.vfov, w / h))
    if args.trace_config:
        sc.set_trace_config(*[int(v) for v in args.trace_config.split(",")])
        sc.set_anyhit_config(*[int(v) for v in (args.anyhit_config or args.trace_config).split(",")])
    else:
        autotune_scene(ar, torch, sc, desc, w / h)
    if integ == "pt":
        it = ar.PathTracer(ar.PathTracerConfig(spp=spp, max_depth=depth, rr_depth=3), seed=0,
                           sample_chunk=args.chunk)
    elif integ == "bdpt":
        it = ar.BidirectionalPathTracer(ar.BdptConfig(spp=spp, max_depth=depth, rr_depth=3), seed=0,
                                        sample_chunk=args.chunk)
    else:
        it = ar.GradientDomainPathTracer(ar.GdptConfig(spp=spp, max_depth=depth, rr_depth=3, alpha=0.2,
                                                       poisson_iterations=50), seed=0, sample_chunk=args.chunk)
    it.render(sc, cam, w, h)
    torch.cuda.synchronize()
    times = []
    for _ in range(max(1, args.steps)):
        t0 = time.perf_counter()
        it.render(sc, cam, w, h)
        times.append(time.perf_counter() - t0)
        log(f"{integ} room render: {times[-1] * 1e3:.1f} ms")
    dt = min(times)
    st = it.last_stats
    name = f"{integ}_room_{w}x{h}x{spp}spp"
    return {name: {"launch_config": dict(sc.trace_config(), anyhit=sc.anyhit_config()), "sample_chunk": args.chunk,
                   "msamples_s": round(w * h * spp / dt / 1e6, 2), "mpaths_s": round(st.paths / dt / 1e6, 2),
                   "mrays_s": round(st.rays / dt / 1e6, 1), "seconds": round(dt, 4), "rays": int(st.rays)}}


def bench_bdpt_gdpt(ar, torch):
    """The room at 1280x720: BDPT 16 spp, depth 10, power heuristic beta 2;
    GDPT 32 spp, depth 16, base + two shifted paths per sample, screened
    Poisson alpha 0.2 with 50 CG iterations.  The scene is autotuned on
    ray samples, then each integrator refines the launches on a short render
    of its own job (PathTracer.tune_launch) and picks its queue chunk
    (tune_chunk).  Each timed render follows a
    warm-up render; the Poisson solve is also timed alone (device buffers,
    HIP events, allocation of its work planes included) against its
    algorithmic bytes: 120 B per pixel and iteration (two passes over fp32
    RGB planes: 4 and 6 plane accesses of 12 B) plus 204 B per pixel of
    set-up (de-interleave, right-hand side, first residual, output)."""
    wl = ar.workloads
    desc = wl.procedural_room(width=1280, height=720)
    sc = ar.Scene.from_desc(desc)
    cd = desc.camera
    w, h = 1280, 720
    cam = ar.PinholeCamera.from_array(wl.camera_lookat(cd.lookfrom, cd.lookat, cd.vup, cd.vfov, w / h))
    out = {}
    tuned = autotune_scene(ar, torch, sc, desc, w / h)
    sample_c, sample_a = ar.config_key(sc.trace_config()), ar.config_key(sc.anyhit_config())
    bd = ar.BidirectionalPathTracer(ar.BdptConfig(spp=16, max_depth=10, rr_depth=3), seed=0)
    # refine both launches on one 4-spp chunk of the BDPT job itself (subpath
    # extension = closest hit, connection = any hit)
    bd_refined = bd.tune_launch(sc, cam, w, h, spp=4)
    bd_chunk = bd.tune_chunk(sc, cam, w, h)
    bd.render(sc, cam, w, h)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bd.render(sc, cam, w, h)
    dt = time.perf_counter() - t0
    st = bd.last_stats
    out["bdpt_room_1280x720x16spp"] = {"launch_config": dict(sc.trace_config(), anyhit=sc.anyhit_config()),
                                       "launch_refined": bd_refined, "chunk_samples": bd_chunk // (w * h),
                                       "msamples_s": round(w * h * 16 / dt / 1e6, 2),
                                       "mrays_s": round(st.rays / dt / 1e6, 1), "seconds": round(dt, 4),
                                       "rays": int(st.rays)}
    log(f"bdpt room: {out['bdpt_room_1280x720x16spp']}")
    gd = ar.GradientDomainPathTracer(ar.GdptConfig(spp=32, max_depth=16, rr_depth=3, alpha=0.2,
                                                   poisson_iterations=50), seed=0)
    # GDPT traces path-tracer stages: start again from the sample-tuned
    # launches and refine them on one 8-spp chunk of its own job
    sc.set_trace_config(*sample_c)
    sc.set_anyhit_config(*sample_a)
    gd_refined = gd.tune_launch(sc, cam, w, h, spp=8)
    gd_chunk = gd.tune_chunk(sc, cam, w, h)
    gd.render(sc, cam, w, h)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gd.render(sc, cam, w, h)
    dt = time.perf_counter() - t0
    st = gd.last_stats
    g = {"mpaths_s": round(st.paths / dt / 1e6, 2), "mrays_s": round(st.rays / dt / 1e6, 1),
         "seconds": round(dt, 4), "paths": int(st.paths),
         "launch_config": dict(sc.trace_config(), anyhit=sc.anyhit_config()), "launch_refined": gd_refined,
         "chunk_paths": gd_chunk}
    # the Poisson solve alone on the render's own films
    films = [torch.from_numpy(np.ascontiguousarray(gd.aux[k], np.float32)).cuda() for k in range(3)]
    dst = torch.empty_like(films[0])
    stream = torch.cuda.Stream()
    res = __import__("ctypes").c_double()
    lib = ar.lib()
    call = lambda: ar._check(lib.ar_poisson_solve(C_void(films[0].data_ptr()), C_void(films[1].data_ptr()),
                                                  C_void(films[2].data_ptr()), w, h, 0.2, 50, 1,
                                                  C_void(dst.data_ptr()), __import__("ctypes").byref(res),
                                                  ar.AR_ASYNC, C_void(stream.cuda_stream)))
    with torch.cuda.stream(stream):
        call()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(5):
            call()
        e1.record(stream)
        torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    alg = (120 * 50 + 204) * w * h
    g["poisson_cg50_ms"] = round(ms, 3)
    g["poisson_achieved_gbs"] = round(alg / (ms / 1e3) / 1e9, 1)
    g["poisson_frac_of_hbm"] = round(alg / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
    out["gdpt_room_1280x720x32spp"] = g
    log(f"gdpt room: {g}")
    return out


def _crop(img, x0, y0, n):
    return img[y0:y0 + n, x0:x0 + n]


def bench_cpu_and_quality(ar, torch):
    """CPU baselines of the path-sampling configurations (the oracle, a C++
    restatement of the reference CPU path, on all host threads; Poisson SOR
    single-threaded as the reference runs it), each on a bounded sample of
    the workload, and relMSE at equal spp: the GPU's and the oracle's 16-spp
    renders against a 4096-spp oracle render of the same pixels
    (mean over pixel-channels of (I - R)^2 / (R^2 + 1e-2), 4 seeds each)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    wl = ar.workloads
    threads = cpu_threads()
    model = cpu_model()
    out = {}

    def cam_for(desc, w, h):
        cd = desc.camera
        return wl.camera_lookat(cd.lookfrom, cd.lookat, cd.vup, cd.vfov, w / h)
    # C1: the whole Cornell workload (256^2 x 16 spp, depth 8)
    cor = wl.cornell_box()
    osc = orc.OracleScene(cor)
    c1 = cam_for(cor, 256, 256)
    _, _, _, secs = osc.render_pt(c1, 256, 256, 16, 8, 3, rng_mode=0, nthreads=threads)
    out["pt_cornell_256x256x16spp"] = {"value": round(256 * 256 * 16 / secs / 1e6, 3), "unit": "Mpaths/s",
                                       "cores": threads, "kind": "port", "seconds": round(secs, 3),
                                       "sample": f"the whole workload, {threads} threads, {model}"}
    log(f"cpu: C1 {out['pt_cornell_256x256x16spp']}")
    # relMSE on the central 64 x 64 pixels of C1
    x0 = y0 = 96
    n = 64
    ref, _, _, rsecs = osc.render_pt(c1, 256, 256, 4096, 8, 3, rng_mode=1, seed=12345, nthreads=threads,
                                     rows=(y0, y0 + n), cols=(x0, x0 + n))
    ref = _crop(ref, x0, y0, n)
    sc = ar.Scene.from_desc(cor)
    cam = ar.PinholeCamera.from_array(c1)

    def relmse(x):
        return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))
    rg = [relmse(_crop(ar.PathTracer(ar.PathTracerConfig(spp=16, max_depth=8, rr_depth=3), seed=100 + k)
                       .render(sc, cam, 256, 256).pixels, x0, y0, n)) for k in range(4)]
    ro = [relmse(_crop(osc.render_pt(c1, 256, 256, 16, 8, 3, rng_mode=1, seed=200 + k, nthreads=threads,
                                     rows=(y0, y0 + n), cols=(x0, x0 + n))[0], x0, y0, n)) for k in range(4)]
    out["relmse_cornell_16spp"] = {"gpu": round(float(np.mean(rg)), 6), "cpu": round(float(np.mean(ro)), 6),
                                   "ratio": round(float(np.mean(rg) / np.mean(ro)), 4),
                                   "reference": f"oracle 4096 spp on pixels [{x0},{x0 + n})^2 of 256^2 "
                                                f"({rsecs:.1f} s)", "seeds": 4}
    log(f"cpu: relMSE {out['relmse_cornell_16spp']}")
    # C3 crops: four 64 x 64 crops of the 1920x1080 room at 64 spp, depth 16
    room = wl.procedural_room()
    orm = orc.OracleScene(room)
    c3 = cam_for(room, 1920, 1080)
    crops = [(480, 270), (1376, 270), (480, 746), (1376, 746)]
    secs = 0.0
    for (cx, cy) in crops:
        secs += orm.render_pt(c3, 1920, 1080, 64, 16, 3, rng_mode=0, nthreads=threads, rows=(cy, cy + 64),
                              cols=(cx, cx + 64))[3]
    out["pt_room_1920x1080x64spp"] = {"value": round(4 * 64 * 64 * 64 / secs / 1e6, 3), "unit": "Mpaths/s",
                                      "cores": threads, "kind": "port", "seconds": round(secs, 3),
                                      "sample": f"crops 64x64 at {crops} of 1920x1080, 64 spp, {threads} threads"}
    log(f"cpu: C3 {out['pt_room_1920x1080x64spp']}")
    # C4: BDPT 16 spp depth 10 and GDPT 32 spp depth 16 on a 128 x 128 crop of 1280x720
    c4 = cam_for(room, 1280, 720)
    win = (576, 704, 296, 424)
    _, _, secs = orc.render_bdpt(orm, c4, 1280, 720, 16, 10, 3, 0, 2.0, rng_mode=0, nthreads=threads, window=win)
    out["bdpt_room_1280x720x16spp"] = {"value": round(128 * 128 * 16 / secs / 1e6, 4), "unit": "Msamples/s",
                                       "cores": threads, "kind": "port", "seconds": round(secs, 3),
                                       "sample": f"pixels {win} (x0, x1, y0, y1) of 1280x720, 16 spp, depth 10"}
    log(f"cpu: C4 BDPT {out['bdpt_room_1280x720x16spp']}")
    p, gx, gy, secs = orc.render_gdpt(orm, c4, 1280, 720, 32, 16, 3, rng_mode=0, nthreads=threads, window=win)
    out["gdpt_room_1280x720x32spp"] = {"value": round(3 * 128 * 128 * 32 / secs / 1e6, 4), "unit": "Mpaths/s",
                                       "cores": threads, "kind": "port", "seconds": round(secs, 3),
                                       "sample": f"pixels {win} of 1280x720, 32 spp x 3 paths, depth 16"}
    log(f"cpu: C4 GDPT {out['gdpt_room_1280x720x32spp']}")
    # the reference's screened Poisson (red-black SOR, 200 sweeps, one thread) at 1280x720
    full = [np.zeros((720, 1280, 3), np.float32) for _ in range(3)]
    rng = np.random.default_rng(7)
    for f in full:
        f[:] = rng.random(f.shape, dtype=np.float32)
    _, secs = orc.poisson_sor(full[0], full[1], full[2], 0.2, iterations=200, omega=1.6, sign=1)
    out["poisson_sor200_1280x720"] = {"value": round(secs * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
                                      "sample": "one solve on random fp32 films (SOR time does not depend on "
                                                "the values)"}
    log(f"cpu: SOR-200 {out['poisson_sor200_1280x720']}")
    return out


def bench_serving(args, ar, torch):
    """Continuous batching against the resident room scene.  Job mix: square
    side U{64..1024 step 64}, spp round(exp(U[0, ln 256])), PT 60% / BDPT 25% /
    GDPT 15%, seed = job index.

    1. Capacity: the first `serve_jobs` jobs of the mix are rendered alone;
       their mean solo time gives the offered load at the specified rate,
       rho = rate x mean solo time.  Above 1 no server keeps up and the
       queue and p99 grow without bound over the run.
    2. Serving: the same jobs arrive as a Poisson stream whose window makes
       the offered load exactly `load` (sum of solo times / window); each
       job's latency (server clock, submit to film ready) over its solo time
       gives p50 / p99 of latency / solo."""
    wl = ar.workloads
    desc = wl.procedural_room()
    sc = ar.Scene.from_desc(desc)
    cd = desc.camera

    def integ(j):
        if j.integrator == 0:
            return ar.PathTracer(ar.PathTracerConfig(spp=j.spp, max_depth=16, rr_depth=3), seed=j.seed)
        if j.integrator == 1:
            return ar.BidirectionalPathTracer(ar.BdptConfig(spp=j.spp, max_depth=10, rr_depth=3), seed=j.seed)
        return ar.GradientDomainPathTracer(ar.GdptConfig(spp=j.spp, max_depth=16, poisson_iterations=50),
                                           seed=j.seed)

    def cam(j):
        return ar.PinholeCamera.from_array(wl.camera_lookat(cd.lookfrom, cd.lookat, cd.vup, cd.vfov,
                                                            j.width / j.height))

    def solo(j):
        it = integ(j)
        it.render(sc, cam(j), j.width, j.height)
        return it.last_stats.seconds

    spec_jobs = wl.serving_jobs(rate=args.rate, duration=args.duration, seed=0)
    solo(spec_jobs[0])  # warm-up (workspace, code objects)
    jobs = spec_jobs[: args.serve_jobs]
    solos = []
    for k, j in enumerate(jobs):
        solos.append(solo(j))
        if k % 20 == 19:
            log(f"serving: solo {k + 1}/{len(jobs)}")
    solos = np.asarray(solos)
    mean_solo = float(solos.mean())
    rho_spec = args.rate * mean_solo
    # arrival window for the target load: sum of solo times / window = load
    window = float(solos.sum()) / args.load
    rate = len(jobs) / window
    rng = np.random.default_rng(1)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(jobs)))
    log(f"serving: {len(spec_jobs)} jobs at {args.rate}/s specified; mean solo {mean_solo:.3f}s -> offered load "
        f"{rho_spec:.2f}; serving the first {len(jobs)} at {rate:.3f}/s (load {args.load})")
    srv = ar.RenderServer(sc, hbm_budget_bytes=200 << 30, max_active_jobs=args.slots,
                          max_concurrent_jobs=args.lanes, pool_paths=args.pool_paths,
                          pool_min_paths=args.pool_min_paths, pool_fill_paths=args.pool_fill,
                          chunk_paths=args.chunk_paths,
                          bdpt_chunk_samples=args.bdpt_chunk)
    warm = [srv.submit(integ(j), cam(j), j.width, j.height) for j in jobs[: args.slots]]
    for w in warm:
        srv.wait(w)
    t0 = time.perf_counter()
    ids, last_log = [], 0.0
    for j, at in zip(jobs, arrivals):
        now = time.perf_counter() - t0
        if at > now:
            time.sleep(at - now)
        ids.append((j, srv.submit(integ(j), cam(j), j.width, j.height)))
        if now - last_log > 10.0:
            log(f"serving: {len(ids)}/{len(jobs)} jobs submitted at t={now:.1f}s")
            last_log = now
    lat, waits, paths, detail = [], [], 0, []
    for k, (j, jid) in enumerate(ids):
        _, st = srv.wait(jid)
        lat.append(st["latency"])
        waits.append(st["start"] - st["submit"])
        paths += st["paths"]
        detail.append([int(j.index), int(jid), int(j.integrator), int(j.width), int(j.spp), round(solos[k] * 1e3, 2),
                       round(st["submit"], 4), round(st["start"], 4), round(st["finish"], 4)])
    wall = time.perf_counter() - t0
    srv.close()
    lat = np.asarray(lat)
    ratio = lat / np.maximum(solos, 1e-6)
    worst = []
    for k in np.argsort(-ratio)[:8]:
        j = jobs[k]
        worst.append({"job": int(j.index), "integrator": ['pt', 'bdpt', 'gdpt'][j.integrator],
                      "size": [int(j.width), int(j.height)], "spp": int(j.spp), "solo_ms": round(solos[k] * 1e3, 2),
                      "latency_ms": round(lat[k] * 1e3, 2), "admission_wait_ms": round(waits[k] * 1e3, 2),
                      "ratio": round(float(ratio[k]), 2)})
        log(f"serving: job {j.index} {['pt', 'bdpt', 'gdpt'][j.integrator]} {j.width}x{j.height} x {j.spp} spp: solo "
            f"{solos[k] * 1e3:.1f} ms, latency {lat[k] * 1e3:.1f} ms (admission wait {waits[k] * 1e3:.1f} ms, "
            f"{ratio[k]:.1f}x)")
    return {"jobs": len(jobs), "rate_per_s": round(rate, 3), "arrival_window_s": round(window, 2),
            "slots": args.slots, "lanes": args.lanes, "pool_paths": args.pool_paths,
            "pool_min_paths": args.pool_min_paths, "pool_fill_paths": args.pool_fill, "chunk_paths": args.chunk_paths, "bdpt_chunk_samples": args.bdpt_chunk, "specified_rate_per_s": args.rate, "specified_jobs": len(spec_jobs),
            "mean_solo_s": round(mean_solo, 4), "offered_load_at_specified_rate": round(rho_spec, 2),
            "offered_load": args.load,
            "job_mix": {"pt": sum(j.integrator == 0 for j in jobs), "bdpt": sum(j.integrator == 1 for j in jobs),
                        "gdpt": sum(j.integrator == 2 for j in jobs)},
            "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
            "p50_latency_over_solo": round(float(np.percentile(ratio, 50)), 3),
            "p99_latency_over_solo": round(float(np.percentile(ratio, 99)), 3),
            "p99_solo_s": round(float(np.percentile(solos, 99)), 4),
            "p99_latency_over_p99_solo": round(float(np.percentile(lat, 99) / np.percentile(solos, 99)), 3),
            "worst_jobs_by_ratio": worst,
            "mpaths_s": round(paths / wall / 1e6, 2), "wall_s": round(wall, 2),
            **({"jobs_detail": detail} if getattr(args, "detail", False) else {})}


def bench_server_alone(args, ar, torch):
    """Each job shape rendered alone with ar_render and submitted alone to a
    server (default configuration: PT / GDPT jobs of at least 2^22 paths in
    the pool, the rest in chunks), best of three each: what serving costs a
    job when nothing else runs."""
    wl = ar.workloads
    desc = wl.procedural_room()
    sc = ar.Scene.from_desc(desc)
    cd = desc.camera

    def integ(kind, spp):
        if kind == 0:
            return ar.PathTracer(ar.PathTracerConfig(spp=spp, max_depth=16, rr_depth=3), seed=1)
        if kind == 1:
            return ar.BidirectionalPathTracer(ar.BdptConfig(spp=spp, max_depth=10, rr_depth=3), seed=1)
        return ar.GradientDomainPathTracer(ar.GdptConfig(spp=spp, max_depth=16, poisson_iterations=50), seed=1)

    def cam(side):
        return ar.PinholeCamera.from_array(wl.camera_lookat(cd.lookfrom, cd.lookat, cd.vup, cd.vfov, 1.0))
    shapes = [(0, 384, 32), (0, 960, 22), (0, 1024, 64), (2, 704, 32), (1, 704, 16), (1, 128, 132)]
    integ(0, 4).render(sc, cam(256), 256, 256)
    srv = ar.RenderServer(sc, hbm_budget_bytes=200 << 30, max_active_jobs=4, pool_paths=args.pool_paths,
                          pool_min_paths=args.pool_min_paths, bdpt_chunk_samples=args.bdpt_chunk)
    srv.wait(srv.submit(integ(0, 4), cam(256), 256, 256))
    out = []
    for kind, side, spp in shapes:
        solo, served = [], []
        for _ in range(3):
            if not args.served_only:
                it = integ(kind, spp)
                it.render(sc, cam(side), side, side)
                solo.append(it.last_stats.seconds)
            _, st = srv.wait(srv.submit(integ(kind, spp), cam(side), side, side))
            served.append(st["latency"])
        row = {"integrator": ["pt", "bdpt", "gdpt"][kind], "side": side, "spp": spp,
               "served_ms": round(min(served) * 1e3, 2)}
        if solo:
            row.update(solo_ms=round(min(solo) * 1e3, 2), ratio=round(min(served) / min(solo), 3))
        log(f"server alone: {row}")
        out.append(row)
    srv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="traversal", choices=["traversal", "serving", "server-alone", "pt", "render",
                                                          "copy", "pmc-probe", "builders"],
                    help="copy: a plain 1 GiB device copy, the reference point that scales rocprof's "
                         "FETCH_SIZE / WRITE_SIZE counters to bytes")
    ap.add_argument("--trace-config", default=None,
                    help="WORKGROUP,LDS_STACK,BLOCKS_PER_CU[,REFILL_MIN[,SEARCH_MIN[,PARK_SLOTS[,BALANCE]]]]: "
                         "fixed traversal launch instead of autotuning")
    ap.add_argument("--anyhit-config", default=None,
                    help="W,S,P[,R[,SM[,PS[,BAL]]]]: any-hit launch (with --trace-config; default: the same as "
                         "closest hit)")
    ap.add_argument("--integrator", default="pt", choices=sorted(ROOM_JOBS),
                    help="render workload: the integrator rendered on the room")
    ap.add_argument("--chunk", type=int, default=0, help="render workload: samples per pixel per queue chunk")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the path-tracer throughput fields")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter runs for roofline.traffic")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg of the default run")
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--duration", type=float, default=60.0, help="specified arrival window (s)")
    ap.add_argument("--load", type=float, default=0.7, help="serving: target offered load when the rate is above it")
    ap.add_argument("--serve-jobs", type=int, default=60, help="serving: jobs of the specified mix actually served")
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--lanes", type=int, default=0, help="serving: jobs on the device at once (0 = the server's default)")
    ap.add_argument("--pool-paths", type=lambda v: int(v, 0), default=0,
                    help="serving: paths in flight in the shared path-tracer queues (0 = default, 0xFFFFFFFF = off)")
    ap.add_argument("--served-only", action="store_true", help="server-alone: skip the ar_render renders")
    ap.add_argument("--detail", action="store_true",
                    help="serving: per-job [index, server id, integrator, side, spp, solo ms, submit, start, finish]")
    ap.add_argument("--pool-fill", type=lambda v: int(v, 0), default=0,
                    help="serving: live paths of the other pooled jobs while one drains (0 = default)")
    ap.add_argument("--chunk-paths", type=lambda v: int(v, 0), default=0,
                    help="serving: PT / GDPT paths per chunk of a job outside the pool (0 = default)")
    ap.add_argument("--bdpt-chunk", type=lambda v: int(v, 0), default=0,
                    help="serving: BDPT samples per chunk (0 = default)")
    ap.add_argument("--pool-min-paths", type=lambda v: int(v, 0), default=0,
                    help="serving: PT / GDPT jobs of at least this many paths share the queues (0 = default)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; sharding over several GPUs is designed "
                                   "(sample ranges or pixel tiles, one ordered film reduction) but not launched", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    ar = load_pkg()
    ar.init(0)
    if args.workload == "pt":
        res = {"pt_cornell_256x256x16spp": bench_pt(ar, torch, "cornell"),
               "pt_room_1920x1080x64spp": bench_pt(ar, torch, "room")}
        res.update(bench_bdpt_gdpt(ar, torch))
        print(json.dumps(res), flush=True)
        return
    if args.workload == "render":
        print(json.dumps(bench_render_room(args, ar, torch)), flush=True)
        return
    if args.workload == "builders":
        print(json.dumps(bench_builders(ar, torch)), flush=True)
        return
    if args.workload == "copy":
        print(json.dumps(bench_copy(args, torch)), flush=True)
        return
    if args.workload == "pmc-probe":
        pmc_probe(args, ar, torch)
        return
    if args.workload == "server-alone":
        print(json.dumps({"server_alone": bench_server_alone(args, ar, torch)}), flush=True)
        return
    if args.workload == "serving":
        srv = bench_serving(args, ar, torch)
        out = {"metric": METRIC, "value": srv["p99_latency_s"], "unit": "s (p99 job latency)", "n_gpus": 1,
               "steps": srv["jobs"], "warmup": args.slots, "ms_per_step": None, "higher_is_better": False,
               "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
               "config": {"workload": "serving_room_280k_tris", "parallelism": "single-gpu"}, "serving": srv}
        print(json.dumps(out), flush=True)
        return
    out = bench_traversal(args, ar, torch)
    if not args.no_extra:
        out["extra"] = {"pt_cornell_256x256x16spp": bench_pt(ar, torch, "cornell"),
                        "pt_room_1920x1080x64spp": bench_pt(ar, torch, "room")}
        out["extra"].update(bench_bdpt_gdpt(ar, torch))
        out["extra"]["tree_builders_soup1M"] = bench_builders(ar, torch)
        if not args.no_cpu_baseline:
            out["extra"]["cpu_baselines"] = bench_cpu_and_quality(ar, torch)
        if not args.no_serving:
            sargs = argparse.Namespace(**vars(args))
            sargs.load, sargs.serve_jobs = 0.5, 60
            out["extra"]["serving_room"] = bench_serving(sargs, ar, torch)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
