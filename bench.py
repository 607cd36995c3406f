"""Benchmark of the ChIMES design-matrix / fit / force-evaluation path on one GPU.

Workload (one "step"): assembly of the full design matrix A (force, stress and
energy rows) of 100 frames of 512-atom carbon at 2.2 g/cc with 2+3+4-body
Chebyshev orders 16/9/5 (A = 154200 rows x 1038 columns, FP64), synthetic
seeded inputs.  `value` is A rows per second for that whole batch, timed around
`cg_assemble` (frames handed over, neighbour build, cluster enumeration,
2/3/4-body assembly, A resident in HBM afterwards); after the warm-up calls the
context replays the assembly from a captured HIP graph.  The 4-body kernel's
time for the roofline comes from the same number of eager steps (HIP events
around its launch).

Also reported: the FP32 assembly rate; the fit time of the FP64 path, of the
FP32-A path (FP64-product Gram + FP64 refinement) and of the split-BF16x3 Gram
variant, with their force-RMS agreement; force-evaluation latency p50/p99 of
1k-atom C/H/N/O configurations (orders 12/6/3), one call at a time and as a
Poisson stream of 64-4096-atom requests through the batching server; the
4-body kernel's roofline position (HIP events on the context stream; HBM
traffic from the committed counter summary); and the CPU oracle timed on the
host cores on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--frames F] [--no-cpu]
    python bench.py --counters     (one assembly + calibration copy, for rocprofv3 --pmc)
"""
import argparse
import ctypes
import importlib.util
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "chimes_lsq-myllfork_amd")
METRIC = "A-matrix rows/s + fit time at FP64-oracle tol vs plain FP64; force-eval p50/p99"
PEAK_FP64_TFLOPS = 78.6     # MI355X dense FP64 (vector and matrix)
COPY_PROBE_BYTES = 1 << 30  # calibration copy of the counter passes
COUNTERS = os.path.join(ROOT, "profiles", "counters_k_assemble_4b_carbon_100x512.md")


def measured_traffic():
    """HBM bytes per launch of the 4-body kernel from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by the calibration copy), or None."""
    try:
        for line in open(COUNTERS):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


GRAPHS = True  # assembly graph capture (bench.py --no-graphs turns it off)


def new_device(ch, setup):
    dev = ch.Device(setup)
    if not GRAPHS:
        dev.assembly_graphs(False)
    return dev


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_frames(ch, frames):
    return [ch.PyFrame(f["natoms"], f["xyz"], f["types"], f["cell"], f["forces"], f["stress"], f["energy"],
                       f["index"]) for f in frames]


def setup_from(ch, text):
    d = tempfile.mkdtemp(prefix="chimes_bench_")
    p = os.path.join(d, "fm_setup.in")
    open(p, "w").write(text)
    return ch.Setup(p)


def cpu_baseline(ch, setup, frames, seconds_hint=25.0):
    """Time the oracle (CPU FP64 restatement) on the host cores: one frame per OpenMP thread
    on every core this job is given (OMP_NUM_THREADS, else the machine's count, at most 16),
    then one frame on a single core."""
    orc = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    orc.orc_assemble_many.restype = ctypes.c_longlong
    orc.orc_assemble_many.argtypes = [ctypes.POINTER(ch.Model), ctypes.POINTER(ch.Frame), ctypes.c_int, ctypes.c_int]
    nproc = os.cpu_count() or 1
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, nproc)
    cores = max(1, min(cores, 16, len(frames)))

    def timed(nf, threads):
        cf = (ch.Frame * nf)(*[f.c() for f in frames[:nf]])
        t0 = time.perf_counter()
        rows = orc.orc_assemble_many(setup.model_ptr, cf, nf, threads)
        return rows, time.perf_counter() - t0

    rows, dt = timed(cores, cores)  # one frame per thread: ~12 s of work per frame per core
    rows1, dt1 = timed(1, 1)
    return {"value": rows / dt, "unit": "A rows/s", "cores": cores, "kind": "port",
            "sample": "%d of the benchmark's 512-atom carbon frames (%d rows), one frame per OpenMP thread, "
                      "%.1f s" % (cores, rows, dt),
            "nproc": nproc,
            "one_core": {"value": rows1 / dt1, "unit": "A rows/s",
                         "sample": "1 frame (%d rows) on one core, %.1f s" % (rows1, dt1)}}


def serving_latency(ch, synth, n_req=400, n_conf=16, natoms=1000, seed=105):
    """Force/energy/stress latency of 1k-atom C/H/N/O configurations, one call at a
    time (n_conf configurations of one shape, cycled over n_req calls)."""
    text, frames, _ = synth.chno(nframes=n_conf, nmin=natoms, nmax=natoms, seed=seed, target_density=0.1)
    s = setup_from(ch, text)
    fr = make_frames(ch, frames)
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1e-3, s.model.ncols)
    dev = ch.Device(s)
    for f in fr[:3]:
        dev.force(x, f)
    lat = []
    t0 = time.perf_counter()
    for i in range(n_req):
        t1 = time.perf_counter()
        dev.force(x, fr[i % n_conf])
        lat.append((time.perf_counter() - t1) * 1e3)
    tot = time.perf_counter() - t0
    replays = dev.force_stats()[2]
    dev.close()
    lat = np.array(lat)
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "atoms_per_s": natoms * n_req / tot, "requests": n_req, "configurations": n_conf,
            "graph_replays": replays, "natoms": natoms, "orders": "12/6/3", "ncols": int(s.model.ncols)}


def md_steps(ch, synth, n_steps=60, natoms=1000, seed=107):
    """MD-style force calls: one 1k-atom C/H/N/O configuration (fixed cell and atom
    count) with atoms moved between calls.  The context captures the call into a HIP
    graph and replays it; the same steps on a second context that alternates with a
    different shape stay eager, for comparison."""
    text, frames, _ = synth.chno(nframes=1, nmin=natoms, nmax=natoms, seed=seed, target_density=0.1)
    s = setup_from(ch, text)
    f0 = make_frames(ch, frames)[0]
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1e-3, s.model.ncols)
    other = ch.PyFrame(64, synth.random_sequential(rng, 64, np.diag([12.0] * 3), 0.9),
                       rng.integers(0, 4, 64).astype(np.int32), np.diag([12.0] * 3))
    xyz = f0.xyz.reshape(-1, 3)
    steps = [ch.PyFrame(natoms, xyz + rng.normal(0, 0.01, xyz.shape), f0.types, f0.cell) for _ in range(n_steps)]
    out = {"natoms": natoms, "steps": n_steps, "orders": "12/6/3"}
    for tag, alternate in (("graph", False), ("eager", True)):
        dev = ch.Device(s)
        lat = []
        for i, f in enumerate(steps):
            t1 = time.perf_counter()
            dev.force(x, f)
            if i >= 5:
                lat.append((time.perf_counter() - t1) * 1e3)
            if alternate:
                dev.force(x, other)
        out[tag + "_p50_ms"] = float(np.percentile(lat, 50))
        out[tag + "_p99_ms"] = float(np.percentile(lat, 99))
        out[tag + "_replays"] = dev.force_stats()[2]
        dev.close()
    return out


def serving_sweep(ch, synth, qps_points=(250, 500, 1000, 1500, 2000), n_req=1000, workers=4, seed=106):
    """Poisson streams of C/H/N/O configurations through the continuous-batching
    server, 1000 requests per offered rate.  Three requests in four are drawn from
    200 configurations of 64-4096 atoms (log-uniform), every fourth is one of 16
    1000-atom configurations; latency is measured in the library from submit to
    completion and reported per size class."""
    text, frames, _ = synth.chno(nframes=200, nmin=64, nmax=4096, seed=seed, target_density=0.1)
    _, frames1k, _ = synth.chno(nframes=16, nmin=1000, nmax=1000, seed=seed + 100, target_density=0.1)
    s = setup_from(ch, text)
    mixed = make_frames(ch, frames)
    onek = make_frames(ch, frames1k)
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1e-3, s.model.ncols)
    srv = ch.Server(s, x, workers=workers)
    # warm-up: every configuration once, so that device buffers have reached
    # their working sizes before the timed streams
    for t in [srv.submit(f) for f in mixed + onek]:
        srv.poll(t)
    classes = (("64-255", 64, 256), ("256-1023", 256, 1024), ("1024-4096", 1024, 4097))
    points = []
    for qps in qps_points:
        reqs = [onek[(i // 4) % len(onek)] if i % 4 == 3 else mixed[int(rng.integers(len(mixed)))]
                for i in range(n_req)]
        arrivals = np.cumsum(rng.exponential(1.0 / qps, n_req))
        c0 = srv.counters()
        tickets = []
        t0 = time.perf_counter()
        for i, f in enumerate(reqs):
            dt = t0 + arrivals[i] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            tickets.append(srv.submit(f))
        lat = np.array([srv.poll(t)[3] for t in tickets])
        wall = time.perf_counter() - t0
        c1 = srv.counters()
        ev = {k: c1[k] - c0[k] for k in c0}
        n = np.array([f.natoms for f in reqs])
        is1k = np.arange(n_req) % 4 == 3
        pt = {"qps_offered": qps, "qps_served": n_req / wall, "requests": n_req,
              "atoms_per_s": float(n.sum() / wall), "p50_ms": float(np.percentile(lat, 50)),
              "p99_ms": float(np.percentile(lat, 99)), "mean_batch": ev["requests"] / max(1, ev["batches"]),
              "events": ev,
              "drain_ms": (wall - arrivals[-1]) * 1e3,
              "p99_ms_by_atoms": {}, "one_k": {"p50_ms": float(np.percentile(lat[is1k], 50)),
                                               "p99_ms": float(np.percentile(lat[is1k], 99)),
                                               "requests": int(is1k.sum())}}
        for name, lo, hi in classes:
            sel = (~is1k) & (n >= lo) & (n < hi)
            if sel.any():
                pt["p99_ms_by_atoms"][name] = float(np.percentile(lat[sel], 99))
        log("serving %5.0f req/s offered: %.0f served, p50 %.2f ms, p99 %.2f ms, 1k-atom p99 %.2f ms, "
            "mean batch %.2f, events %s" % (qps, pt["qps_served"], pt["p50_ms"], pt["p99_ms"], pt["one_k"]["p99_ms"],
                                            pt["mean_batch"], ev))
        points.append(pt)
    srv.close()
    # sustainable: the highest offered rate whose queue drained within 50 ms of
    # the last arrival and whose 1000-atom p99 stayed below 5 ms
    ok = [p for p in points if p["drain_ms"] < 50.0 and p["one_k"]["p99_ms"] < 5.0]
    return {"workers": workers, "natoms": "64-4096 log-uniform + 1000 (every 4th)", "orders": "12/6/3",
            "sustainable_qps": max((p["qps_offered"] for p in ok), default=None), "points": points}


def assemble_rate(ch, dev, fr, rows, prec, steps=2, warmup=3):
    """Rows/s of repeated assemblies of one batch: after the warm-up calls (exact
    list sizes, then capacity-sized, then captured) the timed calls replay the
    captured graph."""
    for _ in range(warmup):
        dev.assemble(fr, prec=prec, copy=False)
    dev.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dev.assemble(fr, prec=prec, copy=False)
    dev.synchronize()
    return rows / ((time.perf_counter() - t0) / steps)


def family_ms(dev, fr, prec):
    """Per-family kernel times of one eager assembly (a replayed graph carries no
    per-launch events)."""
    dev.assembly_graphs(False)
    dev.assemble(fr, prec=prec, copy=False)
    ms, _ = dev.detail()
    dev.assembly_graphs(GRAPHS)
    return dict(zip(["2B", "3B", "4B"], ms))


def config_c3(ch, synth, nframes=600):
    """Precision form of the carbon workload at 600 frames: FP32 A + FP64-product
    Gram + FP64 refinement against the plain FP64 path (assembly and fit)."""
    text, frames, _ = synth.carbon(nframes=nframes, natoms=512, seed=103)
    s = setup_from(ch, text)
    fr = make_frames(ch, frames)
    rows = sum(s.frame_rows(f.natoms, f.index) for f in fr)
    dev = new_device(ch, s)
    out = {"workload": "carbon %d frames x 512 atoms, orders 16/9/5" % nframes, "rows": rows,
           "cols": int(s.model.ncols)}
    for prec, tag in [(ch.CG_FP64, "fp64"), (ch.CG_FP32, "fp32")]:
        rate = assemble_rate(ch, dev, fr, rows, prec, steps=1)
        _, b = dev.assemble(fr, prec=prec, copy=False)
        dev.fit(b, eps=1e-5, refine=0)
        t0 = time.perf_counter()
        x, r = dev.fit(b, eps=1e-5, refine=4)
        tf = time.perf_counter() - t0
        out[tag] = {"assembly_rows_per_s": rate, "fit_s": tf, "rms": r.rms, "nvars": r.nvars,
                    "assembly_plus_fit_rows_per_s": rows / (rows / rate + tf), "x": x}
    # split-BF16x3 Gram on the FP32 A (FP64 refinement guarded against growth)
    t0 = time.perf_counter()
    xb, rb = dev.fit(b, eps=1e-5, refine=4, gram=ch.CG_GRAM_BF16X3)
    tb = time.perf_counter() - t0
    out["fp32_bf16x3_gram"] = {"fit_s": tb, "gram_ms": rb.ms_gram, "rms": rb.rms, "nvars": rb.nvars,
                               "iters": rb.iters,
                               "assembly_plus_fit_rows_per_s": rows / (rows / out["fp32"]["assembly_rows_per_s"] + tb)}
    x64, x32 = out["fp64"].pop("x"), out["fp32"].pop("x")
    for tag, x in (("fp32", x32), ("fp32_bf16x3_gram", xb)):
        out[tag]["rms_rel_err_vs_fp64"] = abs(out[tag]["rms"] - out["fp64"]["rms"]) / out["fp64"]["rms"]
        out[tag]["coef_rel_err_vs_fp64"] = float(np.linalg.norm(x - x64) / np.linalg.norm(x64))
        out[tag]["meets_1e-6_rms_target"] = bool(out[tag]["rms_rel_err_vs_fp64"] <= 1e-6)
    out["fp32_over_fp64_rows_per_s"] = (out["fp32"]["assembly_plus_fit_rows_per_s"] /
                                         out["fp64"]["assembly_plus_fit_rows_per_s"])
    out["fp32_bf16x3_over_fp64_rows_per_s"] = (out["fp32_bf16x3_gram"]["assembly_plus_fit_rows_per_s"] /
                                                out["fp64"]["assembly_plus_fit_rows_per_s"])
    dev.close()
    return out


def config_c4(ch, synth, nframes=256):
    """Batched assembly of frames of uneven size and species in one call: C/H/N/O,
    32-2048 atoms (log-uniform), triclinic cells, orders 12/6/3."""
    text, frames, _ = synth.chno(nframes=nframes, nmin=32, nmax=2048, seed=104)
    s = setup_from(ch, text)
    fr = make_frames(ch, frames)
    rows = sum(s.frame_rows(f.natoms, f.index) for f in fr)
    dev = new_device(ch, s)
    out = {"workload": "C/H/N/O %d frames, 32-2048 atoms, orders 12/6/3" % nframes, "rows": rows,
           "cols": int(s.model.ncols), "atoms": sum(f.natoms for f in fr)}
    out["fp64_rows_per_s_default_tiles"] = assemble_rate(ch, dev, fr, rows, ch.CG_FP64, steps=2)
    choice, ms = dev.autotune(fr)
    cfgs, classes = ch.Device.TILE_CONFIGS, ch.Device.TILE_CLASSES
    out["autotune"] = {
        "choice": {fam: {classes[q]: cfgs[choice[w][q]] for q in range(len(classes)) if choice[w][q] >= 0}
                   for w, fam in ((1, "3B"), (2, "4B"))},
        "kernel_ms": {cfgs[k]: {fam: {classes[q]: float(ms[k][w][q]) for q in range(len(classes)) if choice[w][q] >= 0}
                                for w, fam in ((1, "3B"), (2, "4B"))} for k in range(len(cfgs))}}
    for prec, tag in [(ch.CG_FP64, "fp64_rows_per_s"), (ch.CG_FP32, "fp32_rows_per_s")]:
        out[tag] = assemble_rate(ch, dev, fr, rows, prec, steps=2)
        out[tag.replace("rows_per_s", "family_ms")] = family_ms(dev, fr, prec)
    out["clusters"] = dict(zip(["pairs", "triplets", "quadruplets"], dev.detail()[1]))
    dev.assembly_deterministic(True)
    out["fp64_deterministic_rows_per_s"] = assemble_rate(ch, dev, fr, rows, ch.CG_FP64, steps=2)
    out["fp64_deterministic_family_ms"] = family_ms(dev, fr, ch.CG_FP64)
    dev.assembly_deterministic(False)
    calls, repeats, replays, captures = dev.assembly_stats()
    out["assembly_calls"] = {"calls": calls, "exact_repeats": repeats, "graph_replays": replays,
                             "graph_captures": captures}
    dev.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fit", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-md", action="store_true",
                    help="skip the MD-style leg (force calls replayed from HIP graphs), e.g. under a kernel tracer")
    ap.add_argument("--no-configs", action="store_true", help="skip the 600-frame and mixed-batch legs")
    ap.add_argument("--no-graphs", action="store_true",
                    help="keep every assembly eager (no graph capture), e.g. under a kernel tracer")
    ap.add_argument("--counters", nargs="?", const="carbon", choices=["carbon", "chno"],
                    help="one FP64 assembly of the carbon workload (or of the 256-frame mixed C/H/N/O batch) "
                         "plus a 1 GiB device copy, for rocprofv3 --pmc runs")
    args = ap.parse_args()
    global GRAPHS
    GRAPHS = not args.no_graphs
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the design-matrix path is single-device "
              "(sharding frames over GPUs is designed but not built)", file=sys.stderr)
        sys.exit(2)
    ch = _load("chimes", os.path.join(PKG, "chimes.py"))
    synth = _load("synth", os.path.join(PKG, "synth.py"))
    if args.counters == "chno":
        text, frames, _ = synth.chno(nframes=256, nmin=32, nmax=2048, seed=104)
        setup = setup_from(ch, text)
        dev = new_device(ch, setup)
        dev.assemble(make_frames(ch, frames), prec=ch.CG_FP64, copy=False)
        dev.copy_probe(COPY_PROBE_BYTES)
        dev.close()
        log("counter pass done: mixed-batch assembly + %d-byte copy" % COPY_PROBE_BYTES)
        return
    log("generating %d carbon frames" % args.frames)
    text, frames, _ = synth.carbon(nframes=args.frames, natoms=512, seed=102)
    setup = setup_from(ch, text)
    fr = make_frames(ch, frames)
    rows = sum(setup.frame_rows(f.natoms, f.index) for f in fr)
    dev = new_device(ch, setup)
    if args.counters:
        dev.assemble(fr, prec=ch.CG_FP64, copy=False)
        dev.copy_probe(COPY_PROBE_BYTES)
        dev.close()
        log("counter pass done: assembly + %d-byte copy" % COPY_PROBE_BYTES)
        return

    def step(prec):
        dev.assemble(fr, prec=prec, copy=False)

    # warm-up: the first call sizes its lists exactly, the next ones from the
    # per-atom counts; a repeated shape is captured into a graph and replayed
    for _ in range(max(args.warmup, 3)):
        step(ch.CG_FP64)
    dev.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(ch.CG_FP64)
    dev.synchronize()
    dt = time.perf_counter() - t0
    ms_step = dt / args.steps * 1e3
    value = rows / (dt / args.steps)
    calls, repeats, replays, captures = dev.assembly_stats()
    # the 4-body kernel's launch time: HIP events on the context stream around
    # its launch in the same number of eager steps (graphs off)
    dev.assembly_graphs(False)
    k4 = []
    asm = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(ch.CG_FP64)
        ms, ncl = dev.detail()
        k4.append(ms[2])
        asm.append(dev.timing())
    fam_eager = list(ms)  # per-family kernel times of the last eager step
    dev.synchronize()
    ms_step_eager = (time.perf_counter() - t0) / args.steps * 1e3
    # bitwise-reproducible mode (waves add to the LDS tile in turn): its cost
    dev.assembly_deterministic(True)
    k4d = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(ch.CG_FP64)
        k4d.append(dev.detail()[0][2])
    dev.synchronize()
    ms_step_det = (time.perf_counter() - t0) / args.steps * 1e3
    dev.assembly_deterministic(False)
    dev.assembly_graphs(GRAPHS)
    ms4 = float(np.mean(k4))
    # algorithmic work of the 4-body kernel: 4 flops per pair per allowed power set
    # (product, prefix/suffix force coefficient, accumulation), 6 pairs per cluster
    nallowed = setup.model.c4.total_allowed
    flops4 = 24.0 * ncl[2] * nallowed
    achieved = flops4 / (ms4 * 1e-3) / 1e12
    log("FP64: %.3f ms/step (%d graph replays), %.4e rows/s; eager %.3f ms/step; 4-body kernel %.3f ms "
        "(%d quads), %.2f TFLOP/s" % (ms_step, replays, value, ms_step_eager, ms4, ncl[2], achieved))
    # FP32 assembly (its own graph after the warm-up)
    for _ in range(max(args.warmup, 3)):
        step(ch.CG_FP32)
    dev.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(ch.CG_FP32)
    dev.synchronize()
    value32 = rows / ((time.perf_counter() - t0) / args.steps)
    log("FP32: %.4e rows/s" % value32)
    extra = {"fp32_rows_per_s": value32, "kernel_ms": {"enum": asm[-1][0], "assembly": asm[-1][1]},
             "assembly_calls": {"graph_replays": replays, "graph_captures": captures, "exact_repeats": repeats,
                                "calls": calls},
             "eager_ms_per_step": ms_step_eager, "eager_rows_per_s": rows / (ms_step_eager * 1e-3),
             "deterministic": {"ms_per_step": ms_step_det, "rows_per_s": rows / (ms_step_det * 1e-3),
                               "kernel_4b_ms": float(np.mean(k4d))},
             "family_ms": dict(zip(["2B", "3B", "4B"], fam_eager)),
             "clusters": dict(zip(["pairs", "triplets", "quadruplets"], ncl))}
    if not args.no_fit:
        _, b = dev.assemble(fr, prec=ch.CG_FP64, copy=False)
        dev.fit(b, eps=1e-5, refine=0)  # warm-up: solver workspaces and kernels
        t0 = time.perf_counter()
        x64, r64 = dev.fit(b, eps=1e-5, refine=3)
        t_fit64 = time.perf_counter() - t0
        t_asm64 = rows / value
        dev.assemble(fr, prec=ch.CG_FP32, copy=False)
        t0 = time.perf_counter()
        x32, r32 = dev.fit(b, eps=1e-5, refine=4)
        t_fit32 = time.perf_counter() - t0
        t_asm32 = rows / value32
        t0 = time.perf_counter()
        xb, rb = dev.fit(b, eps=1e-5, refine=4, gram=ch.CG_GRAM_BF16X3)
        t_fitb = time.perf_counter() - t0
        extra["fit"] = {
            "fp64": {"fit_s": t_fit64, "gram_ms": r64.ms_gram, "solve_ms": r64.ms_solve, "refine_ms": r64.ms_refine,
                     "rms": r64.rms, "nvars": r64.nvars, "assembly_plus_fit_rows_per_s": rows / (t_asm64 + t_fit64)},
            "fp32_A_fp64_gram_refined": {
                "fit_s": t_fit32, "gram_ms": r32.ms_gram, "solve_ms": r32.ms_solve, "refine_ms": r32.ms_refine,
                "rms": r32.rms, "nvars": r32.nvars, "rms_rel_err_vs_fp64": abs(r32.rms - r64.rms) / r64.rms,
                "coef_rel_err_vs_fp64": float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64)),
                "assembly_plus_fit_rows_per_s": rows / (t_asm32 + t_fit32)},
            "fp32_A_bf16x3_gram_refined": {
                "fit_s": t_fitb, "gram_ms": rb.ms_gram, "rms": rb.rms, "nvars": rb.nvars, "iters": rb.iters,
                "rms_rel_err_vs_fp64": abs(rb.rms - r64.rms) / r64.rms},
        }
        log("fit FP64 %.3f s rms %.6e | FP32 A + FP64 Gram + refine %.3f s rms %.6e | split-BF16x3 Gram rms %.6e" % (
            t_fit64, r64.rms, t_fit32, r32.rms, rb.rms))
    dev.close()
    if not args.no_serve:
        one = serving_latency(ch, synth)
        log("serving 1k atoms, one call at a time: p50 %.3f ms p99 %.3f ms" % (one["p50_ms"], one["p99_ms"]))
        sweep = serving_sweep(ch, synth)
        log("serving sweep: sustainable %s req/s" % sweep["sustainable_qps"])
        extra["serving"] = {"single_1k_atoms": one, "stream_sweep": sweep}
        if not args.no_md:
            md = md_steps(ch, synth)
            log("MD-style steps, 1k atoms: graph replay p50 %.3f ms p99 %.3f ms (%d replays) | eager p50 %.3f ms "
                "p99 %.3f ms" % (md["graph_p50_ms"], md["graph_p99_ms"], md["graph_replays"], md["eager_p50_ms"],
                                 md["eager_p99_ms"]))
            extra["serving"]["md_steps_1k_atoms"] = md
    if not args.no_configs:
        c3 = config_c3(ch, synth)
        log("600-frame carbon: FP64 %.3g rows/s + fit %.3f s | FP32 %.3g rows/s + fit %.3f s, RMS rel err %.2e" % (
            c3["fp64"]["assembly_rows_per_s"], c3["fp64"]["fit_s"], c3["fp32"]["assembly_rows_per_s"],
            c3["fp32"]["fit_s"], c3["fp32"]["rms_rel_err_vs_fp64"]))
        c4 = config_c4(ch, synth)
        log("mixed C/H/N/O batch (%d atoms, %d rows x %d cols): FP64 %.3g rows/s (default tiles %.3g), FP32 %.3g "
            "rows/s; autotuned %s" % (c4["atoms"], c4["rows"], c4["cols"], c4["fp64_rows_per_s"],
                                      c4["fp64_rows_per_s_default_tiles"], c4["fp32_rows_per_s"],
                                      c4["autotune"]["choice"]))
        extra["carbon_600_frames_precision"] = c3
        extra["chno_mixed_batch"] = c4
    cpu = None
    if not args.no_cpu:
        log("timing the CPU oracle on the host")
        cpu = cpu_baseline(ch, setup, fr)
        log("CPU: %.1f rows/s on %d cores (machine count %d), %.1f rows/s on one core" % (
            cpu["value"], cpu["cores"], cpu["nproc"], cpu["one_core"]["value"]))
    out = {
        "metric": METRIC, "value": value, "unit": "A rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded carbon, uniform random positions, "
        "1.1 A minimum separation; Gaussian forces/stress/energy)",
        "config": {"workload": "carbon 2+3+4-body design-matrix assembly, %d frames x 512 atoms, 2.2 g/cc, "
                               "orders 16/9/5, cutoffs 4.5/4.0/3.5 A, %d rows x %d cols" % (
                                   args.frames, rows, setup.model.ncols),
                   "frames": args.frames, "natoms": 512, "rows": rows, "cols": int(setup.model.ncols)},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": measured_traffic(),
                     "pipe": "FP64 VALU (same 78.6 TFLOP/s dense peak as FP64 MFMA)",
                     "kernel": "k_assemble<4,5,0,double> (4-body rows)", "kernel_ms": ms4,
                     "algorithmic_flops": flops4},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
