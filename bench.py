#!/usr/bin/env python3
"""Throughput of the fused FiVoNAGI time step on one MI355X.

Workload: the HIFU-beam-like single instance -- 4096x1024 FP32 grid, quiescent
medium driven from the left edge (rho' = A g(j) sin(omega T), A = 0.02,
Gaussian aperture of half-width 200 cells, 20 cells per wavelength),
beta = 3.5, delta = 5e-5, CFL 0.9, zero-gradient on the other edges.
A "step" is one step attempt of the whole grid (one launch of the fused
kernel); value = accepted cell-updates per second over the timed steps.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Prints one JSON line.  Timing uses HIP events recorded on the stream the step
kernel runs on (created here through the HIP runtime, handed to the library).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s + % HBM peak (fwd FP32); batched fwd+adjoint gradients/s vs plain"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
BYTES_PER_CELL = 24            # read q (3 x 4 B) + write q (3 x 4 B)
PEAK_CLOCK_HZ = 2.4e9          # MI355X peak engine clock
KERNEL = "fv_step_kernel"


class Hip:
    """Minimal HIP runtime calls through ctypes: streams and events."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        for n in ("hipStreamCreate", "hipEventCreate", "hipEventRecord", "hipEventSynchronize",
                  "hipEventElapsedTime", "hipStreamSynchronize", "hipEventDestroy",
                  "hipStreamDestroy", "hipGetDeviceCount"):
            getattr(self.lib, n).restype = C.c_int

    def check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with hipError {rc}")

    def device_count(self):
        n = C.c_int()
        self.check(self.lib.hipGetDeviceCount(C.byref(n)), "hipGetDeviceCount")
        return n.value

    def stream(self):
        s = C.c_void_p()
        self.check(self.lib.hipStreamCreate(C.byref(s)), "hipStreamCreate")
        return s

    def event(self):
        e = C.c_void_p()
        self.check(self.lib.hipEventCreate(C.byref(e)), "hipEventCreate")
        return e

    def record(self, e, s):
        self.check(self.lib.hipEventRecord(e, s), "hipEventRecord")

    def elapsed_ms(self, e0, e1):
        self.check(self.lib.hipEventSynchronize(e1), "hipEventSynchronize")
        ms = C.c_float()
        self.check(self.lib.hipEventElapsedTime(C.byref(ms), e0, e1), "hipEventElapsedTime")
        return ms.value


def warm_up(ctx, warmup, steps, min_seconds=0.25):
    """Untimed: the warmup attempts, then the timed pattern (its graph
    replays are captured here, not inside the events), each followed by
    attempts up to a phase-0 boundary of the six-launch rotation, so the
    timed attempts replay whole captured graphs from their first launch.
    The timed pattern is repeated until min_seconds of work have run, so a
    short timed region (a few steps, well under a millisecond) starts with
    the GPU at its working clock rather than ramping up from idle."""
    t0 = time.perf_counter()
    for k in range(1000):
        n = warmup if k == 0 else steps
        ctx.enqueue(n)
        if ctx.sync() != 0:
            raise RuntimeError("warmup failed")
        ctx.enqueue((-ctx.info(0)["phase"]) % 6)
        if k >= 1 and time.perf_counter() - t0 >= min_seconds:
            break
    if ctx.sync() != 0:
        raise RuntimeError("warmup failed")


def profile_record(workload_tag):
    """The kernel's measured HBM bytes and VALU figures per launch, from the
    rocprofv3 counter runs recorded under profiles/ ({} if not measured)."""
    path = os.path.join(ROOT, "profiles", f"{KERNEL}_{workload_tag}_traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def traffic_from_profiles(workload_tag):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE,
    each scaled by the counter's reading for a copy of known size), recorded
    under profiles/; None if not measured."""
    path = os.path.join(ROOT, "profiles", f"{KERNEL}_{workload_tag}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def native_oracle():
    """The oracle's sources built for this host's CPU (gcc -O3 -march=native
    -fopenmp) in a temporary directory, so the CPU baseline uses every vector
    extension of the machine it runs on; (path, flags)."""
    import subprocess
    import tempfile
    src = os.path.join(ROOT, "oracle")
    out = os.path.join(tempfile.mkdtemp(prefix="fv_oracle_native_"), "liboracle_native.so")
    flags = ["-std=c11", "-O3", "-march=native", "-fPIC", "-shared", "-fopenmp",
             "-ffp-contract=off"]
    subprocess.run(["gcc", *flags, "-o", out, os.path.join(src, "fvoracle.c"),
                    os.path.join(src, "fvadjoint.c"), "-lm"], check=True)
    return out, " ".join(flags)


def cpu_threads():
    """Every CPU of this host's share: OMP_NUM_THREADS where the host sets
    it (the CPU share of one GPU on a shared GPU host), else every CPU in
    this process's affinity mask."""
    visible = len(os.sched_getaffinity(0))
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return (min(env, visible) if env > 0 else visible), visible


def cpu_baseline(prob, seconds=15.0):
    """The FP32 build of the CPU oracle (a port of the reference step) on this
    host's cores, built for this host, over a bounded number of steps of the
    same grid; and the same on one core."""
    from oracle import pyoracle as po
    lib, flags = native_oracle()
    threads, visible = cpu_threads()
    op = po.Problem(prob.nx, prob.ny, prob.dx, prob.dy, prob.beta, prob.delta, prob.cfl_max,
                    prob.cfl_wish, prob.dt0, prob.T0, bc=prob.bc, cfl_mode=prob.cfl_mode,
                    form=po.FORM_RHO1, omega=prob.omega, edge_amp=prob.edge_amp,
                    nthreads=threads)
    q = prob.q0.astype(np.float32)
    # warm up (threads, first touch), calibrate on a few more steps, then time
    # one bounded run of the same grid
    q, st, rc = po.run(op, q, 3, dtype=np.float32, fast=lib)
    t0 = time.perf_counter()
    q, st, rc = po.run(op, q, 5, dtype=np.float32, state=st, fast=lib)
    per_step = max((time.perf_counter() - t0) / 5, 1e-4)
    steps = max(5, int(seconds / per_step))
    t0 = time.perf_counter()
    q, st, rc = po.run(op, q, steps, dtype=np.float32, state=st, fast=lib)
    el = time.perf_counter() - t0
    if rc != 0:
        raise RuntimeError(f"CPU oracle failed: {rc}")
    value = prob.nx * prob.ny * steps / el / 1e9
    # the same on one core, over a shorter sample
    op.nthreads = 1
    t0 = time.perf_counter()
    q, st, rc = po.run(op, q, 2, dtype=np.float32, state=st, fast=lib)
    per_step1 = max((time.perf_counter() - t0) / 2, 1e-4)
    steps1 = max(2, int(seconds * 0.4 / per_step1))
    t0 = time.perf_counter()
    q, st, rc1 = po.run(op, q, steps1, dtype=np.float32, state=st, fast=lib)
    el1 = time.perf_counter() - t0
    if rc1 != 0:
        raise RuntimeError(f"CPU oracle failed: {rc1}")
    return {"value": value, "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(), "cpus_visible": visible, "build": "gcc " + flags,
            "sample": f"{steps} accepted steps of the same {prob.nx}x{prob.ny} grid from the "
                      f"quiescent start, FP32 CPU oracle (OpenMP over rows, {threads} threads), "
                      f"{el:.1f} s; 1 core: {steps1} further steps, {el1:.1f} s",
            "value_1core": prob.nx * prob.ny * steps1 / el1 / 1e9}


def copy_probe(nbytes=256 << 20, reps=5):
    """A plain device-to-device copy of known size (the HIP runtime's copy
    kernel), run under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to calibrate
    those counters: bytes the counters report per byte actually moved."""
    hip = Hip()
    lib = hip.lib
    lib.hipMalloc.restype = C.c_int
    lib.hipMemcpy.restype = C.c_int
    lib.hipDeviceSynchronize.restype = C.c_int
    a, b = C.c_void_p(), C.c_void_p()
    hip.check(lib.hipMalloc(C.byref(a), C.c_size_t(nbytes)), "hipMalloc")
    hip.check(lib.hipMalloc(C.byref(b), C.c_size_t(nbytes)), "hipMalloc")
    lib.hipMemset.restype = C.c_int
    hip.check(lib.hipMemset(a, 1, C.c_size_t(nbytes)), "hipMemset")
    for _ in range(reps):
        hip.check(lib.hipMemcpy(b, a, C.c_size_t(nbytes), 3), "hipMemcpy D2D")
    hip.check(lib.hipDeviceSynchronize(), "hipDeviceSynchronize")
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps}), flush=True)


def bench_batched(args, hip):
    """64 uneven instances (nx, ny in 256..2048 step 64), one launch per step
    for all of them, work split chosen by the autotuner."""
    import fivonagi_amd as fa
    from fivonagi_amd import workloads as wl
    probs = wl.batched(64)
    stream = hip.stream()
    ctx = fa.Context(probs)
    ctx.set_stream(stream.value)
    rows = ctx.autotune(30)
    split = ctx.info(0)
    warm_up(ctx, args.warmup, args.steps)
    before = [ctx.info(k) for k in range(len(probs))]
    e0, e1 = hip.event(), hip.event()
    hip.record(e0, stream)
    ctx.enqueue(args.steps)
    hip.record(e1, stream)
    ms = hip.elapsed_ms(e0, e1)
    if ctx.sync() != 0:
        raise RuntimeError("timed steps failed")
    after = [ctx.info(k) for k in range(len(probs))]
    cells = sum(p.nx * p.ny * (a["n_accepted"] - b["n_accepted"])
                for p, a, b in zip(probs, after, before))
    total = sum(p.nx * p.ny for p in probs)
    value = cells / (ms / 1e3) / 1e9
    achieved = BYTES_PER_CELL * total / (ms / args.steps / 1e3) / 1e9
    ctx.close()
    return {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"batched forward: 64 instances, {total} cells, per-instance dt",
                   "rows_per_wave": rows, "skew": round(split["skew"], 3),
                   "rotate": split["rotate"], "accepted_cell_updates": cells},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": KERNEL},
        "cpu_baseline": None,
    }


def bench_hifu(args, hip):
    """The hifu-beam app at ETAC = 4 (2048x4096, ISPAM = 30): calcFrame's frame
    loop (fivonagi_step_until per frame of DTPRINT = 1) with the moving window
    and the data collection on; `steps` is a minimum, whole frames are
    timed (their host synchronisations included)."""
    import fivonagi_amd as fa
    from fivonagi_amd import workloads as wl
    prob = wl.hifu(4)
    stream = hip.stream()
    ctx = fa.Context(prob)
    ctx.set_stream(stream.value)
    ctx.set_collect(0, **wl.hifu_collect(prob))
    # warm-up: the packet is launched at T ~ 0; start timing past MDT so the
    # window moves and the collection records
    T = prob.T0
    while T < prob.moving[2] + 1.0:      # frames of DTPRINT = 1, as calcFrame
        T += 1.0
        ctx.step_until(T)
    before = ctx.info(0)
    e0, e1 = hip.event(), hip.event()
    hip.record(e0, stream)
    frames = 0
    while True:
        T += 1.0
        ctx.step_until(T)
        frames += 1
        if ctx.info(0)["n_accepted"] - before["n_accepted"] >= args.steps:
            break
    hip.record(e1, stream)
    ms = hip.elapsed_ms(e0, e1)
    after = ctx.info(0)
    steps = after["n_accepted"] - before["n_accepted"]
    cells = prob.nx * prob.ny * steps
    value = cells / (ms / 1e3) / 1e9
    achieved = BYTES_PER_CELL * prob.nx * prob.ny / (ms / steps / 1e3) / 1e9
    ctx.close()
    return {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": steps, "warmup": before["n_accepted"], "ms_per_step": round(ms / steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "hifu-beam app ETAC=4 (2048x4096, ISPAM=30): frame loop, "
                               "moving window, data collection",
                   "frames": frames, "T_start": before["T"], "T_end": after["T"],
                   "md_x": after["md_x"], "rows_per_wave": after["rows_per_wave"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": KERNEL},
        "cpu_baseline": None,
    }


def taylor_rate_summary(err):
    """Mean regression rate per precision and error type over the Taylor
    study's 25 CFL x angle runs (table 1's rate_reg_mean)."""
    from fivonagi_amd import taylor_study as ts
    from fivonagi_amd.study import rates as run_rates
    out = {}
    for pi, p in enumerate(ts.PRECISIONS):
        for ei, et in enumerate(ts.ERROR_TYPES):
            rr = [run_rates(err, ts.ETAS, ci, pi, ai, ei)[1]
                  for ai in range(len(ts.ANGLES)) for ci in range(len(ts.CFLS))]
            out[f"p{p}_{et}_rate_reg_mean"] = round(float(np.mean(rr)), 4)
    return out


def bench_taylor_study(args, hip):
    """The taylor-angle-xy convergence study as the reference runs it
    (apps/taylor-angle-xy/run.sh): 5 resolutions (NX = 64..1024) x 5 CFLs x 5
    angles, each run to T = 100 with the app's moving window, data collection
    and exports (no whole-grid files), in FP32 and in FP64, then the errors
    (plots/error.py) and the convergence tables (plots/error_table_*.py).
    Each precision's 125 runs are one batched Context.  value = cell updates
    per second of the stepping (step_until calls, their host synchronisations
    included); the exports' host time is reported beside it."""
    import shutil
    import tempfile
    from fivonagi_amd import taylor_study as ts
    root = args.study_dir or tempfile.mkdtemp(prefix="taylor_study_")
    res, dep = os.path.join(root, "results"), os.path.join(root, "deploy")
    os.makedirs(res, exist_ok=True)
    os.makedirs(dep, exist_ok=True)
    t0 = time.perf_counter()
    stats = ts.run_sweep(res, log=lambda m: print(m, file=sys.stderr, flush=True))
    err = ts.extract_errors(res, dep)
    ts.write_tables(dep, err)
    wall = time.perf_counter() - t0
    rates = taylor_rate_summary(err)
    cells = sum(v["cell_updates"] for v in stats.values())
    gpu_s = sum(v["gpu_s"] for v in stats.values())
    if not args.study_dir:
        shutil.rmtree(root, ignore_errors=True)
    per_p = {f"p{p}": {"runs": v["instances"], "cell_updates": v["cell_updates"],
                       "stepping_s": round(v["gpu_s"], 3), "wall_s": round(v["wall_s"], 3),
                       "gcell_per_s": round(v["cell_updates"] / v["gpu_s"] / 1e9, 3)}
             for p, v in stats.items()}
    return {
        "metric": METRIC, "value": round(cells / gpu_s / 1e9, 3), "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": None, "warmup": 0, "ms_per_step": None,
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "taylor-angle-xy convergence study: eta 5..82 x cfl 60..99 x "
                               "5 angles x p1/p2 to T=100, exports, errors, tables",
                   "per_precision": per_p, "seconds_total": round(wall, 2), **rates},
        "roofline": None, "cpu_baseline": None,
    }


def bench_adjoint(args, hip):
    """One gradient of J (time-integrated rho'^2 over a 64x64 box at x=1500)
    with respect to 512 edge parameters on 2048x1024: forward (logging each
    step's control block) + discrete adjoint.  Checkpointing is automatic:
    every state is kept when the run's states fit in half the free HBM (no
    recompute), else one every ceil(sqrt(n)) steps.  The forward alone is
    timed on a second context from the same start, so the record splits the
    gradient into forward and adjoint time."""
    import fivonagi_amd as fa
    from fivonagi_amd import workloads as wl
    prob, box, rows = wl.adjoint_single()
    nsteps = args.steps
    cells = prob.nx * prob.ny
    keep_all = 4 * 3 * cells * (nsteps + 1)
    with fa.Context(prob) as c:          # warm-up: capture and first-touch outside the timing
        c.step_adjoint(min(nsteps, 32), box)
    ctx = fa.Context(prob)
    t0 = time.perf_counter()
    J, ga, gp = ctx.step_adjoint(nsteps, box, rows_per_wave=args.rows_per_wave)
    el = time.perf_counter() - t0
    ctx.close()
    with fa.Context(prob) as c:
        t0 = time.perf_counter()
        c.step(nsteps)
        tf = time.perf_counter() - t0
    g = np.concatenate([ga[0][rows], gp[0][rows]])
    adj_ms = (el - tf) * 1e3 / nsteps
    return {
        "metric": METRIC, "value": round(1.0 / el, 4), "unit": "gradients/s", "n_gpus": 1,
        "steps": nsteps, "warmup": 0, "ms_per_step": round(el * 1e3 / nsteps, 4),
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"adjoint 2048x1024, 512 edge parameters, {nsteps} steps, "
                               "automatic checkpointing", "J": J[0],
                   "grad_norm": float(np.linalg.norm(g)), "seconds": round(el, 3),
                   "state_history_bytes": keep_all,
                   "forward_ms_per_step": round(tf * 1e3 / nsteps, 4),
                   "adjoint_ms_per_step": round(adj_ms, 4)},
        "roofline": None, "cpu_baseline": None,
    }


ADJ_BYTES_PER_CELL_STEP = 84   # forward 24 + recompute 24 + adjoint 36 (read q, lambda; write lambda)


def bench_batched_adjoint(args, hip):
    """32 instances of 1024x1024 (own edge profiles, wavelengths and focal
    boxes): the batched forward+adjoint (one launch per pass for all
    instances, graph-replayed forward) against 32 sequential plain-form
    solves (one instance at a time, one launch per forward step with a host
    read of dt).  The plain time is split by timing each solve again with the
    graph-replayed forward: the difference is the per-step launch and
    synchronisation cost, the rest is kernel time of one instance at a time."""
    import fivonagi_amd as fa
    from fivonagi_amd import workloads as wl
    probs, boxes = wl.adjoint_batch(32)
    nsteps = args.steps
    ctx = fa.Context(probs)
    t0 = time.perf_counter()
    J, ga, gp = ctx.step_adjoint(nsteps, boxes)
    tb = time.perf_counter() - t0
    ctx.close()
    n_plain = min(args.plain_instances, len(probs))

    def solve(k, plain):
        c = fa.Context(probs[k])
        t0 = time.perf_counter()
        jk, _, _ = c.step_adjoint(nsteps, boxes[k], plain_forward=plain)
        el = time.perf_counter() - t0
        c.close()
        return el, jk[0]
    solve(0, False)                      # untimed: first use of the single-instance kernels
    # plain and graph-forward solves interleaved; the fastest of each kind,
    # scaled to the 32 instances (after the batched run the device can clock
    # down under the sustained load, which slows the later solves of both
    # kinds alike)
    tps, tgs, Jp = [], [], []
    for k in range(n_plain):
        el, jk = solve(k, True)
        tps.append(el)
        Jp.append(jk)
        el, _ = solve(k, False)
        tgs.append(el)
        print(f"solve {k + 1}/{n_plain}: plain {tps[-1]:.3f} s, graph-forward {tgs[-1]:.3f} s",
              file=sys.stderr, flush=True)
    tp = float(np.min(tps)) * len(probs)
    tg = float(np.min(tgs)) * len(probs)
    rel = max(abs(a - b) / max(abs(b), 1e-300) for a, b in zip(J[:n_plain], Jp))
    cells = sum(p.nx * p.ny for p in probs)
    achieved = ADJ_BYTES_PER_CELL_STEP * cells * nsteps / tb / 1e9
    return {
        "metric": METRIC, "value": round(len(probs) / tb, 4), "unit": "gradients/s",
        "n_gpus": 1, "steps": nsteps, "warmup": 0, "ms_per_step": round(tb * 1e3 / nsteps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"batched forward+adjoint: 32 x 1024x1024, {nsteps} steps",
                   "batched_seconds": round(tb, 3),
                   "plain_seconds_for_32": round(tp, 3),
                   "plain_instances_timed": n_plain,
                   "speedup_vs_plain": round(tp / tb, 3),
                   "plain_split": {
                       "graph_forward_seconds_for_32": round(tg, 3),
                       "per_step_launch_and_sync_seconds": round(tp - tg, 3),
                       "speedup_from_batching_alone": round(tg / tb, 3)},
                   "J_rel_diff_batched_vs_plain": rel},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "forward + recompute + adjoint march kernels",
                     "bytes_per_cell_step": ADJ_BYTES_PER_CELL_STEP},
        "cpu_baseline": None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--nx", type=int, default=4096)
    ap.add_argument("--ny", type=int, default=1024)
    ap.add_argument("--rows-per-wave", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fp64", action="store_true",
                    help="edge_drive in FP64 (the reference's PRECISION=2); not the headline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain copy of known size (counter calibration)")
    ap.add_argument("--workload", default="edge_drive",
                    choices=["edge_drive", "batched", "adjoint", "batched_adjoint", "hifu",
                             "taylor_study"])
    ap.add_argument("--study-dir", default=None,
                    help="taylor_study: keep the runs and tables here (default: a temporary "
                         "directory, removed)")
    ap.add_argument("--plain-instances", type=int, default=32,
                    help="batched_adjoint: plain solves timed (scaled to 32)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.workload != "edge_drive":
        if args.gpus != 1:
            print("bench.py measures one GPU only", file=sys.stderr)
            sys.exit(2)
        hip = Hip()
        fn = {"batched": bench_batched, "adjoint": bench_adjoint,
              "batched_adjoint": bench_batched_adjoint, "hifu": bench_hifu,
              "taylor_study": bench_taylor_study}[args.workload]
        print(json.dumps(fn(args, hip)), flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the step "
              "would shard by instances or by y-slabs, but no multi-GPU path is built.",
              file=sys.stderr)
        sys.exit(2)

    import fivonagi_amd as fa
    from fivonagi_amd import workloads as wl

    hip = Hip()
    if hip.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    prob = wl.edge_drive(args.nx, args.ny, rows_per_wave=args.rows_per_wave)
    stream = hip.stream()
    ctx = fa.Context(prob, fp64=args.fp64)
    ctx.set_stream(stream.value)
    info0 = ctx.info(0)

    warm_up(ctx, args.warmup, args.steps)
    before = ctx.info(0)
    e0, e1 = hip.event(), hip.event()
    hip.record(e0, stream)
    ctx.enqueue(args.steps)
    hip.record(e1, stream)
    ms = hip.elapsed_ms(e0, e1)
    rc = ctx.sync()
    if rc != 0:
        raise RuntimeError(f"timed steps failed: {rc}")
    after = ctx.info(0)
    accepted = after["n_accepted"] - before["n_accepted"]
    attempts = after["n_attempts"] - before["n_attempts"]
    cells = prob.nx * prob.ny
    secs = ms / 1e3
    value = cells * accepted / secs / 1e9
    ms_per_step = ms / args.steps
    # dominant (only) kernel: one launch per attempt
    bpc = BYTES_PER_CELL * (2 if args.fp64 else 1)
    achieved = bpc * cells / (ms_per_step / 1e3) / 1e9
    tag = f"edge_drive_{prob.nx}x{prob.ny}" + ("_fp64" if args.fp64 else "")
    traffic = traffic_from_profiles(tag)
    rec = profile_record(tag)
    valu = None
    if "valu_active_quad_cycles_per_dispatch" in rec:
        # VALU issue: quad-cycles per SIMD per launch (counter run) against the
        # cycles of this run's launch at the 2.4 GHz peak clock
        cyc = 4.0 * rec["valu_active_quad_cycles_per_dispatch"] / rec["simds"]
        valu = {"issue_cycles_per_simd": round(cyc),
                "frac_at_peak_clock": round(cyc / (ms_per_step * 1e-3 * PEAK_CLOCK_HZ), 4),
                "lane_slots_per_cell": round(64.0 * rec["valu_active_quad_cycles_per_dispatch"]
                                             / cells, 1),
                "peak_clock_ghz": PEAK_CLOCK_HZ / 1e9}
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.fp64 else "f32",
        "data": "synthetic",
        "config": {
            "workload": f"edge_drive {prob.nx}x{prob.ny} {'fp64' if args.fp64 else 'fp32'} "
                        f"forward (A=0.02, aperture "
                        f"half-width 200 cells, 20 cells/wavelength, beta=3.5, delta=5e-5, CFL 0.9)",
            "accepted_steps": accepted, "attempts": attempts,
            "rows_per_wave": after["rows_per_wave"], "skew": round(after["skew"], 3),
            "rotate": after["rotate"], "waves_per_step": after["n_items"],
            "T_start": before["T"], "T_end": after["T"],
        },
        "hbm_peak_frac": round(achieved / HBM_PEAK_GBS, 4),
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": KERNEL,
            "bytes_per_launch": bpc * cells,
            "avg_launch_ms": round(ms_per_step, 5),
            "valu": valu,
        },
        "cpu_baseline": None,
    }
    ctx.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(prob, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
