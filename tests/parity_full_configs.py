"""Parity at the full configurations, against the FP64 oracle (rho'-direct
form).  Not collected by pytest: the oracle needs minutes of CPU time for the
4096x1024 case.  Run on a machine with the GPU, from the repository root:

    python -u tests/parity_full_configs.py

  C1  400x400 Gaussian pulse, 1000 accepted steps, FP32 and FP64 contexts;
  C3  4 of the 64 batched uneven instances (the smallest ones), 2000 accepted
      steps each, advanced inside the full 64-instance batch;
  C2  4096x1024 edge drive, 8000 accepted steps, FP32 and FP64 contexts.

For each: per-field relative L-infinity max|q - q_oracle| / max|q_oracle|,
accepted/rejected counts, and T.  FP32 results are also set beside the FP32
oracle's own distance from the FP64 oracle (what FP32 arithmetic costs on
that run).  The oracle advances in chunks with a progress line after each.

For C2 the FP32 run is also taken apart (the drift analysis):
  - T and cfl of every accepted step, on the GPU (fivonagi_info after each
    one-step call) and in the FP64 oracle (its data collection rows 0 and 1),
    as per-1000-step means of the relative dt difference;
  - the field difference fitted to a shift along x;
  - the GPU run again with the FP64 oracle's dt sequence prescribed: its field
    error, and the per-1000-step means of the relative difference of its cfl
    (with dt equal, that of the maximum wave speed);
  - the maximum wave speed of the FP64 oracle's final state, evaluated with the
    GPU's FP32 form (float32 arithmetic, fma exact through float64) and in
    float64.
"""
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import fivonagi_amd as fa                       # noqa: E402
from fivonagi_amd import workloads as wl        # noqa: E402
from oracle import pyoracle as po               # noqa: E402
from tests.conftest import rel_linf             # noqa: E402
from tests.oracle_bridge import to_oracle       # noqa: E402

THREADS = min(os.cpu_count() or 1, 16)


def oracle_run(prob, n, dtype=np.float64, chunk=250, label="", record=False):
    """Returns (q, state, records); records = (T, cfl) of every accepted
    step when record is set (data collection, one frame per chunk)."""
    op = to_oracle(prob, form=po.FORM_RHO1, nthreads=THREADS)
    if record:
        op.collect = dict(jcut=2, mic=(0, 0, 0, 0), eta=20.0)
        chunk = min(chunk, prob.nx)
    q = prob.q0_f64.copy()
    st = None
    done = 0
    T, cfl = [], []
    t0 = time.perf_counter()
    while done < n:
        k = min(chunk, n - done)
        if record:
            op.n0frame = done
        q, st, rc = po.run(op, q, k, dtype=dtype, state=st, fast=True)
        if rc != 0:
            raise RuntimeError(f"oracle failed: {rc}")
        if record:
            T.append(op.measure[0, :k].copy())
            cfl.append(op.measure[1, :k].copy())
        done += k
        print(f"  oracle {label} {np.dtype(dtype).name}: {done}/{n} steps, "
              f"{time.perf_counter() - t0:.0f} s", flush=True)
    rec = (np.concatenate(T), np.concatenate(cfl)) if record else None
    return np.asarray(q, np.float64), st, rec


def line(name, e, st, sto):
    print(f"{name}: rel Linf q1={e[0]:.3e} q2={e[1]:.3e} q3={e[2]:.3e}  "
          f"accepted {st['n_accepted']}/{sto.n_accepted} rejected "
          f"{st['n_rejected']}/{sto.n_rejected}  T {st['T']:.9f}/{sto.T:.9f}", flush=True)


def single(name, prob, n, record=False):
    res = {}
    for fp64 in (False, True):
        with fa.Context(prob, fp64=fp64) as ctx:
            ctx.step(n)
            res[fp64] = ctx.collect(0)
    qo, sto, rec = oracle_run(prob, n, label=name, record=record)
    q32o, _, _ = oracle_run(prob, n, dtype=np.float32, label=name)
    line(f"{name} GPU FP32", rel_linf(res[False][0], qo), res[False][1], sto)
    line(f"{name} GPU FP64", rel_linf(res[True][0], qo), res[True][1], sto)
    e = rel_linf(q32o, qo)
    print(f"{name} oracle FP32 vs FP64: rel Linf q1={e[0]:.3e} q2={e[1]:.3e} q3={e[2]:.3e}",
          flush=True)
    return res[False][0], qo, q32o, rec


def batched(n=2000, picks=4):
    probs = wl.batched(64)
    with fa.Context(probs) as ctx:
        ctx.step(n)
        outs = [ctx.collect(k) for k in range(len(probs))]
    order = sorted(range(len(probs)), key=lambda k: probs[k].nx * probs[k].ny)[:picks]
    for k in order:
        p = probs[k]
        qo, sto, _ = oracle_run(p, n, label=f"C3[{k}]")
        line(f"C3 instance {k} ({p.nx}x{p.ny}) GPU FP32", rel_linf(outs[k][0], qo), outs[k][1],
             sto)


def gpu_records(prob, n, dts=None):
    """T and cfl after every accepted step (one-step calls), and the final
    state; dts: a prescribed dt sequence."""
    T, cfl = np.zeros(n), np.zeros(n)
    with fa.Context(prob) as ctx:
        if dts is not None:
            ctx.set_dt_sequence(0, dts)
        for k in range(n):
            ctx.step(1)
            i = ctx.info(0)
            T[k], cfl[k] = i["T"], i["last_cfl"]
        q, _ = ctx.collect(0)
    return q, T, cfl


def block_means(x, block=1000):
    return " ".join(f"{x[a:a + block].mean():+.2e}" for a in range(0, len(x), block))


def shift_fit(q1, q1o, dx):
    """Least-squares a in q1 - q1o = -a d(q1o)/dx; returns (a, residual rel Linf)."""
    g = np.gradient(q1o, axis=1) / dx
    d = q1 - q1o
    a = -np.sum(d * g) / np.sum(g * g)
    return a, np.abs(d + a * g).max() / np.abs(q1o).max()


def max_speed(q, hbeta, f32_form):
    """Largest |L1|, |L2| over the x and y interfaces of q (3, ny, nx):
    solver.h riemann() in float32 (rcp/rsq as correctly rounded division and
    square root, fma exact through float64) or the same sums in float64."""
    f32 = np.float32

    def fma(a, b, c):
        return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(f32)

    out = []
    for axis, comp in ((1, 1), (0, 2)):
        a = q.astype(f32) if f32_form else q
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[axis + 1] = slice(0, -1)
        hi[axis + 1] = slice(1, None)
        L, R = a[tuple(lo)], a[tuple(hi)]
        one = L.dtype.type(1)
        unL, unR = L[comp] / (one + L[0]), R[comp] / (one + R[0])
        s1, s2 = L[0] + R[0], unL + unR
        tmp1 = s2 / (s1 + L.dtype.type(2))
        if f32_form:
            tmp3 = fma(tmp1, tmp1 - s2, fma(np.full_like(s1, hbeta), s1, np.ones_like(s1)))
        else:
            tmp3 = tmp1 * (tmp1 - s2) + hbeta * s1 + 1.0
        sq = np.sqrt(tmp3)
        out.append(np.maximum(np.abs(tmp1 + sq), np.abs(tmp1 - sq)).max())
    return float(max(out))


def drift(prob, n, q_gpu, qo, q32o, rec):
    """The C2 FP32 drift analysis (module docstring)."""
    To, co = rec
    qg, Tg, cg = gpu_records(prob, n)
    same = np.array_equal(np.asarray(qg, np.float64), np.asarray(q_gpu, np.float64))
    print(f"C2 drift: one-step calls give the same state as one call: {same}", flush=True)
    dto = np.diff(np.concatenate([[prob.T0], To]))
    dtg = np.diff(np.concatenate([[prob.T0], Tg]))
    print("C2 drift: GPU dt / oracle dt - 1, means per 1000 steps: "
          + block_means((dtg - dto) / dto), flush=True)
    print("C2 drift: T difference at each 1000th step: "
          + " ".join(f"{Tg[k] - To[k]:+.2e}" for k in range(999, n, 1000)), flush=True)
    print("C2 drift: GPU cfl / oracle cfl - 1, means per 1000 steps: "
          + block_means((cg - co) / co), flush=True)
    for name, q1 in (("GPU FP32", np.asarray(qg[0], np.float64)), ("oracle FP32", q32o[0])):
        a, res = shift_fit(q1, qo[0], prob.dx)
        print(f"C2 drift: {name} field difference as a shift along x: {a:+.3e} "
              f"(residual rel Linf {res:.3e})", flush=True)
    dts = np.concatenate([dto, np.full(8, dto[-1])])
    qp, Tp, cp = gpu_records(prob, n, dts)
    e = rel_linf(qp, qo)
    print(f"C2 drift: GPU FP32 with the oracle's dt sequence: rel Linf q1={e[0]:.3e} "
          f"q2={e[1]:.3e} q3={e[2]:.3e}  T {Tp[-1]:.9f}/{To[-1]:.9f}", flush=True)
    print("C2 drift: with equal dt, GPU cfl / oracle cfl - 1, means per 1000 steps: "
          + block_means((cp - co) / co), flush=True)
    hb = prob.beta - 1.0
    s32, s64 = max_speed(qo, hb, True), max_speed(qo, hb, False)
    print(f"C2 drift: max wave speed of the oracle's final state: FP32 form {s32:.10f}, "
          f"FP64 {s64:.10f}, relative {s32 / s64 - 1:+.2e}", flush=True)


if __name__ == "__main__":
    fa.load()
    single("C1 400x400 pulse 1000 steps", wl.gauss_pulse(400, 400), 1000)
    batched()
    c2 = wl.edge_drive(4096, 1024)
    q_gpu, qo, q32o, rec = single("C2 4096x1024 edge drive 8000 steps", c2, 8000, record=True)
    drift(c2, 8000, q_gpu, qo, q32o, rec)
