"""Parity at the full configurations, against the FP64 oracle (rho'-direct
form).  Not collected by pytest: the oracle needs minutes of CPU time for the
4096x1024 case.  Run on a machine with the GPU, from the repository root:

    python -u tests/parity_full_configs.py

  C1  400x400 Gaussian pulse, 1000 accepted steps, FP32 and FP64 contexts;
  C2  4096x1024 edge drive, 8000 accepted steps, FP32 and FP64 contexts;
  C3  4 of the 64 batched uneven instances (the smallest ones), 2000 accepted
      steps each, advanced inside the full 64-instance batch.

For each: per-field relative L-infinity max|q - q_oracle| / max|q_oracle|,
accepted/rejected counts, and T.  FP32 results are also set beside the FP32
oracle's own distance from the FP64 oracle (what FP32 arithmetic costs on
that run).  The oracle advances in chunks with a progress line after each.
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


def oracle_run(prob, n, dtype=np.float64, chunk=250, label=""):
    op = to_oracle(prob, form=po.FORM_RHO1, nthreads=THREADS)
    q = prob.q0_f64.copy()
    st = None
    done = 0
    t0 = time.perf_counter()
    while done < n:
        k = min(chunk, n - done)
        q, st, rc = po.run(op, q, k, dtype=dtype, state=st, fast=True)
        if rc != 0:
            raise RuntimeError(f"oracle failed: {rc}")
        done += k
        print(f"  oracle {label} {np.dtype(dtype).name}: {done}/{n} steps, "
              f"{time.perf_counter() - t0:.0f} s", flush=True)
    return np.asarray(q, np.float64), st


def line(name, e, st, sto):
    print(f"{name}: rel Linf q1={e[0]:.3e} q2={e[1]:.3e} q3={e[2]:.3e}  "
          f"accepted {st['n_accepted']}/{sto.n_accepted} rejected "
          f"{st['n_rejected']}/{sto.n_rejected}  T {st['T']:.9f}/{sto.T:.9f}", flush=True)


def single(name, prob, n):
    res = {}
    for fp64 in (False, True):
        with fa.Context(prob, fp64=fp64) as ctx:
            ctx.step(n)
            res[fp64] = ctx.collect(0)
    qo, sto = oracle_run(prob, n, label=name)
    q32o, _ = oracle_run(prob, n, dtype=np.float32, label=name)
    line(f"{name} GPU FP32", rel_linf(res[False][0], qo), res[False][1], sto)
    line(f"{name} GPU FP64", rel_linf(res[True][0], qo), res[True][1], sto)
    e = rel_linf(q32o, qo)
    print(f"{name} oracle FP32 vs FP64: rel Linf q1={e[0]:.3e} q2={e[1]:.3e} q3={e[2]:.3e}",
          flush=True)


def batched(n=2000, picks=4):
    probs = wl.batched(64)
    with fa.Context(probs) as ctx:
        ctx.step(n)
        outs = [ctx.collect(k) for k in range(len(probs))]
    order = sorted(range(len(probs)), key=lambda k: probs[k].nx * probs[k].ny)[:picks]
    for k in order:
        p = probs[k]
        qo, sto = oracle_run(p, n, label=f"C3[{k}]")
        line(f"C3 instance {k} ({p.nx}x{p.ny}) GPU FP32", rel_linf(outs[k][0], qo), outs[k][1],
             sto)


if __name__ == "__main__":
    fa.load()
    single("C1 400x400 pulse 1000 steps", wl.gauss_pulse(400, 400), 1000)
    batched()
    single("C2 4096x1024 edge drive 8000 steps", wl.edge_drive(4096, 1024), 8000)
