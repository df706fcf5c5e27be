"""The workloads at their full lengths against the FP64 oracle.  The host
oracle needs about ten minutes for the 4096x1024 case, so these run only
when FIVONAGI_FULL_LENGTH=1 is set (the same workloads at bounded lengths run
in tests/test_gpu_configs.py every time):

    FIVONAGI_FULL_LENGTH=1 python -u -m pytest tests/test_gpu_full_length.py -v -s

  C2  4096x1024 edge drive, 8000 accepted steps, FP32 and FP64 contexts;
  C3  four of the 64 uneven instances, 2000 accepted steps each, advanced
      inside the full 64-instance batch.

Per-field relative L-infinity against the FP64 oracle (rho'-direct form),
with exact accepted/rejected counts.  FP32 contexts are held to the FP32
oracle's own distance from the FP64 oracle, field by field (what the
reference's arithmetic in FP32 reaches on the same run), FP64 contexts to
1e-10.  The oracle advances in chunks with a progress line after each.
"""
import os
import time

import numpy as np
import pytest

import fivonagi_amd as fa
from fivonagi_amd import workloads as wl
from tests.conftest import rel_linf
from tests.oracle_bridge import to_oracle

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("FIVONAGI_FULL_LENGTH") != "1",
                                 reason="full-length runs: set FIVONAGI_FULL_LENGTH=1")]

THREADS = min(len(os.sched_getaffinity(0)), 16)
TOL64 = 1e-10


def oracle_run(oracle, prob, n, dtype=np.float64, chunk=250, label=""):
    op = to_oracle(prob, form=oracle.FORM_RHO1, nthreads=THREADS)
    q, st, done = prob.q0_f64.copy(), None, 0
    t0 = time.perf_counter()
    while done < n:
        k = min(chunk, n - done)
        q, st, rc = oracle.run(op, q, k, dtype=dtype, state=st, fast=True)
        assert rc == 0
        done += k
        print(f"  oracle {label} {np.dtype(dtype).name}: {done}/{n} steps, "
              f"{time.perf_counter() - t0:.0f} s", flush=True)
    return np.asarray(q, np.float64), st


def report(name, e, st, sto):
    print(f"[full] {name}: rel Linf " + " ".join(f"{v:.3e}" for v in e) +
          f"  accepted {st['n_accepted']}/{sto.n_accepted} rejected "
          f"{st['n_rejected']}/{sto.n_rejected}  T {st['T']:.9f}/{sto.T:.9f}", flush=True)


def test_c2_full_8000(oracle):
    prob = wl.edge_drive()
    n = 8000
    out = {}
    for fp64 in (False, True):
        with fa.Context(prob, fp64=fp64) as ctx:
            ctx.step(n)
            out[fp64] = ctx.collect(0)
    print("[full] C2 GPU runs done", flush=True)
    qo, sto = oracle_run(oracle, prob, n, label="C2")
    q32, st32 = oracle_run(oracle, prob, n, dtype=np.float32, label="C2")
    e32 = rel_linf(q32, qo)
    print(f"[full] C2 FP32 oracle: rel Linf " + " ".join(f"{v:.3e}" for v in e32) +
          f"  T {st32.T:.9f}", flush=True)
    for fp64 in (False, True):
        q, st = out[fp64]
        e = rel_linf(q, qo)
        report(f"C2 4096x1024 {n} steps {'FP64' if fp64 else 'FP32'}", e, st, sto)
        assert st["n_accepted"] == sto.n_accepted and st["n_rejected"] == sto.n_rejected
        if fp64:
            assert max(e) < TOL64
        else:
            assert all(a <= b for a, b in zip(e, e32)), (e, e32)


def test_c3_full_2000(oracle):
    probs = wl.batched(64)
    n = 2000
    order = sorted(range(len(probs)), key=lambda k: probs[k].nx * probs[k].ny)
    pick = order[:4]
    with fa.Context(probs) as ctx:
        ctx.step(n)
        outs = {k: ctx.collect(k) for k in pick}
    for k in pick:
        p = probs[k]
        qo, sto = oracle_run(oracle, p, n, chunk=500, label=f"C3 #{k}")
        q32, _ = oracle_run(oracle, p, n, dtype=np.float32, chunk=1000, label=f"C3 #{k}")
        e32 = rel_linf(q32, qo)
        q, st = outs[k]
        e = rel_linf(q, qo)
        report(f"C3 instance {k} {p.nx}x{p.ny} {n} steps (FP32 oracle " +
               " ".join(f"{v:.2e}" for v in e32) + ")", e, st, sto)
        assert st["n_accepted"] == sto.n_accepted and st["n_rejected"] == sto.n_rejected
        assert all(a <= b for a, b in zip(e, e32)), (k, e, e32)
