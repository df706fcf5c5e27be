"""The workloads the throughput and gradient figures are quoted on, at their
full sizes, against the FP64 oracle (or the FP64 host adjoint).

Step counts are bounded by what the host oracle finishes in seconds on the
GPU host's cores.  Tolerances (per-field relative L-infinity, accepted and
rejected counts exact):
  FP64 contexts: 1e-10.
  FP32, C2: 1e-5 for rho' and rho u; rho v, the transverse momentum only
      the aperture's edges generate (1/40 of |rho u| at its largest), at
      1e-5 of the momentum's largest magnitude |(rho u, rho v)|; relative to
      its own maximum it is reported (the FP32 oracle is at 1.7e-4 there).
  FP32, C1: the FP32 oracle's own distance from the FP64 oracle.  After
      1000 steps the pulse has left the grid down to a peak |rho'| of
      3.7e-6 (from 0.01), and the FP32 error is a nearly uniform offset of
      ~1e-10 (mass round-off of the FP32 state, which the zero-gradient edges
      keep): relative to that peak it is 1e-5 to 6e-5 whichever arithmetic
      form is used, even with every stage in FP64 and only the stored state
      rounded to FP32.

  C1  400x400 Gaussian pulse, zero-gradient edges (FP32 and FP64)
  C2  4096x1024 left-edge drive through a 200-cell Gaussian aperture (FP32),
      with the dt feedback and with the FP64 oracle's dt prescribed
"""
import os

import numpy as np
import pytest

import fivonagi_amd as fa
from fivonagi_amd import workloads as wl
from tests.conftest import rel_linf
from tests.oracle_bridge import to_oracle

pytestmark = pytest.mark.gpu

TOL = 1e-5
TOL64 = 1e-10
THREADS = min(len(os.sched_getaffinity(0)), 16)


def momentum_rel(q, qo):
    """Largest error of the momentum vector (rho u, rho v) relative to the
    momentum's largest magnitude."""
    d = np.hypot(q[1].astype(np.float64) - qo[1], q[2].astype(np.float64) - qo[2])
    return float(d.max() / np.hypot(qo[1], qo[2]).max())


def check_c2(e, em):
    assert e[0] < TOL and e[1] < TOL and em < TOL, (e, em)


def report(name, e):
    print(f"[config] {name}: rel Linf " + " ".join(f"{v:.3e}" for v in e), flush=True)


def oracle_run(oracle, prob, n, dt_seq=None, collect_T=False):
    """FP64 oracle (rho'-direct form) from the unrounded initial state; with
    collect_T the T of every accepted step (data collection row 0)."""
    op = to_oracle(prob, form=oracle.FORM_RHO1, dt_seq=dt_seq, nthreads=THREADS)
    if not collect_T:
        q, st, rc = oracle.run(op, prob.q0_f64.copy(), n, fast=True)
        assert rc == 0
        return q, st, None
    op.collect = dict(jcut=2, mic=(0, 0, 0, 0), eta=20.0)
    q, st, Ts = prob.q0_f64.copy(), None, []
    chunk = min(prob.nx, n)
    for done in range(0, n, chunk):
        op.n0frame = done
        m = min(chunk, n - done)
        q, st, rc = oracle.run(op, q, m, state=st, fast=True)
        assert rc == 0
        Ts.append(op.measure[0, :m].copy())
    return q, st, np.concatenate(Ts)


@pytest.mark.parametrize("fp64", [False, True])
def test_c1_pulse_400(oracle, fp64):
    """C1 in full: 1000 accepted steps of the 400x400 pulse."""
    prob = wl.gauss_pulse(400, 400)
    n = 1000
    with fa.Context(prob, fp64=fp64) as ctx:
        ctx.step(n)
        q, st = ctx.collect(0)
    qo, sto, _ = oracle_run(oracle, prob, n)
    e = rel_linf(q, qo)
    report(f"C1 400x400 pulse {n} steps {'fp64' if fp64 else 'fp32'} (T {st['T']:.10f} vs "
           f"{sto.T:.10f}, mean q1 offset {float((q[0] - qo[0]).mean()):+.2e})", e)
    assert st["n_accepted"] == sto.n_accepted and st["n_rejected"] == sto.n_rejected
    if fp64:
        assert max(e) < TOL64
        return
    op32 = to_oracle(prob, form=oracle.FORM_RHO1, nthreads=THREADS)
    q32, _, rc = oracle.run(op32, prob.q0_f64.copy(), n, dtype=np.float32, fast=True)
    assert rc == 0
    e32 = rel_linf(q32, qo)
    report("C1 400x400 pulse, FP32 oracle", e32)
    assert all(a <= b for a, b in zip(e, e32)), (e, e32)


def test_c2_edge_drive_4096x1024(oracle):
    """C2 at its size, 600 accepted steps: the drive has steepened into
    shocks over the first wavelengths.  Feedback dt (counts and T against
    the oracle's), then the same run with the oracle's dt sequence."""
    prob = wl.edge_drive()
    assert (prob.nx, prob.ny) == (4096, 1024)
    n = 600
    qo, sto, T_o = oracle_run(oracle, prob, n, collect_T=True)
    with fa.Context(prob) as ctx:
        ctx.step(n)
        q, st = ctx.collect(0)
    e, em = rel_linf(q, qo), momentum_rel(q, qo)
    report(f"C2 4096x1024 edge drive {n} steps, feedback dt (T {st['T']:.9f} vs "
           f"{sto.T:.9f}; momentum {em:.3e})", e)
    assert st["n_accepted"] == sto.n_accepted and st["n_rejected"] == sto.n_rejected
    assert st["T"] == pytest.approx(sto.T, rel=1e-7)
    check_c2(e, em)
    dts = np.diff(np.concatenate([[prob.T0], T_o]))
    with fa.Context(prob) as ctx:
        ctx.set_dt_sequence(0, dts)
        ctx.step(n)
        qp, sp = ctx.collect(0)
    ep, emp = rel_linf(qp, qo), momentum_rel(qp, qo)
    report(f"C2 4096x1024 edge drive {n} steps, FP64 oracle's dt (momentum {emp:.3e})", ep)
    assert sp["T"] == pytest.approx(sto.T, rel=1e-12)
    check_c2(ep, emp)


def test_c3_batched_64_uneven(oracle):
    """C3 at its size: the 64 uneven instances (nx, ny in 256..2048) in one
    context, autotuned, 100 accepted steps each in one launch per attempt.
    Six instances (the two smallest, the largest, three others) against the
    FP64 oracle run alone; every instance's counts against its own T."""
    probs = wl.batched(64)
    cells = sum(p.nx * p.ny for p in probs)
    n = 100
    with fa.Context(probs) as ctx:
        ctx.autotune(4)
        ctx.step(n)
        infos = [ctx.info(k) for k in range(len(probs))]
        order = sorted(range(len(probs)), key=lambda k: probs[k].nx * probs[k].ny)
        pick = order[:2] + order[-1:] + [order[20], order[40], order[50]]
        outs = {k: ctx.collect(k) for k in pick}
    print(f"[config] C3 64 instances, {cells} cells, rows per wave {infos[0]['rows_per_wave']}",
          flush=True)
    assert all(i["n_accepted"] == n and i["error"] == 0 for i in infos)
    for k in pick:
        p = probs[k]
        q, st = outs[k]
        qo, sto, _ = oracle_run(oracle, p, n)
        e = rel_linf(q, qo)
        report(f"C3 instance {k} {p.nx}x{p.ny}", e)
        assert st["n_rejected"] == sto.n_rejected
        assert st["T"] == pytest.approx(sto.T, rel=1e-7)
        assert max(e) < TOL


def near_drive_box(prob):
    """A 64x64 box 8 cells from the driven edge, centred in y: reached by the
    drive within the few steps a host adjoint can afford."""
    return (8, prob.ny // 2 - 32, 64, 64)


def test_c4_adjoint_2048x1024(oracle):
    """C4 at its size: 2048x1024, 256 aperture rows x (amplitude, phase) =
    512 parameters, 20 steps with a frozen dt, the box near the drive.  The
    HIP gradient against the FP64 host adjoint (1e-3, max-norm) and against
    FP64 central differences on 4 random directions of the 512 parameters
    (1e-3, best of two step sizes)."""
    prob, _, rows = wl.adjoint_single()
    assert (prob.nx, prob.ny, len(rows)) == (2048, 1024, 256)
    n, box = 20, near_drive_box(prob)
    dts = np.full(n, 0.04)
    with fa.Context(prob) as ctx:
        ctx.set_dt_sequence(0, dts)
        J, ga, gp = ctx.step_adjoint(n, box, checkpoint_every=8)
    J, ga, gp = J[0], ga[0], gp[0]
    q0 = prob.q0.astype(np.float64)
    op = to_oracle(prob, dt_seq=dts, nthreads=THREADS)
    Jo, gao, gpo, _, rc = oracle.adjoint(op, q0, n, box)
    assert rc == 0 and Jo > 0
    g, go = np.concatenate([ga[rows], gp[rows]]), np.concatenate([gao[rows], gpo[rows]])
    e = (abs(J - Jo) / Jo, float(np.abs(g - go).max() / np.abs(go).max()))
    print(f"[config] C4 J {J:.6e} vs {Jo:.6e}: rel err J {e[0]:.2e}, gradient {e[1]:.2e}",
          flush=True)
    assert e[0] < 1e-5 and e[1] < 1e-3
    # the aperture only: parameters outside it are 0 and stay untouched
    assert not ga[np.setdiff1d(np.arange(prob.ny), rows)].any()
    rng = np.random.default_rng(17)
    amp0, ph0 = prob.edge_amp.copy(), prob.edge_phase.copy()
    worst = 0.0
    for trial in range(4):
        d = np.zeros(2 * prob.ny)
        d[np.concatenate([rows, prob.ny + rows])] = rng.standard_normal(2 * len(rows))
        fds = []
        for h in (1e-5, 1e-6):
            Js = []
            for sgn in (1, -1):
                pp = wl.edge_drive(prob.nx, prob.ny, edge_amp=amp0 + sgn * h * d[:prob.ny],
                                   edge_phase=ph0 + sgn * h * d[prob.ny:], beta=prob.beta,
                                   delta=prob.delta)
                Jx, rc = oracle.objective(to_oracle(pp, dt_seq=dts, nthreads=THREADS), q0, n, box)
                assert rc == 0
                Js.append(Jx)
            fds.append((Js[0] - Js[1]) / (2 * h))
        ad = float(np.concatenate([ga, gp]) @ d)
        err = min(abs(f - ad) / max(abs(f), 1e-300) for f in fds)
        worst = max(worst, err)
        assert err < 1e-3, (trial, fds, ad)
    print(f"[config] C4 4 FD directions: worst rel err {worst:.2e}", flush=True)


def test_c5_batched_adjoint_32x1024(oracle):
    """C5 at its size: 32 instances of 1024x1024, each with its own 256-row
    edge profile and wavelength, 12 steps with a frozen dt, boxes near the
    drive.  The batched J and gradients equal the plain-form solves (one
    instance at a time, one launch per forward step with a host read of dt);
    instance 0 against the FP64 host adjoint."""
    probs, _ = wl.adjoint_batch(32)
    assert len(probs) == 32 and all((p.nx, p.ny) == (1024, 1024) for p in probs)
    n = 12
    boxes = [near_drive_box(p) for p in probs]
    dts = np.full(n, 0.04)
    with fa.Context(probs) as ctx:
        for k in range(len(probs)):
            ctx.set_dt_sequence(k, dts)
        J, ga, gp = ctx.step_adjoint(n, boxes)
    worst = 0.0
    for k, (p, b) in enumerate(zip(probs, boxes)):
        with fa.Context(p) as c:
            c.set_dt_sequence(0, dts)
            Jk, gak, gpk = c.step_adjoint(n, b, plain_forward=True)
        assert J[k] > 0
        g, gk = np.concatenate([ga[k], gp[k]]), np.concatenate([gak[0], gpk[0]])
        e = max(abs(J[k] - Jk[0]) / Jk[0], float(np.abs(g - gk).max() / np.abs(gk).max()))
        worst = max(worst, e)
        assert e < 1e-9, (k, e)
    print(f"[config] C5 batched vs 32 plain-form solves: worst rel diff {worst:.2e}", flush=True)
    p = probs[0]
    op = to_oracle(p, dt_seq=dts, nthreads=THREADS)
    Jo, gao, gpo, _, rc = oracle.adjoint(op, p.q0.astype(np.float64), n, boxes[0])
    assert rc == 0
    g, go = np.concatenate([ga[0], gp[0]]), np.concatenate([gao, gpo])
    e = (abs(J[0] - Jo) / Jo, float(np.abs(g - go).max() / np.abs(go).max()))
    print(f"[config] C5 instance 0 vs FP64 host adjoint: J {e[0]:.2e}, gradient {e[1]:.2e}",
          flush=True)
    assert e[0] < 1e-5 and e[1] < 1e-3
