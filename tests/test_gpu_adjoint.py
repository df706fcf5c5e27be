"""HIP adjoint (fivonagi_step_adjoint) against the FP64 host adjoint
(oracle/fvadjoint.c) and FP64 central finite differences of the oracle
forward, with the same prescribed (frozen) dt sequence.

Tolerance: gradient vectors within 1e-3 (relative, max-norm) of the FP64
adjoint; directional derivatives within 1e-3 of the FP64 finite differences
(best of three step sizes: the step is piecewise smooth)."""
import numpy as np
import pytest

import fivonagi_amd as fa
from tests.oracle_bridge import to_oracle
from tests.test_adjoint import (BOX, NSTEPS, NX, NY, best_rel, mk, mk_taylor, mk_zero_gradient,
                                setup_case)

pytestmark = pytest.mark.gpu


def gpu_adjoint(prob, nsteps, dts, box, K=0, plain=False):
    with fa.Context(prob) as ctx:
        ctx.set_dt_sequence(0, dts)
        J, ga, gp = ctx.step_adjoint(nsteps, box, checkpoint_every=K, plain_forward=plain)
        lam = ctx.adjoint_state(0)
    return J[0], ga[0], gp[0], lam


def rel_max(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def case(seed=2):
    rng = np.random.default_rng(seed)
    amp, ph, q0 = setup_case(rng)
    prob = mk(amp, ph)
    prob.q0 = np.ascontiguousarray(q0, np.float32)
    return rng, amp, ph, prob


def test_hip_adjoint_matches_fp64_adjoint(oracle):
    rng, amp, ph, prob = case()
    dts = np.full(NSTEPS + 5, 0.04)
    J, ga, gp, lam = gpu_adjoint(prob, NSTEPS, dts, BOX)
    op = to_oracle(prob, dt_seq=dts)
    Jo, gao, gpo, gqo, rc = oracle.adjoint(op, prob.q0.astype(np.float64), NSTEPS, BOX)
    assert rc == 0
    e = (abs(J - Jo) / Jo, rel_max(ga, gao), rel_max(gp, gpo),
         max(rel_max(lam[k], gqo[k]) for k in range(3)))
    print(f"[adjoint] J {J:.6e} vs {Jo:.6e}; rel err J {e[0]:.2e} amp {e[1]:.2e} "
          f"phase {e[2]:.2e} state {e[3]:.2e}", flush=True)
    assert e[0] < 1e-5 and e[1] < 1e-3 and e[2] < 1e-3 and e[3] < 1e-3


def test_hip_gradient_matches_fp64_finite_differences(oracle):
    rng, amp, ph, prob = case(5)
    dts = np.full(NSTEPS + 5, 0.04)
    J, ga, gp, lam = gpu_adjoint(prob, NSTEPS, dts, BOX)
    q0 = prob.q0.astype(np.float64)
    worst = 0.0
    for trial in range(16):
        d = rng.standard_normal(NY) * (amp > 1e-8)
        use_amp = trial % 2 == 0
        fds = []
        for h in (1e-5, 1e-6, 1e-7):
            a_p, a_m = (amp + h * d, amp - h * d) if use_amp else (amp, amp)
            p_p, p_m = (ph, ph) if use_amp else (ph + h * d, ph - h * d)
            Ja, _ = oracle.objective(to_oracle(mk(a_p, p_p), dt_seq=dts), q0, NSTEPS, BOX)
            Jb, _ = oracle.objective(to_oracle(mk(a_m, p_m), dt_seq=dts), q0, NSTEPS, BOX)
            fds.append((Ja - Jb) / (2 * h))
        g = ga if use_amp else gp
        err = best_rel(fds, g @ d)
        worst = max(worst, err)
        assert err < 1e-3, (trial, fds, g @ d)
    print(f"[adjoint] 16 FD directions: worst rel err {worst:.2e}", flush=True)


def test_checkpoint_interval_does_not_change_the_gradient():
    rng, amp, ph, prob = case(7)
    dts = np.full(NSTEPS + 5, 0.04)
    ref = gpu_adjoint(prob, NSTEPS, dts, BOX, K=NSTEPS)
    for K in (1, 4, 7):
        out = gpu_adjoint(prob, NSTEPS, dts, BOX, K=K)
        assert abs(out[0] - ref[0]) <= 1e-6 * ref[0]
        assert rel_max(out[1], ref[1]) < 1e-4 and rel_max(out[2], ref[2]) < 1e-4


def test_batched_adjoint_matches_single():
    rng, amp, ph, prob = case(9)
    rng2, amp2, ph2, prob2 = case(11)
    dts = np.full(NSTEPS + 5, 0.04)
    boxes = [BOX, (14, 10, 8, 8)]
    with fa.Context([prob, prob2]) as ctx:
        ctx.set_dt_sequence(0, dts)
        ctx.set_dt_sequence(1, dts)
        J, ga, gp = ctx.step_adjoint(NSTEPS, boxes)
    for k, (p, b) in enumerate(zip([prob, prob2], boxes)):
        Js, gas, gps, _ = gpu_adjoint(p, NSTEPS, dts, b)
        assert abs(J[k] - Js) <= 1e-6 * Js
        assert rel_max(ga[k], gas) < 1e-5 and rel_max(gp[k], gps) < 1e-5


def test_feedback_dt_adjoint_runs_and_plain_forward_agrees():
    # CFL feedback (the dt sequence is recorded during the forward and frozen)
    rng, amp, ph, prob = case(13)
    with fa.Context(prob) as ctx:
        J1, ga1, gp1 = ctx.step_adjoint(30, BOX)
    with fa.Context(prob) as ctx:
        J2, ga2, gp2 = ctx.step_adjoint(30, BOX, plain_forward=True)
    assert np.isfinite(ga1[0]).all() and J1[0] > 0
    assert abs(J1[0] - J2[0]) <= 1e-6 * J1[0]
    assert rel_max(ga1[0], ga2[0]) < 1e-5


@pytest.mark.parametrize("kind", ["zero_gradient", "taylor"])
def test_hip_state_adjoint_other_boundaries_matches_fp64(oracle, kind):
    """dJ/dq0 through the zero-gradient copy rules (box on the left and top
    ghost rows) and through the Taylor closed-form ring."""
    rng = np.random.default_rng(4)
    if kind == "zero_gradient":
        amp, ph, q0 = setup_case(rng)
        prob, box, dt = mk_zero_gradient(q0), (1, 30, 12, 9), 0.04
    else:
        prob, box, dt = mk_taylor(), (20, 40, 16, 12), 0.05
    n = 12
    prob.q0 = np.ascontiguousarray(prob.q0, np.float32)
    dts = np.full(n + 5, dt)
    J, ga, gp, lam = gpu_adjoint(prob, n, dts, box)
    op = to_oracle(prob, dt_seq=dts)
    Jo, gao, gpo, gqo, rc = oracle.adjoint(op, prob.q0.astype(np.float64), n, box)
    assert rc == 0
    e = [abs(J - Jo) / Jo] + [rel_max(lam[k], gqo[k]) for k in range(3)]
    print(f"[adjoint {kind}] J rel err {e[0]:.2e}; dJ/dq0 rel err " +
          " ".join(f"{v:.2e}" for v in e[1:]), flush=True)
    assert e[0] < 1e-5 and max(e[1:]) < 1e-3
