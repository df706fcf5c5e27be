"""HIP adjoint (fivonagi_step_adjoint) against the FP64 host adjoint
(oracle/fvadjoint.c) and FP64 central finite differences of the oracle
forward, with the same prescribed (frozen) dt sequence.

Tolerances: J within 1e-5 and the drive gradients within 1e-3 (relative,
max-norm) of the FP64 adjoint of the oracle's own forward run; directional
derivatives within 1e-3 of the FP64 finite differences (best of three step
sizes: the step is piecewise smooth); dJ/dq0 within 1e-5 of the FP64 adjoint
taken along the GPU's own trajectory (orc_adjoint_path_f64).  The step is
only piecewise differentiable: where a limiter ratio sits at a kink to within
rounding (e.g. a wave of a nearly flat region, at the FP32 noise level of the
state, as the upwind neighbour), the FP32 and FP64 runs take different
branches and each gradient is exact for its own path; the path form compares
the GPU with FP64 on the GPU's branches."""
import numpy as np
import pytest

import fivonagi_amd as fa
from tests.oracle_bridge import to_oracle
from tests.test_adjoint import (BOX, NSTEPS, NX, NY, best_rel, mk, mk_taylor, mk_zero_gradient,
                                setup_case)

pytestmark = pytest.mark.gpu


def gpu_adjoint(prob, nsteps, dts, box, K=0, plain=False, check=False):
    with fa.Context(prob) as ctx:
        ctx.set_dt_sequence(0, dts)
        J, ga, gp = ctx.step_adjoint(nsteps, box, checkpoint_every=K, plain_forward=plain,
                                     check_recompute=check)
        lam = ctx.adjoint_state(0)
    return J[0], ga[0], gp[0], lam


def rel_max(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def gpu_trajectory(prob, nsteps, dts):
    """q_0 .. q_nsteps of the GPU forward with the prescribed dt (the run the
    adjoint differentiates: the same kernel and dt, so the same states)."""
    with fa.Context(prob) as ctx:
        ctx.set_dt_sequence(0, dts)
        qs = [ctx.collect(0)[0].astype(np.float64)]
        for _ in range(nsteps):
            ctx.step(1)
            qs.append(ctx.collect(0)[0].astype(np.float64))
    return np.array(qs)


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
    Jp, gap, gpp, gqp, rc = oracle.adjoint_path(op, gpu_trajectory(prob, NSTEPS, dts), BOX)
    assert rc == 0
    e = (abs(J - Jo) / Jo, rel_max(ga, gao), rel_max(gp, gpo),
         max(rel_max(lam[k], gqp[k]) for k in range(3)),
         max(rel_max(lam[k], gqo[k]) for k in range(3)))
    print(f"[adjoint] J {J:.6e} vs {Jo:.6e}; rel err J {e[0]:.2e} amp {e[1]:.2e} "
          f"phase {e[2]:.2e} state {e[3]:.2e} along the GPU path ({e[4]:.2e} against the "
          f"FP64 run's own path)", flush=True)
    assert e[0] < 1e-5 and e[1] < 1e-3 and e[2] < 1e-3 and e[3] < 1e-5
    # along the GPU path the drive gradients agree as closely
    assert rel_max(ga, gap) < 1e-5 and rel_max(gp, gpp) < 1e-5


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
    """Segments are recomputed by the step kernel from the control blocks the
    forward steps ran with, so every checkpoint interval -- K = 1 keeps every
    state of the forward run, K = NSTEPS recomputes all but the first --
    gives the same J, gradients and dJ/dq0, bit for bit."""
    rng, amp, ph, prob = case(7)
    dts = np.full(NSTEPS + 5, 0.04)
    ref = gpu_adjoint(prob, NSTEPS, dts, BOX, K=NSTEPS)
    for K in (1, 4, 7):
        out = gpu_adjoint(prob, NSTEPS, dts, BOX, K=K)
        assert out[0] == ref[0], K
        assert np.array_equal(out[1], ref[1]) and np.array_equal(out[2], ref[2]), K
        assert np.array_equal(out[3], ref[3]), K


@pytest.mark.parametrize("kind", ["edge_drive", "zero_gradient", "taylor", "feedback"])
def test_reverse_pass_recomputes_the_forward_states_bitwise(kind):
    """check_recompute: the reverse kernel recomputes every q_{n+1} from q_n
    with the step's own functions and control block and compares it with the
    state the forward (K = 1: the forward run's own states) or the segment
    recompute (K = n: the step kernel replaying logged control blocks)
    stored; any differing bit fails the call (ERR_CHECK).  Covers every
    limiter and upwind branch the run takes: the reverse pass differentiates
    the executed path."""
    rng = np.random.default_rng(21)
    n = 18
    if kind == "taylor":
        prob, box = mk_taylor(), (20, 40, 16, 12)
    elif kind == "zero_gradient":
        amp, ph, q0 = setup_case(rng)
        prob, box = mk_zero_gradient(q0), (1, 30, 12, 9)
    else:
        amp, ph, q0 = setup_case(rng)
        prob, box = mk(amp, ph), BOX
        prob.q0 = np.ascontiguousarray(q0, np.float32)
    prob.q0 = np.ascontiguousarray(prob.q0, np.float32)
    for K in (1, n):
        with fa.Context(prob) as ctx:
            if kind != "feedback":
                ctx.set_dt_sequence(0, np.full(n + 5, 0.05 if kind == "taylor" else 0.04))
            J, ga, gp = ctx.step_adjoint(n, box, checkpoint_every=K, check_recompute=True)
        assert np.isfinite(J[0]) and J[0] > 0, (kind, K)


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
        # J sums the work items' shares, which differ between the two splits
        assert abs(J[k] - Js) <= 1e-12 * Js
        assert np.array_equal(ga[k], gas) and np.array_equal(gp[k], gps)


def test_feedback_dt_adjoint_runs_and_plain_forward_agrees():
    # CFL feedback (the dt sequence is recorded during the forward and frozen)
    rng, amp, ph, prob = case(13)
    with fa.Context(prob) as ctx:
        J1, ga1, gp1 = ctx.step_adjoint(30, BOX)
    with fa.Context(prob) as ctx:
        J2, ga2, gp2 = ctx.step_adjoint(30, BOX, plain_forward=True)
    assert np.isfinite(ga1[0]).all() and J1[0] > 0
    assert J1[0] == J2[0] and np.array_equal(ga1[0], ga2[0])


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
    Jp, _, _, gqp, rc = oracle.adjoint_path(op, gpu_trajectory(prob, n, dts), box)
    assert rc == 0
    e = [abs(J - Jo) / Jo] + [rel_max(lam[k], gqp[k]) for k in range(3)]
    eo = max(rel_max(lam[k], gqo[k]) for k in range(3))
    print(f"[adjoint {kind}] J rel err {e[0]:.2e}; dJ/dq0 rel err " +
          " ".join(f"{v:.2e}" for v in e[1:]) + f" along the GPU path ({eo:.2e} against the "
          "FP64 run's own path)", flush=True)
    assert e[0] < 1e-5 and max(e[1:]) < 1e-5 and eo < 1e-3


def test_mixed_batch_adjoint_matches_single_instances():
    """One batch of three apps (edge drive, zero gradient, Taylor; three grid
    sizes): the reverse kernel dispatches per work item on the boundary kind.
    Every state kept (K = 1) and segments recomputed (K = 5) agree bit for
    bit, the recompute check passes, and each instance's gradient and dJ/dq0
    are those of the instance run alone (the batch steps it with the kernel
    that carries every app, alone it gets its app's own kernel)."""
    rng = np.random.default_rng(31)
    amp, ph, q0 = setup_case(rng)
    p_ed = mk(amp, ph)
    p_ed.q0 = np.ascontiguousarray(q0, np.float32)
    rng2 = np.random.default_rng(32)
    _, _, q0b = setup_case(rng2)
    p_zg = mk_zero_gradient(q0b)
    p_zg.q0 = np.ascontiguousarray(p_zg.q0, np.float32)
    p_ty = mk_taylor()
    p_ty.q0 = np.ascontiguousarray(p_ty.q0, np.float32)
    probs = [p_ed, p_zg, p_ty]
    boxes = [BOX, (1, 30, 12, 9), (20, 40, 16, 12)]
    dts = [0.04, 0.04, 0.05]
    n = 14
    out = {}
    for K in (1, 5):
        with fa.Context(probs) as ctx:
            for k, dt in enumerate(dts):
                ctx.set_dt_sequence(k, np.full(n + 3, dt))
            J, ga, gp = ctx.step_adjoint(n, boxes, checkpoint_every=K, check_recompute=True)
            lam = [ctx.adjoint_state(k) for k in range(3)]
        out[K] = (J, ga, gp, lam)
    for k in range(3):
        assert out[1][0][k] == out[5][0][k], k
        assert np.array_equal(out[1][1][k], out[5][1][k]) and np.array_equal(out[1][3][k], out[5][3][k])
    J, ga, gp, lam = out[1]
    for k, (p, b, dt) in enumerate(zip(probs, boxes, dts)):
        Js, gas, gps, lams = gpu_adjoint(p, n, np.full(n + 3, dt), b)
        assert J[k] > 0 and abs(J[k] - Js) <= 1e-12 * Js, (k, J[k], Js)
        assert np.array_equal(ga[k], gas) and np.array_equal(gp[k], gps), k
        assert np.array_equal(lam[k], lams), k


def test_adjoint_with_rejected_attempts_in_the_forward():
    """CFL feedback from a first dt several times too long: the forward rejects
    attempts, and the control-block log indexes accepted steps only.  Every
    state kept and segments recomputed give the same result bit for bit, the
    recompute check passes, and J equals the sum the forward's own states give."""
    rng, amp, ph, prob = case(17)
    prob.dt0 = 0.2
    n = 20
    res = {}
    for K in (1, 3, n):
        with fa.Context(prob) as ctx:
            J, ga, gp = ctx.step_adjoint(n, BOX, checkpoint_every=K, check_recompute=True)
            info = ctx.info(0)
            lam = ctx.adjoint_state(0)
        res[K] = (J[0], ga[0], gp[0], lam, info["n_rejected"], info["n_accepted"])
    assert res[1][4] > 0 and res[1][5] == n
    for K in (3, n):
        assert res[K][0] == res[1][0] and np.array_equal(res[K][1], res[1][1])
        assert np.array_equal(res[K][3], res[1][3])
    # J from the forward's own states, step by step, with the dt it chose
    x0, y0, bx, by = BOX
    with fa.Context(prob) as ctx:
        Jf, T_prev = 0.0, ctx.info(0)["T"]
        for _ in range(n):
            ctx.step(1)
            q, st = ctx.collect(0)
            dt = st["T"] - T_prev
            T_prev = st["T"]
            Jf += dt * float(np.sum(q[0, y0:y0 + by, x0:x0 + bx].astype(np.float64) ** 2))
    assert abs(res[1][0] - Jf) <= 1e-9 * Jf, (res[1][0], Jf)


def test_repeated_calls_reuse_the_context_block():
    """Adjoint calls one after another on one context (every state kept,
    checkpoints, a longer run that grows the kept block, a shorter one inside
    it) against each call on a fresh context brought to the same state by
    plain steps first: J, the drive gradients and dJ/dq0 bit for bit, so
    nothing of an earlier call's buffers reaches a later one.  CFL feedback
    (no prescribed dt), so the plain steps and the adjoint's forward take the
    same dt."""
    _, _, _, prob = case()
    n = 12
    calls = [(n, 1), (n, 4), (2 * n, 1), (n, 1)]
    got = []
    with fa.Context(prob) as ctx:
        for steps, K in calls:
            J, ga, gp = ctx.step_adjoint(steps, BOX, checkpoint_every=K)
            got.append((J[0], ga[0], gp[0], ctx.adjoint_state(0)))
    before = 0
    for (steps, K), (J, ga, gp, lam) in zip(calls, got):
        with fa.Context(prob) as ctx:
            if before:
                ctx.step(before)
            Jr, gar, gpr = ctx.step_adjoint(steps, BOX, checkpoint_every=K)
            lamr = ctx.adjoint_state(0)
        assert J == Jr[0], (steps, K)
        assert np.array_equal(ga, gar[0]) and np.array_equal(gp, gpr[0]), (steps, K)
        assert all(np.array_equal(a, b) for a, b in zip(lam, lamr)), (steps, K)
        before += steps


def test_waves_of_several_items_match_one_item_per_wave():
    """At 2560x2048 the automatic reverse split gives waves of about equal
    cost, some of which end one strip and start the next (two items in one
    wave: 30 of the ~2000 waves on a 256-CU device); a fixed height of 24 rows
    gives one item per wave.  Both runs of 12 steps, with the recompute check,
    give the same gradients and dJ/dq0 bit for bit: the grouping of items
    into waves changes nothing.  J is summed per item, so its rounding follows
    the item split (within 1e-14)."""
    from fivonagi_amd import workloads as wl
    prob, _, _ = wl.adjoint_single(nx=2560, ny=2048)
    box = (2, 2048 // 2 - 32, 64, 64)          # at the drive, so J > 0 after 12 steps
    n = 12
    res = []
    for rows in (0, 24):
        with fa.Context(prob) as ctx:
            ctx.set_dt_sequence(0, np.full(n + 3, 0.04))
            J, ga, gp = ctx.step_adjoint(n, box, check_recompute=True, rows_per_wave=rows)
            res.append((J[0], ga[0], gp[0], ctx.adjoint_state(0)))
    assert res[0][0] > 0 and abs(res[0][0] - res[1][0]) <= 1e-14 * res[0][0]
    assert np.abs(res[0][1]).max() > 0
    assert np.array_equal(res[0][1], res[1][1]) and np.array_equal(res[0][2], res[1][2])
    assert np.array_equal(res[0][3], res[1][3])
