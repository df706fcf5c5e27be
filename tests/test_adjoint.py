"""The FP64 discrete adjoint of the oracle (oracle/fvadjoint.c) against
central finite differences of the oracle forward (FP64, frozen/prescribed dt),
and the HIP adjoint (when built) against both.

Frozen-dt policy: the dt of every step is an input (the dt sequence the
forward took); FD runs replay the same sequence, so dJ/dtheta excludes the
dependence of the CFL-chosen dt on the state.  The step is piecewise smooth
(MC limiter kinks, upwind switches), so each FD check takes the best of three
step sizes.
"""
import numpy as np
import pytest

from fivonagi_amd import workloads as wl
from tests.oracle_bridge import to_oracle

NX, NY, NSTEPS = 48, 40, 25
BOX = (10, 14, 10, 12)            # x0, y0, nx, ny (2 units from the drive columns)


def setup_case(rng):
    amp = 0.02 * np.exp(-((np.arange(NY) - 19.5) / 8.0) ** 2) * (1 + 0.1 * rng.standard_normal(NY))
    ph = 0.3 * rng.standard_normal(NY)
    J_, I_ = np.mgrid[0:NY, 0:NX]
    # a non-quiescent start: at a zero state every wave is 0 and the limited
    # correction is not differentiable (phi jumps from 1 to phi(r))
    q0 = np.zeros((3, NY, NX))
    q0[0] = (2e-3 * np.sin(0.31 * I_ + 0.17 * J_) + 1e-3 * np.cos(0.23 * J_ - 0.11 * I_)
             + 5e-3 * np.exp(-((I_ - 24) ** 2 + (J_ - 20) ** 2) / 30.0))
    q0[1] = 1e-3 * np.cos(0.29 * I_ + 0.13 * J_)
    q0[2] = 1e-3 * np.sin(0.19 * J_ + 0.07 * I_)
    return amp, ph, q0


def mk(amp, ph):
    return wl.edge_drive(NX, NY, edge_amp=amp, edge_phase=ph, cells_per_wavelength=12.0)


def best_rel(fds, ad):
    return min(abs(f - ad) / max(abs(f), 1e-300) for f in fds)


def test_line_formulation_matches_oracle_forward(oracle):
    rng = np.random.default_rng(1)
    amp, ph, q0 = setup_case(rng)
    dts = np.full(6, 0.04)
    op = to_oracle(mk(amp, ph), dt_seq=dts)
    ref, _, rc = oracle.run(op, q0, 6)
    q = q0.copy()
    T = op.T0
    for k in range(6):
        q, rc2 = oracle.line_step(op, q, T, dts[k])
        T += dts[k]
    for f in range(3):
        assert np.abs(q[f] - ref[f]).max() <= 1e-12 * np.abs(ref[f]).max()


def test_adjoint_matches_finite_differences(oracle):
    rng = np.random.default_rng(2)
    amp, ph, q0 = setup_case(rng)
    dts = np.full(NSTEPS, 0.04)
    op = to_oracle(mk(amp, ph), dt_seq=dts)
    J, ga, gp, gq, rc = oracle.adjoint(op, q0, NSTEPS, BOX)
    assert rc == 0
    Jf, rc = oracle.objective(op, q0, NSTEPS, BOX)
    assert abs(J - Jf) <= 1e-12 * abs(Jf)
    assert np.abs(ga).max() > 0 and np.abs(gp).max() > 0
    I_ = np.mgrid[0:NY, 0:NX][1]
    J_ = np.mgrid[0:NY, 0:NX][0]
    for trial in range(3):
        d = rng.standard_normal(NY) * (amp > 1e-8)
        fds_a, fds_p = [], []
        for h in (1e-5, 1e-6, 1e-7):
            Ja, _ = oracle.objective(to_oracle(mk(amp + h * d, ph), dt_seq=dts), q0, NSTEPS, BOX)
            Jb, _ = oracle.objective(to_oracle(mk(amp - h * d, ph), dt_seq=dts), q0, NSTEPS, BOX)
            fds_a.append((Ja - Jb) / (2 * h))
            Ja, _ = oracle.objective(to_oracle(mk(amp, ph + h * d), dt_seq=dts), q0, NSTEPS, BOX)
            Jb, _ = oracle.objective(to_oracle(mk(amp, ph - h * d), dt_seq=dts), q0, NSTEPS, BOX)
            fds_p.append((Ja - Jb) / (2 * h))
        assert best_rel(fds_a, ga @ d) < 1e-3, (fds_a, ga @ d)
        assert best_rel(fds_p, gp @ d) < 1e-3, (fds_p, gp @ d)
        D = np.stack([np.sin(rng.uniform(0.05, 0.4) * I_ + rng.uniform(0, 0.4) * J_
                             + rng.uniform(0, 6)) for _ in range(3)])
        fds_q = []
        for h in (1e-4, 1e-5, 1e-6):
            Ja, _ = oracle.objective(op, q0 + h * D, NSTEPS, BOX)
            Jb, _ = oracle.objective(op, q0 - h * D, NSTEPS, BOX)
            fds_q.append((Ja - Jb) / (2 * h))
        assert best_rel(fds_q, (gq * D).sum()) < 1e-3, (fds_q, (gq * D).sum())


def test_adjoint_of_a_linear_functional_of_the_source(oracle):
    # the ghost ring is read-only input to the source and sweeps: the initial
    # state's ghost cells that no boundary rule reads have zero gradient
    rng = np.random.default_rng(3)
    amp, ph, q0 = setup_case(rng)
    dts = np.full(4, 0.04)
    op = to_oracle(mk(amp, ph), dt_seq=dts)
    J, ga, gp, gq, rc = oracle.adjoint(op, q0, 4, BOX)
    # columns 0, 1 are overwritten by the drive every step: no dependence
    assert not gq[:, :, 0:2].any()


def mk_zero_gradient(q0):
    """The adjoint case on the zero-gradient boundary (every copy rule, corners
    included, transposed)."""
    p = wl.gauss_pulse(NX, NY)
    p.q0 = np.ascontiguousarray(q0)
    return p


def mk_taylor():
    """Taylor shock (closed-form ghost ring) with a small smooth perturbation:
    the bare profile is exactly uniform far from the front (tanh saturates),
    and at a uniform state every wave is 0 and the limited correction is not
    differentiable, so finite differences there measure a one-sided quantity."""
    p = wl.taylor(eta_mult=1, theta=0.3, beta=3.5, delta=1e-3)
    J_, I_ = np.mgrid[0:p.ny, 0:p.nx]
    p.q0 = np.ascontiguousarray(p.q0 + 2e-5 * np.stack([
        np.sin(0.31 * I_ + 0.17 * J_), np.cos(0.29 * I_ + 0.13 * J_),
        np.sin(0.19 * J_ + 0.07 * I_)]))
    return p


@pytest.mark.parametrize("kind", ["zero_gradient", "taylor"])
def test_state_adjoint_matches_finite_differences(oracle, kind):
    rng = np.random.default_rng(4)
    if kind == "zero_gradient":
        amp, ph, q0 = setup_case(rng)
        prob, box, dt = mk_zero_gradient(q0), (1, 30, 12, 9), 0.04
    else:
        prob, box, dt = mk_taylor(), (20, 40, 16, 12), 0.05
        q0 = prob.q0.astype(np.float64)
    n = 12
    dts = np.full(n, dt)
    op = to_oracle(prob, dt_seq=dts)
    J, ga, gp, gq, rc = oracle.adjoint(op, q0, n, box)
    assert rc == 0 and J > 0 and np.abs(gq).max() > 0
    J_, I_ = np.mgrid[0:prob.ny, 0:prob.nx]
    for trial in range(3):
        D = np.stack([np.sin(rng.uniform(0.05, 0.4) * I_ + rng.uniform(0, 0.4) * J_
                             + rng.uniform(0, 6)) for _ in range(3)])
        fds = []
        for h in (1e-4, 1e-5, 1e-6):
            s = h * np.abs(q0).max() if kind == "taylor" else h
            Ja, _ = oracle.objective(op, q0 + s * D, n, box)
            Jb, _ = oracle.objective(op, q0 - s * D, n, box)
            fds.append((Ja - Jb) / (2 * s))
        assert best_rel(fds, (gq * D).sum()) < 1e-3, (kind, fds, (gq * D).sum())
    if kind == "taylor":
        # every ghost cell is the closed form: no dependence on q0 there
        ghost = np.zeros((prob.ny, prob.nx), bool)
        ghost[:2, :] = ghost[-2:, :] = ghost[:, :2] = ghost[:, -2:] = True
        assert not gq[:, ghost].any()


def test_adjoint_along_a_given_trajectory_is_the_adjoint(oracle):
    """orc_adjoint_path_f64 on the trajectory the host adjoint's own forward
    computes (line_step) is orc_adjoint_f64, bit for bit: the two share the
    reverse sweep, and the path form only takes the states as given."""
    rng = np.random.default_rng(3)
    amp, ph, q0 = setup_case(rng)
    n = 9
    dts = np.full(n + 2, 0.04)
    op = to_oracle(mk(amp, ph), dt_seq=dts)
    states = [q0.copy()]
    T = op.T0
    for k in range(n):
        q, rc = oracle.line_step(op, states[-1], T, dts[k])
        assert rc == 0
        states.append(q)
        T += dts[k]
    J, ga, gp, gq, rc = oracle.adjoint(op, q0, n, BOX)
    J2, ga2, gp2, gq2, rc2 = oracle.adjoint_path(op, np.array(states), BOX)
    assert rc == 0 and rc2 == 0
    assert J == J2 and np.array_equal(ga, ga2) and np.array_equal(gp, gp2)
    assert np.array_equal(gq, gq2)
    # a different trajectory (a perturbed start, stepped the same way) gives a
    # different linearisation: the states are used, not recomputed
    states[0] = states[0] * (1 + 1e-3)
    J3, ga3, _, gq3, _ = oracle.adjoint_path(op, np.array(states), BOX)
    assert not np.array_equal(gq3, gq)
