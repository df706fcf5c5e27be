"""Pins of the CPU oracle (oracle/fvoracle.c) against facts the reference's
mathematics fixes -- closed forms, identities, invariants and hand-computed
tables -- chosen so that a plausible transcription slip (a dropped term, a
wrong sign, index or operand) fails one of them.  The reference ships no test
vectors of its own (it has no tests), so these are the pins.

Which reference code each group pins:
  Riemann solve      apps/na/fv.h:88-180 (x), :276-368 (y)
  MC limiter         apps/na/fv.h:579-662
  boundary rules     common/boundary.h:27-81, apps/hifu-beam/boundary.h:77-93,
                     apps/taylor-angle-xy/boundary.h:38-55
  source             apps/na/source.h:84-104
  CFL fold           common/main.cu:179-212, :503-524
  whole step         common/main.cu:310-635
"""
import math

import numpy as np
import pytest

from fivonagi_amd import workloads as wl
from tests.oracle_bridge import to_oracle

BETA = 3.5


# ----------------------------------------------------------------- Riemann --

def test_rest_state_speeds(oracle):
    rc, r = oracle.riemann([0, 0, 0], [0, 0, 0], BETA)
    assert rc == 0
    np.testing.assert_array_equal(r["s"], [1.0, -1.0, 0.0])
    assert not r["W"].any() and not r["amdq"].any() and not r["apdq"].any()


@pytest.mark.parametrize("rho1", [-0.1, 0.01, 0.3])
def test_still_medium_sound_speed(oracle, rho1):
    # uniform density, no flow: L1,2 = +-sqrt(1 + 2(beta-1) rho'), L3 = 0
    rc, r = oracle.riemann([rho1, 0, 0], [rho1, 0, 0], BETA)
    c = math.sqrt(1 + 2 * (BETA - 1) * rho1)
    np.testing.assert_allclose(r["s"], [c, -c, 0.0], rtol=1e-14, atol=1e-15)
    assert not r["W"].any()   # uniform state -> no waves


@pytest.mark.parametrize("U", [-0.6, 0.25, 0.9])
def test_uniform_flow_speeds(oracle, U):
    # rho = 1 moving at U: tmp3 = 1 - U^2, L = U +- sqrt(1 - U^2), L3 = U
    rc, r = oracle.riemann([0, U, 0], [0, U, 0], BETA)
    s = math.sqrt(1 - U * U)
    np.testing.assert_allclose(r["s"], [U + s, U - s, U], rtol=1e-14, atol=1e-15)


def test_supersonic_is_imaginary_eigenvalue(oracle):
    # rho = 1 at U = 1.5: tmp3 = 1 - U^2 < 0 -> simError = 2 (fv.h:102)
    rc, _ = oracle.riemann([0, 1.5, 0], [0, 1.5, 0], BETA)
    assert rc == 2


@pytest.mark.parametrize("v", [0.05, -0.2])
def test_tangential_eigenvector_closed_form(oracle, v):
    # rho' = 0, u = 0, v uniform: Z1 = 1, Z2 = 0, Z3 = v, L = +-1, so
    # A1 = v/(1 - v), A2 = v/(1 + v): W31/W11 and W32/W12 (fv.h:110-113)
    rc, r = oracle.riemann([0.0, 0.0, v], [0.001, 0.0, v * 1.001], BETA)
    # tiny jump so the waves exist; Z's move by O(1e-3)
    W = r["W"]
    A1, A2 = W[2, 0] / W[0, 0], W[2, 1] / W[0, 1]
    assert abs(A1 - v / (1 - v)) < 5e-3 * abs(v) and abs(A2 - v / (1 + v)) < 5e-3 * abs(v)
    rc, r = oracle.riemann([0.0, 0.0, v], [0.0, 1e-9, v], BETA)
    W = r["W"]
    np.testing.assert_allclose(W[2, 0] / W[0, 0], v / (1 - v), rtol=1e-6)
    np.testing.assert_allclose(W[2, 1] / W[0, 1], v / (1 + v), rtol=1e-6)


def random_pair(rng):
    qL = [rng.uniform(-0.05, 0.05), rng.uniform(-0.05, 0.05), rng.uniform(-0.05, 0.05)]
    qR = [rng.uniform(-0.05, 0.05), rng.uniform(-0.05, 0.05), rng.uniform(-0.05, 0.05)]
    return qL, qR


def test_waves_sum_to_the_jump(oracle):
    # sum_p W_p = q_R - q_L component-wise (fv.h:116-139 substituted)
    rng = np.random.default_rng(7)
    for _ in range(200):
        qL, qR = random_pair(rng)
        for form in (oracle.FORM_LITERAL, oracle.FORM_RHO1):
            rc, r = oracle.riemann(qL, qR, BETA, form)
            assert rc == 0
            np.testing.assert_allclose(r["W"].sum(axis=1), np.subtract(qR, qL), rtol=1e-9,
                                       atol=1e-15)


def test_waves_solve_the_eigensystem(oracle):
    # alpha solves [r1 r2 r3] alpha = dq with r_p = (1, L_p, A_p), r3 = (0,0,1):
    # an independent linear solve of the same system
    rng = np.random.default_rng(8)
    for _ in range(100):
        qL, qR = random_pair(rng)
        rc, r = oracle.riemann(qL, qR, BETA)
        L1, L2 = r["s"][0], r["s"][1]
        W = r["W"]
        a1, a2, a3 = W[0, 0], W[0, 1], W[2, 2]
        A1, A2 = W[2, 0] / a1, W[2, 1] / a2
        R = np.array([[1, 1, 0], [L1, L2, 0], [A1, A2, 1.0]])
        alpha = np.linalg.solve(R, np.subtract(qR, qL))
        np.testing.assert_allclose([a1, a2, a3], alpha, rtol=1e-8, atol=1e-14)
        np.testing.assert_allclose(W[1, 0], a1 * L1, rtol=1e-14)
        np.testing.assert_allclose(W[1, 1], a2 * L2, rtol=1e-14)
        assert W[0, 2] == 0.0 and W[1, 2] == 0.0


def test_fluctuations_route_by_sign(oracle):
    rng = np.random.default_rng(9)
    for _ in range(100):
        qL, qR = random_pair(rng)
        rc, r = oracle.riemann(qL, qR, BETA)
        s, W = r["s"], r["W"]
        am = sum(s[p] * W[:, p] for p in range(3) if s[p] < 0)
        ap = sum(s[p] * W[:, p] for p in range(3) if s[p] > 0)
        np.testing.assert_allclose(r["amdq"], am, rtol=1e-13, atol=1e-17)
        np.testing.assert_allclose(r["apdq"], ap, rtol=1e-13, atol=1e-17)
        # mass flux: amdq1 + apdq1 = jump of the normal momentum
        np.testing.assert_allclose(r["amdq"][0] + r["apdq"][0], qR[1] - qL[1], rtol=1e-9,
                                   atol=1e-16)


def test_zero_speed_wave_goes_nowhere(oracle):
    # u = 0 on both sides: L3 = 0; a tangential jump is all wave 3 and
    # appears in neither fluctuation
    rc, r = oracle.riemann([0, 0, 0.0], [0, 0, 0.01], BETA)
    assert r["s"][2] == 0.0 and abs(r["W"][2, 2] - 0.01) < 1e-15
    assert r["amdq"][2] == 0.0 and r["apdq"][2] == 0.0


# ---------------------------------------------------------------- limiter --

@pytest.mark.parametrize("r,phi", [(-1.0, 0.0), (0.0, 0.0), (0.2, 0.4), (1 / 3, 2 / 3),
                                   (0.5, 0.75), (1.0, 1.0), (2.0, 1.5), (3.0, 2.0),
                                   (10.0, 2.0)])
def test_mc_limiter_table(oracle, r, phi):
    # hand-computed MC values; wnorm2 = 1 so r = dot
    assert oracle.philim(1.0, r, 99.0, +1.0) == pytest.approx(phi, abs=1e-15)   # s > 0 -> left
    assert oracle.philim(1.0, 99.0, r, -1.0) == pytest.approx(phi, abs=1e-15)   # s < 0 -> right
    assert oracle.philim(1.0, 99.0, r, 0.0) == pytest.approx(phi, abs=1e-15)    # s = 0 -> right


def test_mc_limiter_vanishing_wave(oracle):
    assert oracle.philim(0.0, 5.0, -5.0, 1.0) == 1.0


# --------------------------------------------------------------- boundary --

def ids(nx, ny):
    q = np.zeros((3, ny, nx))
    J, I = np.mgrid[0:ny, 0:nx]
    q[0] = I + 1000.0 * J
    q[1] = -(I + 1000.0 * J)
    q[2] = 0.5 * (I + 1000.0 * J)
    return q


def test_zero_gradient_index_map(oracle):
    nx, ny = 12, 10
    p = wl.gauss_pulse(nx, ny)
    out, rc = oracle.boundary(to_oracle(p), ids(nx, ny), 0.0)
    src = out[0]
    # (i, j) -> source cell, later rules of common/boundary.h win:
    table = {
        (5, 5): (5, 5),                       # interior untouched
        (0, 5): (2, 5), (1, 5): (2, 5),       # left
        (nx - 1, 5): (nx - 3, 5), (nx - 2, 5): (nx - 3, 5),   # right
        (5, 0): (5, 2), (5, 1): (5, 2),       # bottom
        (5, ny - 1): (5, ny - 3), (5, ny - 2): (5, ny - 3),   # top
        (0, 0): (2, 0), (1, 1): (2, 1),       # left beats bottom, keeps its row
        (0, ny - 1): (2, ny - 1),             # left beats top
        (nx - 1, 0): (nx - 1, 2),             # bottom beats right
        (nx - 2, ny - 1): (nx - 2, ny - 3),   # top beats right
    }
    for (i, j), (si, sj) in table.items():
        assert src[j, i] == si + 1000.0 * sj, (i, j)
    assert np.array_equal(out[1], -out[0]) and np.array_equal(out[2], 0.5 * out[0])


def test_edge_drive_values(oracle):
    nx, ny = 16, 12
    amp = np.linspace(0.001, 0.02, ny)
    ph = np.linspace(0.0, 1.0, ny)
    p = wl.edge_drive(nx, ny, edge_amp=amp, edge_phase=ph)
    T = 0.37
    out, rc = oracle.boundary(to_oracle(p), ids(nx, ny), T)
    pk = amp * np.sin(p.omega * T + ph)
    for i in (0, 1):
        np.testing.assert_allclose(out[0][:, i], pk, rtol=1e-15)
        np.testing.assert_allclose(out[1][:, i], pk * (pk + 1), rtol=1e-15)
        assert not out[2][:, i].any()
    # the other edges stay zero-gradient
    assert out[0][5, nx - 1] == (nx - 3) + 1000.0 * 5


def test_hifu_index_map(oracle):
    # apps/hifu-beam/boundary.h:20-133: right/top zero-gradient, left ghost
    # copies column 2 outside the aperture, rows 0,1 mirror rows 4,3 with rho v
    # negated (beam axis), the 2x2 corners are 0
    nx, ny = 14, 12
    p = wl.hifu(nx=nx, ny=ny, ispam=0.0)          # no aperture rows
    out, rc = oracle.boundary(to_oracle(p), ids(nx, ny), 0.0)
    assert rc == 0
    table = {
        (5, 5): (5, 5),                            # interior untouched
        (0, 5): (2, 5), (1, 7): (2, 7),            # left, outside the aperture
        (nx - 1, 5): (nx - 3, 5), (nx - 2, 6): (nx - 3, 6),   # right
        (5, ny - 1): (5, ny - 3), (6, ny - 2): (6, ny - 3),   # top
        (5, 0): (5, 4), (6, 1): (6, 3),            # mirrored beam axis
        (2, 0): (2, 4), (nx - 3, 1): (nx - 3, 3),
    }
    for (i, j), (si, sj) in table.items():
        assert out[0][j, i] == si + 1000.0 * sj, (i, j)
        assert out[1][j, i] == -(si + 1000.0 * sj), (i, j)
        sign = -1.0 if j <= 1 else 1.0
        assert out[2][j, i] == sign * 0.5 * (si + 1000.0 * sj), (i, j)
    for (i, j) in [(0, 0), (1, 1), (0, ny - 1), (1, ny - 2), (nx - 1, 0), (nx - 2, 1),
                   (nx - 1, ny - 1), (nx - 2, ny - 2)]:
        assert out[0][j, i] == out[1][j, i] == out[2][j, i] == 0.0, (i, j)


def test_hifu_packet_special_values(oracle):
    # packet(T, phi) = AMPL sin(ISPW (T + phi)) exp(-(FREC (T+phi)/4)^10) on
    # rows 2 <= j < 2 + int(ISPa/dy); nu = (F, -(j-2))/sqrt((j-2)^2 + F^2),
    # phi = dx (j-2)^2 / (2F) (apps/hifu-beam/boundary.h:10-17, 79-93)
    nx, ny = 16, 64
    p = wl.hifu(nx=nx, ny=ny, ispam=10.0)
    h = p.hifu
    width = int(h["a"] / p.dy)
    assert 2 + width < ny - 2
    # T = 1/4 (FREC = 1): row 2 (phi = 0) sits at the crest, sin = 1,
    # envelope exp(-(1/16)^10), direction (1, 0)
    out, rc = oracle.boundary(to_oracle(p), np.zeros((3, ny, nx)), 0.25)
    pk = h["ampl"] * math.exp(-(1.0 / 16.0) ** 10)
    for i in (0, 1):
        assert out[0][2, i] == pytest.approx(pk, rel=1e-15)
        assert out[1][2, i] == pytest.approx(pk * (pk + 1), rel=1e-15)
        assert out[2][2, i] == 0.0
    # row j: the packet is 0 when T + phi_j = 0 (sin 0), whatever the row
    for j in (3, 7, 2 + width - 1):
        phi = p.dx * (j - 2) ** 2 / (2 * h["f"])
        out, rc = oracle.boundary(to_oracle(p), np.zeros((3, ny, nx)), -phi)
        assert abs(out[0][j, 0]) < 1e-18
        assert abs(out[0][j - 1, 0]) > 1e-9       # the neighbour row is off its zero
    # direction: a unit vector tilted towards the axis, rho v / rho u = -(j-2)/F
    out, rc = oracle.boundary(to_oracle(p), np.zeros((3, ny, nx)), 0.3)
    for j in (3, 4):                     # rows whose packet is inside its envelope
        assert out[2][j, 1] / out[1][j, 1] == pytest.approx(-(j - 2) / h["f"], rel=1e-12)
    # first row past the aperture: plain copy of column 2 (zero here)
    assert out[0][2 + width, 0] == 0.0 and out[0][2 + width - 1, 0] != 0.0
    # once the window has moved, no packet (boundary.h:82): with the window
    # moving from the start, a run that has shifted equals one whose aperture
    # is empty from the shift on (compared on the step after the first shift)


def test_taylor_ghosts_closed_form(oracle):
    p = wl.taylor(1, theta=math.pi / 8)
    t = p.taylor
    T = 1.25
    out, rc = oracle.boundary(to_oracle(p), np.zeros((3, p.ny, p.nx)), T)
    J, I = np.mgrid[0:p.ny, 0:p.nx]
    exact = t["rhoa"] * np.tanh(t["c"] * (I * p.dx * math.cos(t["theta"])
                                          + ((J - t["jc"]) * p.dy + t["yoff"])
                                          * math.sin(t["theta"]) - T))
    ghost = (I <= 1) | (J <= 1) | (I >= p.nx - 2) | (J >= p.ny - 2)
    np.testing.assert_allclose(out[0][ghost], exact[ghost], rtol=1e-14, atol=1e-22)
    np.testing.assert_allclose(out[1][ghost], exact[ghost] * math.cos(t["theta"]), rtol=1e-14,
                               atol=1e-22)
    assert not out[0][~ghost].any()


# ----------------------------------------------------------------- source --

def test_source_quadratic_velocity_closed_form(oracle):
    # rho' = 0, rho u = (i dx)^2 + (j dy)^2: the discrete Laplacian of u is
    # exactly 2 + 2, so q2 gains 4*delta*dt in the interior; q1, q3 untouched
    nx, ny = 20, 14
    p = wl.gauss_pulse(nx, ny, delta=1e-3)
    J, I = np.mgrid[0:ny, 0:nx]
    q = np.zeros((3, ny, nx))
    q[1] = (I * p.dx) ** 2 + (J * p.dy) ** 2
    q[2] = 0.1
    dt = 0.01
    out, rc = oracle.source(to_oracle(p), q, dt)
    inner = (slice(2, ny - 2), slice(2, nx - 2))
    np.testing.assert_allclose(out[1][inner] - q[1][inner], 4 * p.delta * dt, rtol=1e-9)
    assert np.array_equal(out[2], q[2]) and np.array_equal(out[0], q[0])
    ring = np.ones((ny, nx), bool)
    ring[inner] = False
    assert np.array_equal(out[1][ring], q[1][ring])


def test_source_divides_by_density(oracle):
    # rho' = 1 (rho = 2) everywhere: velocity is q2/2
    nx, ny = 12, 12
    p = wl.gauss_pulse(nx, ny, delta=1e-3)
    I = np.mgrid[0:ny, 0:nx][1]
    q = np.zeros((3, ny, nx))
    q[0] = 1.0
    q[1] = 2.0 * (I * p.dx) ** 2
    out, rc = oracle.source(to_oracle(p), q, 0.02)
    np.testing.assert_allclose(out[1][2:-2, 2:-2] - q[1][2:-2, 2:-2], 2 * p.delta * 0.02,
                               rtol=1e-9)


# -------------------------------------------------------------- CFL fold --

@pytest.mark.parametrize("nx,ny", [(64, 64), (128, 32), (4096, 1024), (16, 8)])
def test_fold_subset_is_even_offsets_for_powers_of_two(oracle, nx, ny):
    # every stride of the schedule is an even power of two, so an offset keeps
    # its parity and reaches 0 exactly when it is even
    m = oracle.cfl_subset(nx, ny).ravel()
    np.testing.assert_array_equal(m, (np.arange(nx * ny) % 2 == 0).astype(np.uint8))


def test_fold_subset_400(oracle):
    m = oracle.cfl_subset(400, 400)
    frac = m.mean()
    assert 0.40 < frac < 0.42 and m[0, 0] == 1


def _fold_brute_force(nx, ny):
    """common/main.cu:503-524 run on sets: slot c starts as {c}; every
    reduceCFL launch of stride S (:179-197) merges slot i + S into slot i for
    i < S (slots past the grid hold no cell); cfl[0] ends as the subset."""
    N = nx * ny
    slots = [{c} for c in range(N)]

    def fold(S):
        for i in range(S):
            if i + S < N:
                slots[i] = slots[i] | slots[i + S]
    size = N
    while size > 32:
        size //= 2
        fold((size // 16) * 16)
    while size > 2:
        size //= 2
        fold(size)
    m = np.zeros(N, np.uint8)
    m[sorted(slots[0])] = 1
    return m.reshape(ny, nx)


@pytest.mark.parametrize("nx,ny", [(8, 8), (117, 11), (59, 8), (9, 13), (8, 200), (33, 17),
                                   (100, 72), (130, 40), (64, 64)])
def test_fold_subset_against_brute_force(oracle, nx, ny):
    # the oracle's backward walk and the library's host table both equal a
    # direct simulation of the reference's folds (the library's first table
    # followed each cell on one path and lost the slots a later fold re-reads)
    import fivonagi_amd as fa
    ref = _fold_brute_force(nx, ny)
    np.testing.assert_array_equal(oracle.cfl_subset(nx, ny), ref)
    np.testing.assert_array_equal(fa.cfl_subset(nx, ny), ref)


# ------------------------------------------------------------- whole step --

def run(oracle, prob, n, form=None, dtype=np.float64):
    op = to_oracle(prob, form=oracle.FORM_LITERAL if form is None else form)
    q, st, rc = oracle.run(op, prob.q0.astype(np.float64), n, dtype=dtype)
    assert rc == 0
    return q, st


def test_uniform_state_is_a_fixed_point(oracle):
    p = wl.gauss_pulse(32, 24, amp=0.0)
    q0 = np.zeros((3, 24, 32))
    q0[0] = 0.01
    p.q0 = q0.astype(np.float32)
    q, st = run(oracle, p, 10)
    assert np.array_equal(q, q0.astype(np.float32).astype(np.float64))


def test_mass_conservation_away_from_edges(oracle):
    # the density flux is exactly the momentum jump and the source leaves
    # rho' alone, so sum(rho') only changes through edge interfaces
    p = wl.gauss_pulse(160, 160, sigma_cells=4.0)
    m0 = p.q0[0].astype(np.float64).sum()
    q, st = run(oracle, p, 15)
    assert np.abs(q[0][:4]).max() < 1e-20 and np.abs(q[0][:, :4]).max() < 1e-20
    # rounding of ~4e5 cell updates; a wrong flux term moves it by O(1e-3)
    assert abs(q[0].sum() - m0) < 1e-12 * abs(m0)


def test_mirror_symmetry_small_amplitude(oracle):
    # i <-> nx-1-i maps the scheme onto itself with u -> -u (exact up to the
    # O(amplitude^2) asymmetry of the reference's A_k); at 1e-6 amplitude the
    # residual is rounding-level
    p = wl.gauss_pulse(64, 48, amp=1e-6, sigma_cells=4.0)
    q, st = run(oracle, p, 15)
    s = np.abs(q[0]).max()
    assert np.abs(q[0] - q[0][:, ::-1]).max() < 1e-9 * s
    assert np.abs(q[1] + q[1][:, ::-1]).max() < 1e-9 * s
    assert np.abs(q[2] - q[2][:, ::-1]).max() < 1e-9 * s
    assert np.abs(q[0] - q[0][::-1, :]).max() < 1e-9 * s


def test_linear_wave_travels_at_c0(oracle):
    # y-uniform right-going pulse (rho u = rho'), tiny amplitude: the centroid
    # of rho' moves at c0 = 1
    nx, ny = 400, 8
    x = np.arange(nx)
    r = 1e-6 * np.exp(-((x - 100.0) / 8.0) ** 2)
    q0 = np.zeros((3, ny, nx))
    q0[0] = r[None, :]
    q0[1] = r[None, :]
    p = wl.gauss_pulse(nx, ny)
    p.q0 = q0
    q, st = run(oracle, p, 100)
    row = q[0][ny // 2]
    cx = (row * x).sum() / row.sum()
    moved = (cx - 100.0) / (1.0 / p.dx)     # in length units
    assert abs(moved - st.T) < 0.02 * st.T


def test_taylor_error_falls_with_resolution(oracle):
    # the reference's own known-answer app (apps/taylor-angle-xy): the L1
    # error against the closed form at the same time must fall at close to
    # second order when the resolution doubles (eta = 10 -> 20 cells/unit)
    errs = []
    for eta in (2, 4):
        p = wl.taylor(eta)
        q, st = run(oracle, p, 20 * eta)
        t = p.taylor
        I = np.arange(p.nx)
        exact = t["rhoa"] * np.tanh(t["c"] * (I * p.dx - st.T))
        errs.append(np.abs(q[0][p.ny // 2] - exact).mean() / abs(t["rhoa"]))
    assert errs[1] < 0.4 * errs[0], errs


def test_literal_and_rho1_forms_agree_in_fp64(oracle):
    p = wl.edge_drive(96, 64, halfwidth_cells=10.0)
    a, _ = run(oracle, p, 60, form=oracle.FORM_LITERAL)
    b, _ = run(oracle, p, 60, form=oracle.FORM_RHO1)
    for k in range(3):
        assert np.abs(a[k] - b[k]).max() <= 1e-10 * np.abs(a[k]).max()


def test_rejection_is_a_clean_retry(oracle):
    # a huge first dt is rejected and the state is left untouched; the run
    # then matches one started from the dt the retry used
    p = wl.gauss_pulse(48, 48, sigma_cells=4.0, dt0=1.0)
    q, st = run(oracle, p, 5)
    assert st.n_rejected >= 1 and st.n_accepted == 5
    p2 = wl.gauss_pulse(48, 48, sigma_cells=4.0, dt0=1.0)
    op = to_oracle(p2)
    _, st1, _ = oracle.run(op, p2.q0.astype(np.float64), 0)   # no attempts
    assert st1.n_attempts == 0


# ---------------------------------------------------- moving window, frames --

def test_moving_window_is_a_translation(oracle):
    # moveDomain (common/main.cu:109-176) only relabels cells: away from the
    # edges the moving run equals the fixed run read at (i + MDX, j + MDY),
    # bit for bit, and MDX = int((T - MDT) * MDVX / dx) after every step
    nst, r = 12, 48                       # r: reach of the edges in 12 steps
    md = (1.0, 0.5, 0.0)
    fixed = wl.gauss_pulse(200, 160, sigma_cells=8.0)
    moving = wl.gauss_pulse(200, 160, sigma_cells=8.0, moving=md)
    dts = [0.8 * fixed.dx] * nst           # prescribed: 0.8 and 0.4 cells per step
    outs = []
    for p in (fixed, moving):
        q, st, rc = oracle.run(to_oracle(p, dt_seq=dts), p.q0.astype(np.float64), nst)
        assert rc == 0 and st.n_accepted == nst
        outs.append((q, st))
    (qf, stf), (qm, stm) = outs
    assert (stf.mdx, stf.mdy) == (0, 0)
    mx, my = stm.mdx, stm.mdy
    assert mx == int((stm.T - md[2]) * md[0] / fixed.dx) == 9
    assert my == int((stm.T - md[2]) * md[1] / fixed.dy) == 4
    sub_m = qm[:, r:160 - r - my, r:200 - r - mx]
    sub_f = qf[:, r + my:160 - r, r + mx:200 - r]
    assert np.abs(sub_f[0]).max() > 1e-3          # the pulse is inside the window
    assert np.array_equal(sub_m, sub_f)
    # a wrong direction is not a translation
    assert not np.array_equal(qm[:, r:160 - r - my, r:200 - r - mx],
                              qf[:, r:160 - r - my, r:200 - r - mx])


def test_moving_window_keeps_the_taylor_shock_in_view(oracle):
    # apps/taylor-angle-xy with its moving window (parameters.h:92-95): the
    # window follows the shock at unit speed along theta, so long after the
    # window starts the state still matches the closed form taken at
    # (i + MDX, j + MDY) as closely as the fixed run does before it starts
    theta = math.pi / 6
    p = wl.taylor(2, theta=theta, moving=True)
    t = p.taylor
    I = np.arange(p.nx)[None, :]
    J = np.arange(p.ny)[:, None]

    def err(q, st):
        exact = t["rhoa"] * np.tanh(t["c"] * ((I + st.mdx) * p.dx * math.cos(theta)
                                              + ((J + st.mdy - t["jc"]) * p.dy + t["yoff"])
                                              * math.sin(theta) - st.T))
        return np.abs(q[0][8:-8, 8:-8] - exact[8:-8, 8:-8]).mean() / abs(t["rhoa"])

    md_t = p.moving[2]
    op = to_oracle(p, T_stop=md_t)           # up to the start of the window
    q1, st1, rc = oracle.run(op, p.q0.astype(np.float64), 10 ** 6)
    assert rc == 0 and st1.T >= md_t and (st1.mdx, st1.mdy) == (0, 0)
    e_before = err(q1, st1)
    op = to_oracle(p, T_stop=md_t + 5.0)     # five more time units
    q2, st2, rc = oracle.run(op, q1, 10 ** 6, state=st1)
    assert rc == 0 and st2.T >= md_t + 5.0
    assert st2.mdx == int((st2.T - md_t) * math.cos(theta) / p.dx) > 20
    assert st2.mdy == int((st2.T - md_t) * math.sin(theta) / p.dy) > 10
    assert err(q2, st2) < 2.0 * e_before + 1e-3, (err(q2, st2), e_before)


def test_run_stops_at_T_stop(oracle):
    # the frame loop while (T < Tprint) (common/main.cu:310): the run stops
    # at the first accepted step that reaches T_stop, as a step count does
    p = wl.gauss_pulse(64, 48, sigma_cells=5.0)
    q, st, rc = oracle.run(to_oracle(p, T_stop=0.3), p.q0.astype(np.float64), 10 ** 6)
    assert rc == 0 and st.T >= 0.3
    n = st.n_accepted
    q2, st2, rc = oracle.run(to_oracle(p), p.q0.astype(np.float64), n - 1)
    assert st2.T < 0.3
    q3, st3, rc = oracle.run(to_oracle(p), q2, 1, state=st2)
    assert st3.T == st.T and np.array_equal(q3, q)


# ------------------------------------------- taylor-angle-xy data collection --
# apps/taylor-angle-xy/data_collect.h:7-98: rows 0/1 = T and cfl of the step,
# row 2 = rho' interpolated in y on the line through the front's centre along
# the propagation direction, row 3 = the closed form at the nearer row.

def _taylor_collect_problem(oracle, theta, fp64=True):
    p = wl.taylor(1, theta=theta)
    op = to_oracle(p)
    op.collect = dict(wl.taylor_collect(p, fp64=fp64), f32=not fp64)
    op.measure = np.zeros((p.ny, p.nx))
    return p, op


def test_taylor_cut_on_the_exact_profile_at_zero_angle(oracle):
    # theta = 0: ja = jcenter exactly, so the cut is row jcenter of the state,
    # columns icenter - L .. icenter + L, and the state equal to the closed
    # form makes both rows the same numbers
    p, op = _taylor_collect_problem(oracle, 0.0)
    t, T = p.taylor, 3.3
    X = np.arange(p.nx) * p.dx
    prof = t["rhoa"] * np.tanh(t["c"] * (X - T))
    q1 = np.tile(prof, (p.ny, 1))
    assert oracle.collect(op, q1, T, n_accepted=5, cfl=0.8) == 0
    m = op.measure
    L = op.collect["cut_half"]
    ic = int(T / p.dx)                                  # 16
    assert L == 25 and ic == 16
    i = np.arange(max(ic - L, 0), ic + L + 1)
    iw = i - (ic - L)
    np.testing.assert_array_equal(m[2, iw], q1[p.ny // 2, i])
    np.testing.assert_allclose(m[3, iw], prof[i], rtol=1e-14, atol=1e-22)
    assert np.all(m[2, :iw[0]] == 0)                    # i < 0 has no thread
    assert m[0, 4] == T and m[1, 4] == 0.8              # entry n - 1 - n0frame


def test_taylor_cut_interpolates_on_the_line(oracle):
    # a state linear in i and j is interpolated exactly: row 2 returns
    # 1000 i + ja, which fixes the weights, the line and the offsets
    theta = math.pi / 8
    p, op = _taylor_collect_problem(oracle, theta)
    t, T, mdx, mdy = p.taylor, 8.0, 3, 1
    I, J = np.meshgrid(np.arange(p.nx), np.arange(p.ny))
    q1 = 1000.0 * I + J
    assert oracle.collect(op, q1, T, n_accepted=1, mdx=mdx, mdy=mdy) == 0
    m = op.measure
    L = op.collect["cut_half"]
    ic = int(T * math.cos(theta) / p.dx) - mdx
    jc = p.ny // 2 + int((T - t["mdt"]) * math.sin(theta) / p.dy) - mdy
    i = np.arange(ic - L, ic + L + 1)
    ja = jc + (i - ic) * math.tan(theta)
    np.testing.assert_allclose(m[2, i - (ic - L)], 1000.0 * i + ja, rtol=0, atol=1e-4)
    jn = np.floor(ja + 0.5)                             # j - ja in (-1/2, 1/2]
    an = t["rhoa"] * np.tanh(t["c"] * ((i + mdx) * p.dx * math.cos(theta) +
                                       ((jn + mdy - p.ny // 2) * p.dy + t["mdt"] *
                                        math.sin(theta)) * math.sin(theta) - T))
    np.testing.assert_allclose(m[3, i - (ic - L)], an, rtol=1e-12, atol=0)


def test_taylor_cut_float_build_matches_double(oracle):
    # PRECISION=1 arithmetic (float T, constants and closed form) picks the
    # same cells and agrees to float rounding
    theta = math.pi / 16
    T = 7.25
    p, o64 = _taylor_collect_problem(oracle, theta, fp64=True)
    _, o32 = _taylor_collect_problem(oracle, theta, fp64=False)
    t = p.taylor
    X, Y = np.meshgrid(np.arange(p.nx) * p.dx, (np.arange(p.ny) - p.ny // 2) * p.dy)
    q1 = t["rhoa"] * np.tanh(t["c"] * (X * math.cos(theta) + (Y + t["yoff"]) *
                                       math.sin(theta) - T))
    assert oracle.collect(o64, q1, T, 1) == 0 and oracle.collect(o32, q1, T, 1) == 0
    a, b = o64.measure[2:4], o32.measure[2:4]
    assert np.array_equal(a != 0, b != 0)
    amp = abs(t["rhoa"])
    assert np.abs(a - b).max() < 1e-5 * amp
    # the interpolated profile and the closed form differ by the half-cell
    # offset of the closed form's row and the interpolation error only
    assert np.abs(a[0] - a[1]).max() < 0.2 * amp


def test_taylor_collection_zeroes_before_T0_and_overflows(oracle):
    p, op = _taylor_collect_problem(oracle, 0.0)
    op.measure[:] = 1.0
    q1 = np.zeros((p.ny, p.nx))
    assert oracle.collect(op, q1, -0.5, 3) == 0 and not op.measure.any()
    op.n0frame = 0
    assert oracle.collect(op, q1, 1.0, p.nx + 1) == oracle.ERR_COLLECT


# -------------------------------------------------------- dt rule (update_dt) --
# common/main.cu:199-230: cfl = max over the counted cells of |s| dt/d (both
# sweeps); accept when cfl < CFLMAX; then dt <- dt CFLWHISH / cfl, and on a
# rejection dt <- 0.9 dt when that is smaller.  T grows by the dt of accepted
# steps only (:232-238).  A uniform flow leaves the state unchanged and has
# closed-form speeds (test_uniform_flow_speeds): along x |U| + sqrt(1 - U^2),
# along y (no normal flow) 1.

def _uniform_flow(U, dx, dy, dt0, cfl_wish, nx=32, ny=24):
    p = wl.gauss_pulse(nx, ny, amp=0.0, dt0=dt0)
    p.dx, p.dy, p.cfl_wish = dx, dy, cfl_wish
    q0 = np.zeros((3, ny, nx))
    q0[1] = U                                  # rho = 1 moving at U along x
    p.q0 = q0
    return p


def _uniform_cfl_per_dt(U, dx, dy):
    return max((abs(U) + math.sqrt(1 - U * U)) / dx, 1.0 / dy)


@pytest.mark.parametrize("dy", [0.06, 0.03])       # the x sweep, then the y sweep, decides
def test_dt_rule_accepted_step(oracle, dy):
    U, dx, dt0, cw = 0.3, 0.05, 0.004, 0.9
    p = _uniform_flow(U, dx, dy, dt0, cw)
    q, st, rc = oracle.run(to_oracle(p), p.q0, 1)
    assert rc == 0 and st.n_accepted == 1 and st.n_rejected == 0
    c = _uniform_cfl_per_dt(U, dx, dy)
    assert st.last_cfl == pytest.approx(c * dt0, rel=1e-14)
    assert st.dt == pytest.approx(cw / c, rel=1e-14)          # dt0 * cw / (c dt0)
    assert st.T == pytest.approx(dt0, rel=1e-15)
    # cfl_max plays no part in the new dt of an accepted step
    p.cfl_max = 3.0
    _, st2, _ = oracle.run(to_oracle(p), p.q0, 1)
    assert st2.dt == st.dt


def test_dt_rule_grows_without_the_rejection_cap(oracle):
    # cfl 0.3 is accepted: the next dt is 3 dt0 (the 0.9 cap is for
    # rejections only), and the step after it runs at cfl = cfl_wish
    U, dx, dy, cw = 0.3, 0.05, 0.06, 0.9
    c = _uniform_cfl_per_dt(U, dx, dy)
    dt0 = 0.3 / c
    p = _uniform_flow(U, dx, dy, dt0, cw)
    _, st, rc = oracle.run(to_oracle(p), p.q0, 2)
    assert rc == 0 and st.n_rejected == 0
    assert st.T == pytest.approx(dt0 + 3 * dt0, rel=1e-14)
    assert st.last_cfl == pytest.approx(cw, rel=1e-14)


@pytest.mark.parametrize("cw,factor", [(0.9, 0.9 / 1.05), (0.99, 0.9)])
def test_dt_rule_after_a_rejection(oracle, cw, factor):
    # cfl 1.05 >= CFLMAX = 1: rejected, the state left alone; the retry runs
    # at dt0 cw / 1.05 when that is below 0.9 dt0, else at 0.9 dt0.  Its T
    # shows the dt it ran with.
    U, dx, dy = 0.3, 0.05, 0.06
    c = _uniform_cfl_per_dt(U, dx, dy)
    dt0 = 1.05 / c
    p = _uniform_flow(U, dx, dy, dt0, cw)
    q, st, rc = oracle.run(to_oracle(p), p.q0, 1)
    assert rc == 0 and st.n_rejected == 1 and st.n_accepted == 1 and st.n_attempts == 2
    assert st.T == pytest.approx(factor * dt0, rel=1e-14)
    assert st.dt == pytest.approx(cw / c, rel=1e-14)
    np.testing.assert_array_equal(q, p.q0)


def test_dt_rule_on_a_step_of_a_pulse(oracle):
    # a non-uniform state on a 48x40 grid (a fold subset that is not the even
    # offsets): the cfl of the step is the largest |s| dt/d over the counted
    # cells (subset, i >= 1 for x interfaces, j >= 1 for y ones), with the y
    # speeds on the state after the ghost cells and the x speeds on the y
    # sweep's result, each from the per-interface solve
    p = wl.gauss_pulse(48, 40, sigma_cells=3.0, amp=0.05, dt0=0.02)
    op = to_oracle(p)
    q0 = p.q0.astype(np.float64)
    _, st, rc = oracle.run(op, q0, 1)
    assert rc == 0 and st.n_accepted == 1 and st.n_rejected == 0
    sub = oracle.cfl_subset(p.nx, p.ny)
    assert 0 < sub.mean() < 0.5 and not np.array_equal(sub, _even(p.nx, p.ny))
    qb, _ = oracle.boundary(op, q0, p.T0)
    qy, _, _ = oracle.sweep(op, 1, qb, p.dt0)
    cmax = 0.0
    for j in range(p.ny):
        for i in range(1, p.nx):
            if not sub[j, i]:
                continue
            if j >= 1:   # y interface (i, j - 1/2): normal rho v, tangential rho u
                _, r = oracle.riemann([qb[0, j - 1, i], qb[2, j - 1, i], qb[1, j - 1, i]],
                                      [qb[0, j, i], qb[2, j, i], qb[1, j, i]], p.beta)
                cmax = max(cmax, np.abs(r["s"]).max() * p.dt0 / p.dy)
            _, r = oracle.riemann([qy[0, j, i - 1], qy[1, j, i - 1], qy[2, j, i - 1]],
                                  [qy[0, j, i], qy[1, j, i], qy[2, j, i]], p.beta)
            cmax = max(cmax, np.abs(r["s"]).max() * p.dt0 / p.dx)
    assert st.last_cfl == pytest.approx(cmax, rel=1e-14)
    assert st.dt == pytest.approx(p.dt0 * p.cfl_wish / cmax, rel=1e-14)


def _even(nx, ny):
    return (np.arange(nx * ny) % 2 == 0).reshape(ny, nx).astype(np.uint8)


# ----------------------------------------------- hifu-beam data collection --
# apps/hifu-beam/data_collect.h:7-107, after the accepted step n of a frame
# that started at n0 (entry nf = n - 1 - n0), on row JCUT of the new state:
#   row 0 [nf] = T, row 1 [nf] = cfl;
#   rows 2..5: microphone k hears u1 at the cell i with i + MDX == mic[k]
#     (row 2 + k, entry nf); at nf == 0 the row is cleared first, and a
#     microphone outside the window has its row cleared;
#   u1max: every cell with |(i + MDX)/ETA - T| <= 0.5 keeps max(u1, m) at the
#     flat index i + MDX + 6 NX (written only inside the grid).

def _hifu_collect_problem(oracle, nx=96, ny=12, mic=(10, 30, 50, 200), eta=20.0):
    p = wl.hifu(nx=nx, ny=ny, ispam=2.0)
    op = to_oracle(p)
    op.collect = dict(jcut=2, mic=mic, eta=eta)
    op.measure = np.zeros((ny, nx))
    return p, op


def test_hifu_collection_microphones_and_peak(oracle):
    p, op = _hifu_collect_problem(oracle)
    nx, ny, eta = p.nx, p.ny, 20.0
    I = np.arange(nx)
    u1 = 1e-3 * np.sin(I / 5.0)
    q1 = np.zeros((ny, nx))
    q1[2] = u1                                   # JCUT = 2
    q1[3] = 1.0                                  # other rows are not read
    op.measure[:] = -5.0                         # nf = 0 clears the microphone rows
    op.measure[6:] = 0.0
    T = 2.0                                      # cells 30..50: |i/20 - 2| <= 0.5
    assert oracle.collect(op, q1, T, n_accepted=1, cfl=0.7) == 0
    m = op.measure
    assert m[0, 0] == T and m[1, 0] == 0.7
    for k, mk in enumerate((10, 30, 50)):
        row = np.zeros(nx)
        row[0] = u1[mk]
        np.testing.assert_array_equal(m[2 + k], row)
    assert not m[5].any()                        # microphone 200 is outside the grid
    peak = np.zeros(nx)
    w = np.abs(I / eta - T) <= 0.5
    peak[w] = np.maximum(u1[w], 0.0)
    np.testing.assert_array_equal(m[6], peak)
    assert w.sum() == 21 and not m[7:].any()
    # the next step of the frame, the window moved 4 cells: microphone k
    # hears column mic[k] - 4 at entry 1, the maximum keeps the larger value
    u1b = 1.5e-3 * np.cos(I / 7.0)
    q1[2] = u1b
    T2 = 2.1
    assert oracle.collect(op, q1, T2, n_accepted=2, mdx=4, cfl=0.8) == 0
    assert m[0, 1] == T2 and m[1, 1] == 0.8 and m[0, 0] == T
    for k, mk in enumerate((10, 30, 50)):
        assert m[2 + k, 0] == u1[mk] and m[2 + k, 1] == u1b[mk - 4]
    w2 = np.abs((I + 4) / eta - T2) <= 0.5
    peak2 = peak.copy()
    idx = I[w2] + 4
    peak2[idx] = np.maximum(u1b[w2], peak[idx])
    np.testing.assert_array_equal(m[6], peak2)


def test_hifu_collection_peak_past_the_grid_and_overflow(oracle):
    # u1max index i + MDX + 6 NX is written only while it is inside the grid
    # (NX NY cells); before T > 0 the grid is zeroed; more than NX steps in a
    # frame is simError 4
    p, op = _hifu_collect_problem(oracle, nx=64, ny=8)
    nx = p.nx
    q1 = np.zeros((p.ny, nx))
    q1[2] = 0.25
    mdx = 2 * nx - 10                            # i + mdx + 6 nx >= 8 nx for i >= 10
    T = (5 + mdx) / 20.0                         # cells 0..15 in the peak window
    assert oracle.collect(op, q1, T, n_accepted=1, mdx=mdx) == 0
    flat = op.measure.reshape(-1)
    written = np.nonzero(flat[6 * nx:])[0]
    np.testing.assert_array_equal(written, np.arange(0, 10) + mdx)
    assert np.all(flat[6 * nx + mdx: 6 * nx + mdx + 10] == 0.25)
    op.measure[:] = 3.0
    assert oracle.collect(op, q1, -0.1, n_accepted=2) == 0 and not op.measure.any()
    op.n0frame = 0
    assert oracle.collect(op, q1, 1.0, n_accepted=nx + 1) == oracle.ERR_COLLECT
