"""The five workloads of the time step, as seeded synthetic inputs.

Each builder returns fivonagi_amd.Problem objects (parameters + initial state):

  gauss_pulse   400x400 Gaussian density pulse, zero-gradient edges, 1000 steps
  edge_drive    4096x1024 left-edge sinusoidal drive through a Gaussian aperture,
                zero-gradient on the other edges, 8000 steps (the HIFU-beam-like
                single instance the throughput metric is quoted on)
  batched       64 instances of uneven size (nx, ny in 256..2048, step 64),
                Gaussian pulses or windowed plane waves, 2000 steps each
  taylor        the closed-form Taylor shock of apps/taylor-angle-xy

Units follow the reference apps: length in wavelengths of the 1-unit period
(20 cells per unit, apps/hifu-beam/parameters.h:28-30, :96-97), c0 = 1, rho0 = 1.
"""
import math

import numpy as np

from . import BC_EDGE_DRIVE, BC_HIFU, BC_TAYLOR, BC_ZEROGRAD, CFL_REFERENCE_FOLD, Problem

CELLS_PER_UNIT = 20.0


def gauss_pulse(nx=400, ny=400, amp=0.01, sigma_cells=10.0, beta=3.5, delta=1e-4,
                cfl_mode=CFL_REFERENCE_FOLD, dt0=0.005, rows_per_wave=0, moving=None):
    """Gaussian rho' pulse centred at ((nx-1)/2, (ny-1)/2), zero velocity.
    moving: (MDVX, MDVY, MDT) of a moving window, or None."""
    x = np.arange(nx) - (nx - 1) / 2.0
    y = np.arange(ny) - (ny - 1) / 2.0
    X, Y = np.meshgrid(x, y)
    q = np.zeros((3, ny, nx))
    q[0] = amp * np.exp(-(X ** 2 + Y ** 2) / (2.0 * sigma_cells ** 2))
    d = 1.0 / CELLS_PER_UNIT
    return Problem(nx, ny, d, d, beta, delta, q, cfl_max=1.0, cfl_wish=0.9, dt0=dt0,
                   T0=0.0, bc=BC_ZEROGRAD, cfl_mode=cfl_mode, rows_per_wave=rows_per_wave,
                   moving=moving)


def edge_profile(ny, amp=0.02, halfwidth_cells=200.0, centre=None):
    """amp exp(-((j - jc)/w)^2) per row, with the C library's exp (as the
    app's C header apps/edge-drive/parameters.h computes it)."""
    jc = (ny - 1) / 2.0 if centre is None else centre
    return np.array([amp * math.exp(-((j - jc) / halfwidth_cells) * ((j - jc) / halfwidth_cells))
                     for j in range(ny)])


def edge_drive(nx=4096, ny=1024, amp=0.02, halfwidth_cells=200.0, cells_per_wavelength=20.0,
               beta=3.5, delta=5e-5, cfl_mode=CFL_REFERENCE_FOLD, dt0=0.005,
               edge_amp=None, edge_phase=None, rows_per_wave=0, moving=None):
    """Quiescent medium driven from the left ghost columns:
    rho' = amp_j sin(omega T + phase_j), omega = 2*pi/(wavelength in time units).
    moving: (MDVX, MDVY, MDT) of a moving window (the drive stops once the
    window has moved), or None."""
    d = 1.0 / CELLS_PER_UNIT
    omega = 2.0 * math.pi * CELLS_PER_UNIT / cells_per_wavelength
    q = np.zeros((3, ny, nx))
    if edge_amp is None:
        edge_amp = edge_profile(ny, amp, halfwidth_cells)
    return Problem(nx, ny, d, d, beta, delta, q, cfl_max=1.0, cfl_wish=0.9, dt0=dt0, T0=0.0,
                   bc=BC_EDGE_DRIVE, cfl_mode=cfl_mode, omega=omega, edge_amp=edge_amp,
                   edge_phase=edge_phase, rows_per_wave=rows_per_wave, moving=moving)


def plane_packet(nx, ny, amp, theta, wavelength_cells=20.0, width_cells=60.0):
    """Right-going (along theta) Gaussian-windowed linear plane wave."""
    x = np.arange(nx, dtype=np.float64)
    y = np.arange(ny, dtype=np.float64)
    X, Y = np.meshgrid(x, y)
    s = X * math.cos(theta) + Y * math.sin(theta)
    sc = 0.5 * (nx * math.cos(theta) + ny * math.sin(theta))
    r = amp * np.sin(2 * math.pi * s / wavelength_cells) * np.exp(-((s - sc) / width_cells) ** 2)
    q = np.zeros((3, ny, nx))
    q[0] = r
    q[1] = r * math.cos(theta)
    q[2] = r * math.sin(theta)
    return q


def batched(n=64, seed=1234, lo=256, hi=2048, step=64, cfl_mode=CFL_REFERENCE_FOLD,
            rows_per_wave=0):
    rng = np.random.default_rng(seed)
    sizes = np.arange(lo, hi + 1, step)
    probs = []
    d = 1.0 / CELLS_PER_UNIT
    for _ in range(n):
        nx, ny = int(rng.choice(sizes)), int(rng.choice(sizes))
        amp = float(rng.uniform(1e-3, 2e-2))
        beta = float(rng.uniform(3.0, 4.0))
        delta = float(10 ** rng.uniform(-5.0, -3.0))
        if rng.uniform() < 0.5:
            sig = float(rng.uniform(5.0, 20.0))
            x = np.arange(nx) - (nx - 1) / 2.0
            y = np.arange(ny) - (ny - 1) / 2.0
            X, Y = np.meshgrid(x, y)
            q = np.zeros((3, ny, nx))
            q[0] = amp * np.exp(-(X ** 2 + Y ** 2) / (2 * sig ** 2))
        else:
            q = plane_packet(nx, ny, amp, float(rng.uniform(0.0, math.pi / 2)))
        probs.append(Problem(nx, ny, d, d, beta, delta, q, cfl_max=1.0, cfl_wish=0.9,
                             dt0=0.005, T0=0.0, bc=BC_ZEROGRAD, cfl_mode=cfl_mode,
                             rows_per_wave=rows_per_wave))
    return probs


TAYLOR_C = 2.1972245773362196   # apps/taylor-angle-xy/parameters.h:105


def taylor(eta_mult=1, theta=0.0, beta=1.0, delta=1e-7, xmax=12.5, cfl_mode=CFL_REFERENCE_FOLD,
           moving=False, cflw=0.9):
    """apps/taylor-angle-xy: NX = NY = 64*ETAC, XMAX = YMAX = 12.5; the exact
    Taylor profile is both the initial state and the Dirichlet ghost data.
    cflw: CFLWHISH (CFLW/100).  The app's window velocity and start,
    MDVX = cos(theta), MDVY = sin(theta), MDT = XMAX/2/cos(theta)
    (apps/taylor-angle-xy/parameters.h:92-94), locate the measured cut;
    moving: the window also moves (moveDomain)."""
    nx = ny = 64 * eta_mult
    dx = dy = xmax / (nx - 1)
    rhoa = -delta / beta
    mdt = xmax * 0.5 / math.cos(theta)
    tide = math.ceil(4 + (xmax * 0.5 - mdt * math.sin(theta)) * math.sin(theta))
    yoff = mdt * math.sin(theta)
    jc = ny // 2
    X = np.arange(nx)[None, :].astype(np.float64)
    Y = np.arange(ny)[:, None].astype(np.float64)
    T0 = -float(tide)
    r = rhoa * np.tanh(TAYLOR_C * (X * dx * math.cos(theta) + ((Y - jc) * dy + yoff)
                                   * math.sin(theta) - T0))
    q = np.stack([r, r * math.cos(theta), r * math.sin(theta)])
    t = dict(rhoa=rhoa, c=TAYLOR_C, theta=theta, yoff=yoff, jc=jc, xmax=xmax,
             mdvx=math.cos(theta), mdvy=math.sin(theta), mdt=mdt)
    md = (math.cos(theta), math.sin(theta), mdt) if moving else None
    return Problem(nx, ny, dx, dy, beta, delta, q, cfl_max=1.0, cfl_wish=cflw, dt0=0.005,
                   T0=T0, bc=BC_TAYLOR, cfl_mode=cfl_mode, taylor=t, moving=md)


def taylor_collect(prob, fp64=False):
    """fivonagi_set_collect options of the taylor-angle-xy app: ETA = NX/XMAX,
    the cut half-length static_cast<int>(5.0*ETA*cos(ISPTHETA)) and
    tan(ISPTHETA), with ETA and the angle's cosine in the build's precision
    (float for PRECISION=1, apps/taylor-angle-xy/data_collect.h:49-55)."""
    t = prob.taylor
    if fp64:
        eta = prob.nx / t["xmax"]
        cos = math.cos(t["theta"])
    else:
        eta = float(np.float32(prob.nx) / np.float32(t["xmax"]))
        cos = float(np.cos(np.float32(t["theta"])))
    return dict(eta=eta, cut_half=int(5.0 * eta * cos), tan_theta=math.tan(t["theta"]))


def hifu(etac=1, ispam=30.0, cflw=0.9, nx=None, ny=None, moving=True,
         cfl_mode=CFL_REFERENCE_FOLD, rows_per_wave=0):
    """apps/hifu-beam as configured by its parameters.h: NX = 512*ETAC,
    NY = 1024*ETAC, XMAX = 25, YMAX = 50 (wavelengths), BETA = 4.8,
    DIFFDELTA = 2.974e-4, AMPL = 4.217e-4, FREC = 1, aperture ISPa = ISPAM
    wavelengths, focal distance ISPF = 50 wavelengths (in points), quiescent
    start at T = -TIDE, moving window MDVX = 1, MDVY = 0, MDT = XMAX/2.
    nx, ny override the grid (same physical size)."""
    nx = 512 * etac if nx is None else nx
    ny = 1024 * etac if ny is None else ny
    xmax, ymax, frec = 25.0, 50.0, 1.0
    dx, dy = xmax / (nx - 1), ymax / (ny - 1)
    lam = 2.0 * math.pi / (2.0 * math.pi * frec)           # ISPlambda = 1
    a = ispam * lam
    f = 50.0 * lam / dx
    tide = math.floor(2 + 4 / frec + (a * a * dx) / (2.0 * f * dy * dy))
    q = np.zeros((3, ny, nx))
    h = dict(ampl=4.217e-4, frec=frec, a=a, f=f)
    md = (1.0, 0.0, xmax * 0.5) if moving else None
    return Problem(nx, ny, dx, dy, 4.8, 2.974e-4, q, cfl_max=1.0, cfl_wish=cflw, dt0=0.005,
                   T0=-float(tide), bc=BC_HIFU, cfl_mode=cfl_mode, rows_per_wave=rows_per_wave,
                   moving=md, hifu=h)


def hifu_collect(prob):
    """The hifu app's data-collection options (apps/hifu-beam/data_collect.h:
    45-50, parameters.h): JCUT = 2, ETA = NX/XMAX, microphones at
    round(ISPF*0.5f), round(ISPF*0.9f), ISPF, round(ISPF*1.1f), with ISPF the
    float the FP32 build holds (50*ISPlambda/(XMAX/(NX-1)) rounded to float;
    int() truncates, C round() rounds halves away from zero)."""
    f32 = np.float32
    ispf = f32(50.0 * 1.0 / float(f32(25.0) / f32(prob.nx - 1)))

    def c_round(v):
        return int(math.floor(float(v) + 0.5))

    mic = [c_round(ispf * f32(0.5)), c_round(ispf * f32(0.9)), int(ispf),
           c_round(ispf * f32(1.1))]
    return dict(jcut=2, mic=mic, eta=prob.nx / 25.0)


def aperture_params(ny, n_rows, rng, amp=0.02, halfwidth_rows=None, noise=0.05):
    """Edge-drive parameters on an aperture of n_rows rows centred on the edge:
    amp_j = amp * (g_j + noise * N(0,1)), phase_j = noise * N(0,1), zero
    outside the aperture; g is a Gaussian of half-width n_rows/4."""
    hw = n_rows / 4.0 if halfwidth_rows is None else halfwidth_rows
    j0 = (ny - n_rows) // 2
    rows = np.arange(j0, j0 + n_rows)
    jc = (ny - 1) / 2.0
    a = np.zeros(ny)
    ph = np.zeros(ny)
    a[rows] = amp * (np.exp(-((rows - jc) / hw) ** 2) + noise * rng.standard_normal(n_rows))
    ph[rows] = noise * rng.standard_normal(n_rows)
    return a, ph, rows


def adjoint_single(seed=4, nx=2048, ny=1024):
    """The single-instance adjoint workload: 2048x1024, 256 aperture rows x
    (amplitude, phase) = 512 parameters, 20 cells per wavelength; J over a 64x64
    box at x = 1500 (centred in y), 4000 steps, checkpoints every 64 steps."""
    rng = np.random.default_rng(seed)
    a, ph, rows = aperture_params(ny, 256, rng)
    prob = edge_drive(nx, ny, edge_amp=a, edge_phase=ph, beta=3.5, delta=5e-5)
    box = (1500, ny // 2 - 32, 64, 64)
    return prob, box, rows


def adjoint_batch(n=32, seed=5, size=1024):
    """The batched forward+adjoint workload: n instances of size x size, each
    with a 256-row edge profile, its own drive wavelength (10-40 cells) and its
    own 64x64 focal box; beta = 3.5, delta = 1e-4."""
    rng = np.random.default_rng(seed)
    probs, boxes = [], []
    for _ in range(n):
        a, ph, rows = aperture_params(size, 256, rng)
        lam = float(rng.uniform(10.0, 40.0))
        probs.append(edge_drive(size, size, edge_amp=a, edge_phase=ph, cells_per_wavelength=lam,
                                beta=3.5, delta=1e-4))
        bx = int(rng.integers(size // 2, size - 64 - 2))
        by = int(rng.integers(size // 4, 3 * size // 4 - 64))
        boxes.append((bx, by, 64, 64))
    return probs, boxes
