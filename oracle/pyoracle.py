"""ctypes binding of the CPU oracle (oracle/fvoracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline,
never as a stand-in for the HIP path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

BC_ZEROGRAD, BC_EDGE_DRIVE, BC_TAYLOR, BC_HIFU = 0, 1, 2, 3
OK, ERR_DT, ERR_EIGEN, ERR_COLLECT, ERR_ARG = 0, 1, 2, 4, 8
CFL_REFERENCE_FOLD, CFL_ALL = 0, 1
FORM_LITERAL, FORM_RHO1 = 0, 1


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double),
        ("beta", C.c_double), ("delta", C.c_double),
        ("cfl_max", C.c_double), ("cfl_wish", C.c_double),
        ("dt0", C.c_double), ("T0", C.c_double),
        ("bc", C.c_int32), ("cfl_mode", C.c_int32), ("form", C.c_int32),
        ("nthreads", C.c_int32),
        ("omega", C.c_double),
        ("edge_amp", C.POINTER(C.c_double)), ("edge_phase", C.POINTER(C.c_double)),
        ("tay_rhoa", C.c_double), ("tay_c", C.c_double),
        ("tay_theta", C.c_double), ("tay_yoff", C.c_double),
        ("tay_jc", C.c_int32), ("pad0", C.c_int32),
        ("dt_seq", C.POINTER(C.c_double)), ("dt_seq_len", C.c_int64),
        ("md_on", C.c_int32), ("stop_at_T", C.c_int32),
        ("md_vx", C.c_double), ("md_vy", C.c_double), ("md_t", C.c_double),
        ("T_stop", C.c_double),
        ("hifu_ampl", C.c_double), ("hifu_frec", C.c_double),
        ("hifu_a", C.c_double), ("hifu_f", C.c_double),
        ("measure", C.POINTER(C.c_double)),
        ("jcut", C.c_int32), ("mic", C.c_int32 * 4), ("pad1", C.c_int32),
        ("eta", C.c_double), ("n0frame", C.c_int64),
        ("cut_half", C.c_int32), ("collect_f32", C.c_int32), ("tan_theta", C.c_double),
    ]


class State(C.Structure):
    _fields_ = [
        ("T", C.c_double), ("dt", C.c_double), ("last_cfl", C.c_double),
        ("n_accepted", C.c_int64), ("n_rejected", C.c_int64),
        ("n_attempts", C.c_int64), ("error", C.c_int32), ("pad", C.c_int32),
        ("mdx", C.c_int32), ("mdy", C.c_int32),
    ]


class Box(C.Structure):
    _fields_ = [("x0", C.c_int32), ("y0", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32)]


def _load(name):
    path = name if os.path.isabs(name) else os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} missing: build it with `make -C oracle` (or __graft_entry__.build())")
    lib = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    fp = C.POINTER(C.c_float)
    P = C.POINTER(Params)
    S = C.POINTER(State)
    lib.orc_state_init.argtypes = [P, S]
    lib.orc_run_f64.argtypes = [P, dp, dp, dp, C.c_int64, S]
    lib.orc_run_f64.restype = C.c_int
    lib.orc_run_f32.argtypes = [P, fp, fp, fp, C.c_int64, S]
    lib.orc_run_f32.restype = C.c_int
    lib.orc_riemann_f64.argtypes = [dp, dp, C.c_double, C.c_int, dp]
    lib.orc_riemann_f64.restype = C.c_int
    lib.orc_philim_f64.argtypes = [C.c_double] * 4
    lib.orc_philim_f64.restype = C.c_double
    lib.orc_cfl_subset.argtypes = [C.c_int32, C.c_int32, C.POINTER(C.c_uint8)]
    lib.orc_boundary_f64.argtypes = [P, dp, dp, dp, dp, dp, dp, C.c_double]
    lib.orc_boundary_f64.restype = C.c_int
    lib.orc_source_f64.argtypes = [P, dp, dp, dp, dp, dp, dp, C.c_double]
    lib.orc_source_f64.restype = C.c_int
    lib.orc_sweep_f64.argtypes = [P, C.c_int, dp, dp, dp, C.c_double, dp]
    lib.orc_sweep_f64.restype = C.c_int
    lib.orc_collect_f64.argtypes = [P, dp, C.POINTER(State)]
    lib.orc_collect_f64.restype = C.c_int
    B = C.POINTER(Box)
    lib.orc_line_step_f64.argtypes = [P, dp, dp, dp, C.c_double, C.c_double]
    lib.orc_line_step_f64.restype = C.c_int
    lib.orc_adjoint_f64.argtypes = [P, dp, C.c_int64, B, dp, dp, dp, dp]
    lib.orc_adjoint_f64.restype = C.c_int
    lib.orc_adjoint_path_f64.argtypes = [P, dp, C.c_int64, B, dp, dp, dp, dp]
    lib.orc_adjoint_path_f64.restype = C.c_int
    lib.orc_objective_f64.argtypes = [P, dp, C.c_int64, B, dp]
    lib.orc_objective_f64.restype = C.c_int
    return lib


_LIBS = {}


def lib(fast=False):
    """fast: the -O3 build; a path: a build of the same sources elsewhere
    (bench.py's host-native CPU baseline)."""
    if isinstance(fast, str):
        name = fast
    else:
        name = "liboracle_fast.so" if fast else "liboracle.so"
    if name not in _LIBS:
        _LIBS[name] = _load(name)
    return _LIBS[name]


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class Problem:
    """Parameters of one grid instance, in the reference's vocabulary."""

    def __init__(self, nx, ny, dx, dy, beta, delta, cfl_max=1.0, cfl_wish=0.9,
                 dt0=0.005, T0=0.0, bc=BC_ZEROGRAD, cfl_mode=CFL_REFERENCE_FOLD,
                 form=FORM_LITERAL, omega=0.0, edge_amp=None, edge_phase=None,
                 taylor=None, dt_seq=None, nthreads=0, moving=None, T_stop=None, hifu=None):
        self.nx, self.ny = int(nx), int(ny)
        self.dx, self.dy = float(dx), float(dy)
        self.beta, self.delta = float(beta), float(delta)
        self.cfl_max, self.cfl_wish = float(cfl_max), float(cfl_wish)
        self.dt0, self.T0 = float(dt0), float(T0)
        self.bc, self.cfl_mode, self.form = int(bc), int(cfl_mode), int(form)
        self.omega = float(omega)
        self.edge_amp = None if edge_amp is None else np.ascontiguousarray(edge_amp, np.float64)
        self.edge_phase = None if edge_phase is None else np.ascontiguousarray(edge_phase, np.float64)
        self.taylor = taylor  # dict(rhoa, c, theta, yoff, jc)
        self.dt_seq = None if dt_seq is None else np.ascontiguousarray(dt_seq, np.float64)
        self.nthreads = int(nthreads)
        self.moving = None if moving is None else tuple(float(v) for v in moving)
        self.T_stop = None if T_stop is None else float(T_stop)
        self.hifu = hifu  # dict(ampl, frec, a, f)
        # data collection: dict(jcut, mic, eta) (hifu) or dict(eta, cut_half,
        # tan_theta, f32) (Taylor); measure (ny, nx) float64 is
        # the measurement grid the runs write into, n0frame the frame start
        self.collect = None
        self.measure = None
        self.n0frame = 0

    def cparams(self):
        p = Params()
        p.nx, p.ny, p.dx, p.dy = self.nx, self.ny, self.dx, self.dy
        p.beta, p.delta = self.beta, self.delta
        p.cfl_max, p.cfl_wish, p.dt0, p.T0 = self.cfl_max, self.cfl_wish, self.dt0, self.T0
        p.bc, p.cfl_mode, p.form, p.nthreads = self.bc, self.cfl_mode, self.form, self.nthreads
        p.omega = self.omega
        if self.edge_amp is not None:
            p.edge_amp = _ptr(self.edge_amp, C.c_double)
        if self.edge_phase is not None:
            p.edge_phase = _ptr(self.edge_phase, C.c_double)
        if self.taylor:
            t = self.taylor
            p.tay_rhoa, p.tay_c = t["rhoa"], t["c"]
            p.tay_theta, p.tay_yoff, p.tay_jc = t["theta"], t["yoff"], int(t["jc"])
        if self.hifu:
            h = self.hifu
            p.hifu_ampl, p.hifu_frec, p.hifu_a, p.hifu_f = h["ampl"], h["frec"], h["a"], h["f"]
        if self.dt_seq is not None:
            p.dt_seq = _ptr(self.dt_seq, C.c_double)
            p.dt_seq_len = len(self.dt_seq)
        if self.moving is not None:
            p.md_on = 1
            p.md_vx, p.md_vy, p.md_t = self.moving
        elif self.taylor and "mdt" in self.taylor:
            t = self.taylor
            p.md_vx, p.md_vy, p.md_t = t["mdvx"], t["mdvy"], t["mdt"]
        if self.T_stop is not None:
            p.stop_at_T = 1
            p.T_stop = self.T_stop
        if self.collect:
            if self.measure is None:
                self.measure = np.zeros((self.ny, self.nx))
            c = self.collect
            p.measure = _ptr(self.measure, C.c_double)
            p.jcut, p.eta, p.n0frame = int(c.get("jcut", 0)), float(c["eta"]), int(self.n0frame)
            for k in range(4):
                p.mic[k] = int(c.get("mic", (0, 0, 0, 0))[k])
            p.cut_half = int(c.get("cut_half", 0))
            p.tan_theta = float(c.get("tan_theta", 0.0))
            p.collect_f32 = int(bool(c.get("f32", False)))
        return p


def run(prob, q, nsteps, dtype=np.float64, state=None, fast=False):
    """Advance q (3, ny, nx) by nsteps accepted steps (or until T >= prob.T_stop
    when set).  Returns (q, state, rc)."""
    L = lib(fast)
    p = prob.cparams()
    q = np.ascontiguousarray(q, dtype=dtype).copy()
    assert q.shape == (3, prob.ny, prob.nx)
    if state is None:
        state = State()
        L.orc_state_init(C.byref(p), C.byref(state))
    q1, q2, q3 = (np.ascontiguousarray(q[k]) for k in range(3))
    if dtype == np.float64:
        rc = L.orc_run_f64(C.byref(p), _ptr(q1, C.c_double), _ptr(q2, C.c_double),
                           _ptr(q3, C.c_double), int(nsteps), C.byref(state))
    else:
        rc = L.orc_run_f32(C.byref(p), _ptr(q1, C.c_float), _ptr(q2, C.c_float),
                           _ptr(q3, C.c_float), int(nsteps), C.byref(state))
    return np.stack([q1, q2, q3]), state, rc


def riemann(qL, qR, beta, form=FORM_LITERAL):
    out = np.zeros(18)
    qL = np.ascontiguousarray(qL, np.float64)
    qR = np.ascontiguousarray(qR, np.float64)
    rc = lib().orc_riemann_f64(_ptr(qL, C.c_double), _ptr(qR, C.c_double), float(beta),
                               int(form), _ptr(out, C.c_double))
    return rc, dict(s=out[0:3], W=out[3:12].reshape(3, 3).T.copy(),  # W[m, p]: comp m of wave p
                    amdq=out[12:15], apdq=out[15:18])


def philim(wnorm2, dotl, dotr, s):
    return lib().orc_philim_f64(float(wnorm2), float(dotl), float(dotr), float(s))


def cfl_subset(nx, ny):
    m = np.zeros(nx * ny, np.uint8)
    lib().orc_cfl_subset(nx, ny, _ptr(m, C.c_uint8))
    return m.reshape(ny, nx)


def boundary(prob, q, T):
    p = prob.cparams()
    q = np.ascontiguousarray(q, np.float64)
    o = np.zeros_like(q)
    ins = [np.ascontiguousarray(q[k]) for k in range(3)]
    outs = [np.ascontiguousarray(o[k]) for k in range(3)]
    rc = lib().orc_boundary_f64(C.byref(p), *[_ptr(a, C.c_double) for a in ins],
                                *[_ptr(a, C.c_double) for a in outs], float(T))
    return np.stack(outs), rc


def source(prob, q, dt):
    p = prob.cparams()
    q = np.ascontiguousarray(q, np.float64)
    ins = [np.ascontiguousarray(q[k]) for k in range(3)]
    outs = [np.zeros_like(ins[k]) for k in range(3)]
    rc = lib().orc_source_f64(C.byref(p), *[_ptr(a, C.c_double) for a in ins],
                              *[_ptr(a, C.c_double) for a in outs], float(dt))
    return np.stack(outs), rc


def sweep(prob, axis, q, dt):
    p = prob.cparams()
    qs = [np.ascontiguousarray(q[k], np.float64).copy() for k in range(3)]
    cfl = np.zeros_like(qs[0])
    rc = lib().orc_sweep_f64(C.byref(p), int(axis), *[_ptr(a, C.c_double) for a in qs],
                             float(dt), _ptr(cfl, C.c_double))
    return np.stack(qs), cfl, rc


def line_step(prob, q, T, dt):
    p = prob.cparams()
    qs = [np.ascontiguousarray(q[k], np.float64).copy() for k in range(3)]
    rc = lib().orc_line_step_f64(C.byref(p), *[_ptr(a, C.c_double) for a in qs], float(T),
                                 float(dt))
    return np.stack(qs), rc


def objective(prob, q0, nsteps, box):
    """J = sum_n dt_n sum_box rho'^2 of the oracle forward (prescribed dt)."""
    p = prob.cparams()
    q = np.ascontiguousarray(q0, np.float64)
    J = C.c_double()
    rc = lib().orc_objective_f64(C.byref(p), _ptr(q, C.c_double), int(nsteps),
                                 C.byref(Box(*box)), C.byref(J))
    return J.value, rc


def adjoint(prob, q0, nsteps, box):
    """(J, dJ/d edge_amp, dJ/d edge_phase, dJ/dq0) with frozen (prescribed) dt."""
    p = prob.cparams()
    q = np.ascontiguousarray(q0, np.float64)
    J = C.c_double()
    ga = np.zeros(prob.ny)
    gp = np.zeros(prob.ny)
    gq = np.zeros((3, prob.ny, prob.nx))
    rc = lib().orc_adjoint_f64(C.byref(p), _ptr(q, C.c_double), int(nsteps), C.byref(Box(*box)),
                               C.byref(J), _ptr(ga, C.c_double), _ptr(gp, C.c_double),
                               _ptr(gq, C.c_double))
    return J.value, ga, gp, gq, rc


def adjoint_path(prob, states, box):
    """adjoint() linearised along a given trajectory: states (nsteps + 1, 3,
    ny, nx), q_0 .. q_nsteps of a run with prob's dt sequence."""
    p = prob.cparams()
    st = np.ascontiguousarray(states, np.float64)
    nsteps = st.shape[0] - 1
    J = C.c_double()
    ga = np.zeros(prob.ny)
    gp = np.zeros(prob.ny)
    gq = np.zeros((3, prob.ny, prob.nx))
    rc = lib().orc_adjoint_path_f64(C.byref(p), _ptr(st, C.c_double), int(nsteps),
                                    C.byref(Box(*box)), C.byref(J), _ptr(ga, C.c_double),
                                    _ptr(gp, C.c_double), _ptr(gq, C.c_double))
    return J.value, ga, gp, gq, rc


def collect(prob, q1, T, n_accepted, mdx=0, mdy=0, cfl=0.0):
    """dataCollect (either app) on a state whose rho' is q1 (ny, nx), after
    accepted step n_accepted ending at T, into prob.measure.  Returns rc."""
    p = prob.cparams()
    st = State()
    st.T, st.n_accepted, st.mdx, st.mdy, st.last_cfl = float(T), int(n_accepted), int(mdx), \
        int(mdy), float(cfl)
    q1 = np.ascontiguousarray(q1, np.float64)
    return lib().orc_collect_f64(C.byref(p), _ptr(q1, C.c_double), C.byref(st))
