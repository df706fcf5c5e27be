"""fivonagi_amd -- MI355X-native FiVoNAGI time step.

Python view of the C ABI in include/fivonagi.h (libfivonagi.so, built by
fivonagi_amd.build).  Loading fails loudly when the library is missing: there
is no CPU fallback for the step.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfivonagi.so")

BC_ZEROGRAD, BC_EDGE_DRIVE, BC_TAYLOR, BC_HIFU = 0, 1, 2, 3
CFL_REFERENCE_FOLD, CFL_ALL = 0, 1
OK, ERR_EIGEN, ERR_COLLECT, ERR_ARG, ERR_DTSEQ, ERR_STALL, ERR_CHECK = 0, 2, 4, 8, 9, 10, 11

EXPORTS = (
    "fivonagi_init", "fivonagi_init_f64", "fivonagi_set_stream", "fivonagi_set_dt_sequence",
    "fivonagi_step", "fivonagi_step_until", "fivonagi_collect_f64",
    "fivonagi_enqueue", "fivonagi_sync", "fivonagi_collect", "fivonagi_info",
    "fivonagi_step_adjoint", "fivonagi_adjoint_state", "fivonagi_autotune", "fivonagi_set_skew",
    "fivonagi_set_collect", "fivonagi_measure", "fivonagi_measure_f64", "fivonagi_measure_rows",
    "fivonagi_measure_rows_f64", "fivonagi_cut_errors", "fivonagi_cut_errors_f64",
    "fivonagi_cfl_subset", "fivonagi_step_noalloc",
    "fivonagi_destroy", "fivonagi_last_error",
)


class Params(C.Structure):
    """fivonagi_params (include/fivonagi.h)."""
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double),
        ("beta", C.c_double), ("delta", C.c_double),
        ("cfl_max", C.c_double), ("cfl_wish", C.c_double),
        ("dt0", C.c_double), ("T0", C.c_double),
        ("bc", C.c_int32), ("cfl_mode", C.c_int32),
        ("omega", C.c_double),
        ("edge_amp", C.POINTER(C.c_double)), ("edge_phase", C.POINTER(C.c_double)),
        ("tay_rhoa", C.c_double), ("tay_c", C.c_double),
        ("tay_theta", C.c_double), ("tay_yoff", C.c_double),
        ("tay_jc", C.c_int32), ("rows_per_wave", C.c_int32),
        ("md_on", C.c_int32),
        ("md_vx", C.c_double), ("md_vy", C.c_double), ("md_t", C.c_double),
        ("hifu_ampl", C.c_double), ("hifu_frec", C.c_double),
        ("hifu_a", C.c_double), ("hifu_f", C.c_double),
    ]


class Info(C.Structure):
    """fivonagi_info_t (include/fivonagi.h)."""
    _fields_ = [
        ("T", C.c_double), ("dt", C.c_double), ("last_cfl", C.c_double),
        ("n_accepted", C.c_int64), ("n_rejected", C.c_int64), ("n_attempts", C.c_int64),
        ("error", C.c_int32), ("cur", C.c_int32),
        ("n_items", C.c_int32), ("rows_per_wave", C.c_int32),
        ("skew", C.c_double), ("rotate", C.c_int32), ("phase", C.c_int32),
        ("md_x", C.c_int32), ("md_y", C.c_int32),
        ("fp64", C.c_int32), ("pad2", C.c_int32),
    ]


class Objective(C.Structure):
    """fivonagi_objective (include/fivonagi.h): the box of J."""
    _fields_ = [("box_x0", C.c_int32), ("box_y0", C.c_int32), ("box_nx", C.c_int32),
                ("box_ny", C.c_int32)]


class CollectOpts(C.Structure):
    """fivonagi_collect_opts (include/fivonagi.h)."""
    _fields_ = [("jcut", C.c_int32), ("mic", C.c_int32 * 4), ("eta", C.c_double),
                ("cut_half", C.c_int32), ("pad", C.c_int32), ("tan_theta", C.c_double)]


class NoallocState(C.Structure):
    """fivonagi_noalloc_state (include/fivonagi.h)."""
    _fields_ = [("T", C.c_double), ("dt", C.c_double),
                ("n_accepted", C.c_int64), ("n_rejected", C.c_int64), ("n_attempts", C.c_int64),
                ("md_x", C.c_int32), ("md_y", C.c_int32),
                ("error", C.c_int32), ("pad", C.c_int32)]


class AdjointOpts(C.Structure):
    """fivonagi_adjoint_opts (include/fivonagi.h)."""
    _fields_ = [("checkpoint_every", C.c_int32), ("plain_forward", C.c_int32),
                ("check_recompute", C.c_int32), ("rows_per_wave", C.c_int32)]


_lib = None


def load():
    """Load libfivonagi.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m fivonagi_amd.build` "
            "(the step has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER(Params)
    vp = C.c_void_p
    lib.fivonagi_init.argtypes = [C.POINTER(vp), P, C.c_int, C.POINTER(C.POINTER(C.c_float))]
    lib.fivonagi_init_f64.argtypes = [C.POINTER(vp), P, C.c_int,
                                      C.POINTER(C.POINTER(C.c_double))]
    lib.fivonagi_set_stream.argtypes = [vp, vp]
    lib.fivonagi_set_dt_sequence.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.c_int64]
    lib.fivonagi_step.argtypes = [vp, C.c_int64]
    lib.fivonagi_step_until.argtypes = [vp, C.c_double]
    lib.fivonagi_enqueue.argtypes = [vp, C.c_int64]
    lib.fivonagi_sync.argtypes = [vp]
    fp = C.POINTER(C.c_float)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int64)
    lib.fivonagi_collect.argtypes = [vp, C.c_int, fp, fp, fp, dp, dp, ip, ip]
    lib.fivonagi_collect_f64.argtypes = [vp, C.c_int, dp, dp, dp, dp, dp, ip, ip]
    lib.fivonagi_info.argtypes = [vp, C.c_int, C.POINTER(Info)]
    lib.fivonagi_step_adjoint.argtypes = [vp, C.c_int64, C.POINTER(Objective),
                                          C.POINTER(AdjointOpts), dp, dp, dp]
    lib.fivonagi_adjoint_state.argtypes = [vp, C.c_int, fp, fp, fp]
    lib.fivonagi_autotune.argtypes = [vp, C.c_int, C.POINTER(C.c_int32)]
    lib.fivonagi_set_skew.argtypes = [vp, C.c_double, C.c_int]
    lib.fivonagi_set_collect.argtypes = [vp, C.c_int, C.POINTER(CollectOpts)]
    lib.fivonagi_measure.argtypes = [vp, C.c_int, C.POINTER(C.c_float)]
    lib.fivonagi_measure_f64.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
    lib.fivonagi_measure_rows.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_float)]
    lib.fivonagi_measure_rows_f64.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double)]
    lib.fivonagi_cut_errors.argtypes = [fp, fp, C.c_int, fp]
    lib.fivonagi_cfl_subset.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_uint8)]
    lib.fivonagi_cut_errors_f64.argtypes = [dp, dp, C.c_int, fp]
    lib.fivonagi_step_noalloc.argtypes = [P, C.POINTER(vp), C.c_int64, C.POINTER(NoallocState),
                                          vp]
    lib.fivonagi_destroy.argtypes = [vp]
    lib.fivonagi_destroy.restype = None
    lib.fivonagi_last_error.argtypes = []
    lib.fivonagi_last_error.restype = C.c_char_p
    for name in ("fivonagi_init", "fivonagi_init_f64", "fivonagi_collect_f64",
                 "fivonagi_set_stream", "fivonagi_set_dt_sequence",
                 "fivonagi_step", "fivonagi_step_until", "fivonagi_enqueue", "fivonagi_sync",
                 "fivonagi_collect",
                 "fivonagi_info", "fivonagi_step_adjoint", "fivonagi_adjoint_state",
                 "fivonagi_autotune", "fivonagi_set_skew", "fivonagi_set_collect",
                 "fivonagi_measure", "fivonagi_measure_f64", "fivonagi_measure_rows",
                 "fivonagi_measure_rows_f64", "fivonagi_cut_errors", "fivonagi_cut_errors_f64",
                 "fivonagi_cfl_subset", "fivonagi_step_noalloc"):
        getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


class FivonagiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fivonagi error {code}: {msg}")
        self.code = code


def _check(rc):
    if rc != 0:
        raise FivonagiError(rc, load().fivonagi_last_error().decode(errors="replace"))
    return rc


class Problem:
    """One grid instance: runtime form of an app's parameters.h + init.h.

    q0 has shape (3, ny, nx): rho', rho*u, rho*v (float32 on upload)."""

    def __init__(self, nx, ny, dx, dy, beta, delta, q0, cfl_max=1.0, cfl_wish=0.9,
                 dt0=0.005, T0=0.0, bc=BC_ZEROGRAD, cfl_mode=CFL_REFERENCE_FOLD,
                 omega=0.0, edge_amp=None, edge_phase=None, taylor=None, rows_per_wave=0,
                 moving=None, hifu=None):
        self.nx, self.ny = int(nx), int(ny)
        self.dx, self.dy, self.beta, self.delta = float(dx), float(dy), float(beta), float(delta)
        self.cfl_max, self.cfl_wish, self.dt0, self.T0 = map(float, (cfl_max, cfl_wish, dt0, T0))
        self.bc, self.cfl_mode, self.omega = int(bc), int(cfl_mode), float(omega)
        self.edge_amp = None if edge_amp is None else np.ascontiguousarray(edge_amp, np.float64)
        self.edge_phase = None if edge_phase is None else np.ascontiguousarray(edge_phase, np.float64)
        self.taylor = taylor
        self.rows_per_wave = int(rows_per_wave)
        # moving window: (md_vx, md_vy, md_t) or None (fixed domain)
        self.moving = None if moving is None else tuple(float(v) for v in moving)
        # BC_HIFU: dict(ampl=AMPL, frec=FREC, a=ISPa, f=ISPF)
        self.hifu = hifu
        self.q0 = np.ascontiguousarray(q0, np.float32)
        self.q0_f64 = np.ascontiguousarray(q0, np.float64)   # for FP64 contexts
        assert self.q0.shape == (3, self.ny, self.nx), self.q0.shape

    def cparams(self):
        p = Params()
        p.nx, p.ny, p.dx, p.dy = self.nx, self.ny, self.dx, self.dy
        p.beta, p.delta, p.cfl_max, p.cfl_wish = self.beta, self.delta, self.cfl_max, self.cfl_wish
        p.dt0, p.T0, p.bc, p.cfl_mode, p.omega = self.dt0, self.T0, self.bc, self.cfl_mode, self.omega
        if self.edge_amp is not None:
            p.edge_amp = self.edge_amp.ctypes.data_as(C.POINTER(C.c_double))
        if self.edge_phase is not None:
            p.edge_phase = self.edge_phase.ctypes.data_as(C.POINTER(C.c_double))
        if self.taylor:
            t = self.taylor
            p.tay_rhoa, p.tay_c, p.tay_theta = t["rhoa"], t["c"], t["theta"]
            p.tay_yoff, p.tay_jc = t["yoff"], int(t["jc"])
        p.rows_per_wave = self.rows_per_wave
        if self.moving is not None:
            p.md_on = 1
            p.md_vx, p.md_vy, p.md_t = self.moving
        elif self.taylor and "mdt" in self.taylor:
            # the app's MDVX, MDVY, MDT locate the measured cut also when the
            # window stays put
            t = self.taylor
            p.md_vx, p.md_vy, p.md_t = t["mdvx"], t["mdvy"], t["mdt"]
        if self.hifu:
            h = self.hifu
            p.hifu_ampl, p.hifu_frec, p.hifu_a, p.hifu_f = h["ampl"], h["frec"], h["a"], h["f"]
        return p


class Context:
    """A fivonagi_ctx: one or more instances advanced together on one stream."""

    def __init__(self, problems, fp64=False):
        """fp64: state and arithmetic in double (fivonagi_init_f64), from each
        problem's unrounded initial state."""
        if isinstance(problems, Problem):
            problems = [problems]
        self.problems = list(problems)
        self.fp64 = bool(fp64)
        lib = load()
        n = len(self.problems)
        parr = (Params * n)(*[p.cparams() for p in self.problems])
        qs = []
        for p in self.problems:
            src = p.q0_f64 if self.fp64 else p.q0
            qs += [np.ascontiguousarray(src[k]) for k in range(3)]
        self._keep = qs
        ct = C.c_double if self.fp64 else C.c_float
        qptr = (C.POINTER(ct) * (3 * n))(*[q.ctypes.data_as(C.POINTER(ct)) for q in qs])
        h = C.c_void_p()
        init = lib.fivonagi_init_f64 if self.fp64 else lib.fivonagi_init
        _check(init(C.byref(h), parr, n, qptr))
        self._h = h
        self._dtseq = {}

    def close(self):
        if getattr(self, "_h", None):
            load().fivonagi_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_stream(self, stream_ptr):
        _check(load().fivonagi_set_stream(self._h, C.c_void_p(stream_ptr)))

    def set_dt_sequence(self, inst, dts):
        if dts is None:
            _check(load().fivonagi_set_dt_sequence(self._h, inst, None, 0))
            return
        a = np.ascontiguousarray(dts, np.float64)
        self._dtseq[inst] = a
        _check(load().fivonagi_set_dt_sequence(
            self._h, inst, a.ctypes.data_as(C.POINTER(C.c_double)), len(a)))

    def step(self, nsteps):
        return _check(load().fivonagi_step(self._h, int(nsteps)))

    def step_until(self, T_stop):
        """Attempts until every instance has T >= T_stop (the frame loop)."""
        return _check(load().fivonagi_step_until(self._h, C.c_double(T_stop)))

    def step_rc(self, nsteps):
        """fivonagi_step returning the raw code instead of raising."""
        return load().fivonagi_step(self._h, int(nsteps))

    def enqueue(self, nattempts):
        return _check(load().fivonagi_enqueue(self._h, int(nattempts)))

    def sync(self):
        return load().fivonagi_sync(self._h)

    def collect(self, inst=0):
        """(q (3, ny, nx), state): float64 for FP64 contexts, else float32."""
        p = self.problems[inst]
        dtype, ct = (np.float64, C.c_double) if self.fp64 else (np.float32, C.c_float)
        fn = load().fivonagi_collect_f64 if self.fp64 else load().fivonagi_collect
        q = np.zeros((3, p.ny, p.nx), dtype)
        T, dt = C.c_double(), C.c_double()
        na, nr = C.c_int64(), C.c_int64()
        fpp = [np.ascontiguousarray(q[k]) for k in range(3)]
        _check(fn(self._h, inst, *[a.ctypes.data_as(C.POINTER(ct)) for a in fpp],
                  C.byref(T), C.byref(dt), C.byref(na), C.byref(nr)))
        return np.stack(fpp), dict(T=T.value, dt=dt.value, n_accepted=na.value,
                                   n_rejected=nr.value)

    def info(self, inst=0):
        i = Info()
        _check(load().fivonagi_info(self._h, inst, C.byref(i)))
        return {k: getattr(i, k) for k, _ in Info._fields_}

    def step_adjoint(self, nsteps, boxes, checkpoint_every=0, plain_forward=False,
                     check_recompute=False, rows_per_wave=0):
        """Forward nsteps (recording each step's control block), then the
        adjoint.  boxes: one (x0, y0, nx, ny) per instance.  check_recompute:
        the reverse pass also recomputes every state and raises
        (ERR_CHECK) if one differs in any bit from the stored one.  Returns
        (J, grad_amp, grad_phase) with one entry per instance (arrays of ny)."""
        n = len(self.problems)
        if isinstance(boxes, tuple):
            boxes = [boxes] * n
        objs = (Objective * n)(*[Objective(*b) for b in boxes])
        opts = AdjointOpts(int(checkpoint_every), int(bool(plain_forward)),
                           int(bool(check_recompute)), int(rows_per_wave))
        J = np.zeros(n)
        tot = sum(p.ny for p in self.problems)
        ga = np.zeros(tot)
        gp = np.zeros(tot)
        dp_ = C.POINTER(C.c_double)
        _check(load().fivonagi_step_adjoint(self._h, int(nsteps), objs, C.byref(opts),
                                            J.ctypes.data_as(dp_), ga.ctypes.data_as(dp_),
                                            gp.ctypes.data_as(dp_)))
        outa, outp, off = [], [], 0
        for p in self.problems:
            outa.append(ga[off:off + p.ny].copy())
            outp.append(gp[off:off + p.ny].copy())
            off += p.ny
        return J, outa, outp

    def adjoint_state(self, inst=0):
        """dJ/d(state at the start of the last step_adjoint call), (3, ny, nx)."""
        p = self.problems[inst]
        ls = [np.zeros((p.ny, p.nx), np.float32) for _ in range(3)]
        _check(load().fivonagi_adjoint_state(
            self._h, inst, *[a.ctypes.data_as(C.POINTER(C.c_float)) for a in ls]))
        return np.stack(ls)

    def set_collect(self, inst, jcut=2, mic=(0, 0, 0, 0), eta=1.0, on=True, cut_half=0,
                    tan_theta=0.0):
        """Data collection (apps/hifu-beam/data_collect.h on row jcut; for a
        Taylor instance apps/taylor-angle-xy/data_collect.h, the cut of
        2 cut_half + 1 points along the propagation direction)."""
        if not on:
            _check(load().fivonagi_set_collect(self._h, inst, None))
            return
        o = CollectOpts()
        o.jcut, o.eta = int(jcut), float(eta)
        for k in range(4):
            o.mic[k] = int(mic[k])
        o.cut_half, o.tan_theta = int(cut_half), float(tan_theta)
        _check(load().fivonagi_set_collect(self._h, inst, C.byref(o)))

    def measure(self, inst=0, rows=None):
        """The measurement grid (ny, nx) of instance inst (float64 in an FP64
        context); rows: only the first rows rows."""
        p = self.problems[inst]
        rows = p.ny if rows is None else int(rows)
        if self.fp64:
            m = np.zeros((rows, p.nx), np.float64)
            _check(load().fivonagi_measure_rows_f64(self._h, inst, rows,
                                                    m.ctypes.data_as(C.POINTER(C.c_double))))
        else:
            m = np.zeros((rows, p.nx), np.float32)
            _check(load().fivonagi_measure_rows(self._h, inst, rows,
                                                m.ctypes.data_as(C.POINTER(C.c_float))))
        return m

    def set_skew(self, skew, rotate=False):
        """Height skew of the work items (< 0: automatic) and the per-strip
        rotation of its levels; results do not depend on either."""
        _check(load().fivonagi_set_skew(self._h, C.c_double(skew), int(bool(rotate))))

    def autotune(self, trial_steps=20):
        """Pick the rows per wave with the fastest step; state is unchanged."""
        h = C.c_int32()
        _check(load().fivonagi_autotune(self._h, int(trial_steps), C.byref(h)))
        return h.value


def cut_errors(numeric, analytic):
    """Relative L1, L2, Linf errors of a Taylor cut as the reference's export
    computes them (float accumulators; fivonagi_cut_errors).  numeric and
    analytic: the cut's entries, float32 (FP32 run) or float64 (FP64 run)."""
    numeric = np.ascontiguousarray(numeric)
    analytic = np.ascontiguousarray(analytic, numeric.dtype)
    e = np.zeros(3, np.float32)
    ef = e.ctypes.data_as(C.POINTER(C.c_float))
    if numeric.dtype == np.float64:
        dp = C.POINTER(C.c_double)
        _check(load().fivonagi_cut_errors_f64(numeric.ctypes.data_as(dp),
                                              analytic.ctypes.data_as(dp), len(numeric), ef))
    else:
        numeric = numeric.astype(np.float32)
        analytic = analytic.astype(np.float32)
        fp = C.POINTER(C.c_float)
        _check(load().fivonagi_cut_errors(numeric.ctypes.data_as(fp),
                                          analytic.ctypes.data_as(fp), len(numeric), ef))
    return e


def cfl_subset(nx, ny):
    """(ny, nx) uint8 mask of the cells the reference's CFL fold carries into
    cfl[0] (fivonagi_cfl_subset)."""
    m = np.zeros((ny, nx), np.uint8)
    _check(load().fivonagi_cfl_subset(int(nx), int(ny), m.ctypes.data_as(C.POINTER(C.c_uint8))))
    return m


def step_noalloc(prob, q_dev, nsteps, state=None, stream=None):
    """fivonagi_step_noalloc: nsteps accepted steps of prob's instance on a
    state the caller keeps on the device.  q_dev: six device addresses (the
    state's three fields, then a work buffer of the same size); state: a
    NoallocState (None: start at prob's T0 and dt0).  Returns the state after
    the call; the fields end in q_dev[0..2]."""
    if state is None:
        state = NoallocState()
        state.T, state.dt = prob.T0, prob.dt0
    p = prob.cparams()
    qs = (C.c_void_p * 6)(*[C.c_void_p(int(a)) for a in q_dev])
    _check(load().fivonagi_step_noalloc(C.byref(p), qs, int(nsteps), C.byref(state),
                                        C.c_void_p(stream)))
    return state
