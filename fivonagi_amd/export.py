"""On-disk frame exports of the hifu-beam and taylor-angle-xy apps, in the
reference's formats (apps/hifu-beam/data_export.h:14-179,
apps/taylor-angle-xy/data_export.h:14-176), so that their analysis scripts
read our runs unchanged.

hifu-beam, per frame (after the frame's ``Context.step_until``):
  u1-NNNN.dat      the whole rho' grid with its x, y axes (in focal lengths)
  u1-cut-NNNN.dat  rho' on the cut row JCUT
  u1max.dat        the running maximum of rho' along the cut (rewritten)
  microphone_N.dat the four microphones' samples of the frame (appended)
  u1-cfl.dat       every step's cfl and the running deviation from CFLWHISH
taylor-angle-xy, per frame:
  u1-NNNN.dat      the whole rho' grid (axes in units of 1/ETA)
  u1-cut-NNNN.dat  the numeric and analytic rho' along the propagation cut
  u1-error.dat     the cut's relative L1, L2 and Linf errors (appended)
  u1-cfl.dat       as above

Numbers are written as C++ ostreams write them: 6 significant digits (%g),
T with 5 (setprecision(5)); the values are the device's, in the run's
precision (float for FP32 contexts, double for FP64 ones).
"""
import os

import numpy as np


def _g(v, prec=6):
    """operator<< of an ostream with the default float format."""
    return "%.*g" % (prec, float(v))


def _export_cfl(ex, frame, flat, nx, steps):
    """u1-cfl.dat, the same in both apps (apps/hifu-beam/data_export.h:143-175,
    apps/taylor-angle-xy/data_export.h:141-172): for the frame's steps with
    T > 0 but the very first, n, T, frame, cfl and sqrt(sum (cfl - CFLWHISH)^2
    / (frame + 1)).  The difference is taken in the run's precision, squared
    in double and summed into a float (the static cfl_std_tmp); the mean is a
    float division and its square root a double."""
    R = flat.dtype.type
    with open(ex._file("u1-cfl.dat"), "w" if frame == 0 else "a") as f:
        if frame == 0:
            f.write("# n T frame cfl std\n")
            ex.cfl_std_tmp = np.float32(0.0)
        for k in range(steps):
            if flat[k] > 0:
                if ex.first == 0:
                    d = float(R(flat[k + nx]) - R(ex.cfl_wish))
                    ex.cfl_std_tmp = np.float32(float(ex.cfl_std_tmp) + d * d)
                    std = float(ex.cfl_std_tmp / np.float32(frame + 1)) ** 0.5
                    f.write("%d %s %d %s %s\n" % (k + ex.previous_n + 1, _g(flat[k]),
                                                  frame, _g(flat[k + nx]), _g(std)))
                ex.first = 0


def write_u1max(path, flat, nx, mdx, ispf, ampl):
    """u1max.dat of the hifu app (apps/hifu-beam/data_export.h:95-116): the
    running maximum of rho' on the cut at the unmoved grid's columns
    0 .. nx + MDX - 1 (flat index i + 6 nx of the measurement grid), x in
    focal lengths (i / ISPF, in float)."""
    ispf = np.float32(ispf)
    with open(path, "w") as f:
        f.write("# AMPL=" + _g(ampl) + "\n")
        f.write("# x u1max(x,y_cut)\n")
        f.write("".join(_g(np.float32(i) / ispf) + " " + _g(flat[i + 6 * nx]) + "\n"
                        for i in range(nx + mdx)))


class HifuExport:
    """dataExport of apps/hifu-beam/data_export.h:14-179 for one instance of a
    Context with data collection on (``Context.set_collect``).

    ispf: focal distance in points (the axes are in focal lengths); ampl:
    AMPL for the headers; cfl_wish: CFLWHISH for the cfl deviation; jcut: the
    cut row.  The static counters of the reference (previous_n, first,
    cfl_std_tmp) are members."""

    def __init__(self, deploy_path, ispf, ampl, cfl_wish, jcut=2):
        self.path = deploy_path
        self.ispf = np.float32(ispf)
        self.ampl = ampl
        self.cfl_wish = cfl_wish
        self.jcut = int(jcut)
        self.previous_n = 0
        self.first = 1
        self.cfl_std_tmp = np.float32(0.0)
        os.makedirs(deploy_path, exist_ok=True)

    def _file(self, name):
        return os.path.join(self.path, name)

    def __call__(self, frame, u1, measure, T, n, mdx, mdy):
        """Write frame `frame` (0, 1, ...): u1 (ny, nx) state, measure the
        measurement grid (its first rows suffice: up to the u1max entries,
        7 + MDX/nx of them; float32, or float64 from an FP64 run), T after the
        frame, n accepted steps so far, mdx/mdy the moving-window offset.
        u1 = None leaves out the two files made from the state (u1-NNNN.dat,
        u1-cut-NNNN.dat)."""
        nx = measure.shape[-1]
        flat = np.ascontiguousarray(measure).reshape(-1)
        Tstr = _g(flat.dtype.type(T), 5)
        fstr = "%04d" % frame
        if u1 is not None:
            self._state_files(fstr, Tstr, u1, mdx, mdy)
        # u1max on the cut (data_export.h:95-116)
        write_u1max(self._file("u1max.dat"), flat, nx, mdx, self.ispf, self.ampl)
        self._collected_files(frame, flat, nx, Tstr, n)

    def _state_files(self, fstr, Tstr, u1, mdx, mdy):
        ny, nx = u1.shape
        # u1 for the whole domain (data_export.h:35-75)
        with open(self._file("u1-" + fstr + ".dat"), "w") as f:
            f.write("# T=" + Tstr + "\n")
            f.write("# Lines are parallel to y-axis, an columns parallel to x-axis.\n")
            f.write("# First line are y values, and first column x values.\n")
            f.write("# Left-top corner is the number of columns\n")
            ys = (np.arange(ny) - mdy - ny // 2).astype(np.float32) / self.ispf
            f.write(str(ny) + " " + "".join(_g(y) + " " for y in ys) + " \n")
            xs = (np.arange(nx) + mdx).astype(np.float32) / self.ispf
            for i in range(nx):
                f.write(_g(xs[i]) + " " + "".join(_g(v) + " " for v in u1[:, i]) + "\n")
        # u1 on the cut (data_export.h:77-93)
        with open(self._file("u1-cut-" + fstr + ".dat"), "w") as f:
            f.write("# T=" + Tstr + "\n")
            f.write("# AMPL=" + _g(self.ampl) + "\n")
            f.write("# x u1(x,y_cut)\n")
            for i in range(nx):
                f.write(_g(xs[i]) + " " + _g(u1[self.jcut, i]) + " \n")

    def _collected_files(self, frame, flat, nx, Tstr, n):
        # microphones (data_export.h:118-141)
        steps = n - self.previous_n
        for m in range(1, 5):
            name = self._file("microphone_%d.dat" % m)
            with open(name, "w" if frame == 0 else "a") as f:
                if frame == 0:
                    f.write("# n T u1\n")
                f.write("## frame = %d\n" % frame)
                f.write("## T_final = " + Tstr + "\n")
                for k in range(steps):
                    f.write("%d %s %s\n" % (k + self.previous_n + 1, _g(flat[k]),
                                            _g(flat[k + (m + 1) * nx])))
        # cfl values (data_export.h:143-175)
        _export_cfl(self, frame, flat, nx, steps)
        self.previous_n = n


def hifu_frames(ctx, prob, deploy_path, frames, dt_print=1.0, inst=0, collect=None):
    """calcFrame's frame loop (common/main.cu:248-756) for a hifu-beam
    instance: `frames` frames of dt_print each from T0, data collection on,
    each frame exported.  collect: fivonagi_set_collect options (default: the
    app's JCUT = 2, microphones at round(ISPF x {0.5, 0.9}), ISPF,
    round(ISPF x 1.1), ETA = NX/XMAX with XMAX = 25: workloads.hifu_collect)."""
    from .workloads import hifu_collect
    h = prob.hifu
    if collect is None:
        collect = hifu_collect(prob)
    ctx.set_collect(inst, **collect)
    ex = HifuExport(deploy_path, h["f"], h["ampl"], prob.cfl_wish, collect["jcut"])
    Tprint = prob.T0 + dt_print
    for frame in range(frames):
        ctx.step_until(Tprint)
        q, st = ctx.collect(inst)
        info = ctx.info(inst)
        ex(frame, q[0], ctx.measure(inst), st["T"], st["n_accepted"], info["md_x"],
           info["md_y"])
        Tprint += dt_print
    return ex


class TaylorExport:
    """dataExport of apps/taylor-angle-xy/data_export.h:14-176 for one instance
    of a Context with Taylor data collection on (``Context.set_collect`` with
    ``workloads.taylor_collect``).

    nx, ny: the grid; eta: ETA = NX/XMAX; theta: ISPTHETA; cut_half: the cut's
    half-length (the cut file and the errors take entries 0..2 cut_half - 1);
    cfl_wish: CFLWHISH; fp64: the run's precision (the reference's
    PRECISION=2), which sets the type of ETA, cos(ISPTHETA) and T in the
    formulas; grids: also write u1-NNNN.dat (the whole grid, the largest
    file by far).  The static counters of the reference (previous_n, first,
    cfl_std_tmp) are members; errors holds every frame's (L1, L2, Linf)."""

    def __init__(self, deploy_path, nx, ny, eta, theta, cut_half, cfl_wish, fp64=False,
                 grids=True):
        self.path = deploy_path
        self.nx, self.ny = int(nx), int(ny)
        self.R = np.float64 if fp64 else np.float32
        self.eta = self.R(eta)
        self.cos = self.R(np.cos(self.R(theta)))
        self.length = 2 * int(cut_half)
        self.cfl_wish = cfl_wish
        self.grids = grids
        self.previous_n = 0
        self.first = 1
        self.cfl_std_tmp = np.float32(0.0)
        self.errors = []
        # the cut's abscissa: -5 + i/(ETA cos(ISPTHETA)), stored in a float
        # (data_export.h:92-93)
        i = np.arange(self.length, dtype=np.float32)
        den = self.R(np.float32(self.eta)) * self.cos
        self.xi = (-5.0 + (i.astype(self.R) / den).astype(np.float64)).astype(np.float32)
        os.makedirs(deploy_path, exist_ok=True)

    def _file(self, name):
        return os.path.join(self.path, name)

    def __call__(self, frame, measure, T, n, mdx=0, mdy=0, u1=None):
        """Write frame `frame`: measure = the first 4 (or more) rows of the
        measurement grid, T after the frame, n accepted steps so far, mdx/mdy
        the window offset; u1 (ny, nx) the state's rho' (needed with grids)."""
        nx, ny = self.nx, self.ny
        flat = np.ascontiguousarray(measure).reshape(-1)
        Tstr = _g(self.R(T), 5)
        fstr = "%04d" % frame
        if self.grids:                                   # data_export.h:35-75
            with open(self._file("u1-" + fstr + ".dat"), "w") as f:
                f.write("# T=" + Tstr + "\n")
                f.write("# Lines are parallel to y-axis, an columns parallel to x-axis.\n")
                f.write("# First line are y values, and first column x values.\n")
                f.write("# Left-top corner is the number of columns\n")
                e32 = np.float32(self.eta)
                ys = (np.arange(ny) + mdy - ny // 2).astype(np.float32) / e32
                f.write(str(ny) + " " + "".join(_g(y) + " " for y in ys) + " \n")
                xs = (np.arange(nx) + mdx).astype(np.float32) / e32
                for i in range(nx):
                    f.write(_g(xs[i]) + " " + "".join(_g(v) + " " for v in u1[:, i]) + " \n")
        num = flat[2 * nx:2 * nx + self.length]
        ana = flat[3 * nx:3 * nx + self.length]
        with open(self._file("u1-cut-" + fstr + ".dat"), "w") as f:   # :77-103
            f.write("# T=" + Tstr + "\n")
            f.write("# i xi numeric analytic \n")
            f.write("".join("%d %s %s %s  \n" % (i, _g(self.xi[i]), _g(num[i]), _g(ana[i]))
                            for i in range(self.length)))
        from . import cut_errors
        err = cut_errors(num, ana)                       # :96-122
        self.errors.append(tuple(float(v) for v in err))
        with open(self._file("u1-error.dat"), "w" if frame == 0 else "a") as f:   # :124-139
            if frame == 0:
                f.write("# frame T error_L1 error_L2 error_Li\n")
            f.write("%d %s %s %s %s\n" % (frame, Tstr, _g(err[0]), _g(err[1]), _g(err[2])))
        _export_cfl(self, frame, flat, nx, n - self.previous_n)   # :141-172
        self.previous_n = n


def frame_loop(ctx, final_time, dt_print, on_frame):
    """calcFrame's frame loop (common/main.cu:248-756, :263, :652, :741) over
    every instance of a context: frames end at T0 + dt_print, T0 + 2 dt_print,
    ... (fivonagi_step_until, which starts a frame for the collection), each
    instance after its own T0 (with different T0 the instances share the
    frame ends where they overlap, an instance joining once its first one
    comes), until its T > final_time.  on_frame(k, frame, info) runs after
    instance k's frame `frame` (info = Context.info(k)).  Returns the seconds
    spent stepping."""
    import time
    if not final_time > 0:
        # finalTime = 0 runs forever in the reference (main.cu:741)
        raise ValueError("final_time must be > 0")
    probs = ctx.problems
    n = len(probs)
    done = [False] * n
    frames = [0] * n
    Tprint = min(p.T0 for p in probs) + dt_print
    step_s = 0.0
    while not all(done):
        t0 = time.perf_counter()
        ctx.step_until(Tprint)
        step_s += time.perf_counter() - t0
        for k in range(n):
            if done[k] or Tprint < probs[k].T0 + dt_print * (1 - 1e-9):
                continue
            info = ctx.info(k)
            on_frame(k, frames[k], info)
            frames[k] += 1
            if info["T"] > final_time:
                done[k] = True
        Tprint += dt_print
    return step_s


def taylor_frames(ctx, deploy_paths, final_time=100.0, dt_print=1.0, grids=False,
                  collect=None, stats=None):
    """calcFrame's frame loop (common/main.cu:248-756) for every instance of a
    Context of taylor-angle-xy problems: frames of dt_print from T0 (all
    instances share the frame times; each keeps its own dt), data collection
    on, each frame exported into that instance's deploy path, until T >
    final_time (main.cu:741).  collect: per instance set_collect options
    (default: workloads.taylor_collect in the context's precision).
    stats: a dict that receives the seconds spent stepping ("step_s", the
    step_until calls) and in all ("total_s").  Returns the TaylorExport
    objects."""
    import time
    from .workloads import taylor_collect
    t_start = time.perf_counter()
    probs = ctx.problems
    exps = []
    for k, p in enumerate(probs):
        opts = taylor_collect(p, ctx.fp64) if collect is None else collect[k]
        ctx.set_collect(k, **opts)
        exps.append(TaylorExport(deploy_paths[k], p.nx, p.ny, opts["eta"], p.taylor["theta"],
                                 opts["cut_half"], p.cfl_wish, fp64=ctx.fp64, grids=grids))

    def on_frame(k, frame, info):
        u1 = ctx.collect(k)[0][0] if grids else None
        exps[k](frame, ctx.measure(k, rows=4), info["T"], info["n_accepted"],
                info["md_x"], info["md_y"], u1=u1)

    step_s = frame_loop(ctx, final_time, dt_print, on_frame)
    if stats is not None:
        stats["step_s"] = step_s
        stats["total_s"] = time.perf_counter() - t_start
    return exps
