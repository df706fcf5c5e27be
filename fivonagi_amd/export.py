"""On-disk frame exports of the hifu-beam app, in the reference's formats
(apps/hifu-beam/data_export.h:14-179), so that its analysis scripts read our
runs unchanged.

Per frame (after the frame's ``Context.step_until``):
  u1-NNNN.dat      the whole rho' grid with its x, y axes (in focal lengths)
  u1-cut-NNNN.dat  rho' on the cut row JCUT
  u1max.dat        the running maximum of rho' along the cut (rewritten)
  microphone_N.dat the four microphones' samples of the frame (appended)
  u1-cfl.dat       every step's cfl and the running deviation from CFLWHISH

Numbers are written as C++ ostreams write them: 6 significant digits (%g),
T with 5 (setprecision(5)); floats are the device's FP32 values.
"""
import os

import numpy as np


def _g(v, prec=6):
    """operator<< of an ostream with the default float format."""
    return "%.*g" % (prec, float(v))


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
        """Write frame `frame` (0, 1, ...): u1 (ny, nx) float32 state, measure
        (ny, nx) float32 measurement grid, T after the frame, n accepted steps
        so far, mdx/mdy the moving-window offset."""
        ny, nx = u1.shape
        flat = np.ascontiguousarray(measure, np.float32).reshape(-1)
        Tstr = _g(T, 5)
        fstr = "%04d" % frame
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
        # u1max on the cut (data_export.h:95-116)
        with open(self._file("u1max.dat"), "w") as f:
            f.write("# AMPL=" + _g(self.ampl) + "\n")
            f.write("# x u1max(x,y_cut)\n")
            for i in range(nx + mdx):
                x = np.float32(i) / self.ispf
                f.write(_g(x) + " " + _g(flat[i + 6 * nx]) + "\n")
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
        with open(self._file("u1-cfl.dat"), "w" if frame == 0 else "a") as f:
            if frame == 0:
                f.write("# n T frame cfl std\n")
                self.cfl_std_tmp = np.float32(0.0)
            for k in range(steps):
                if flat[k] > 0:
                    if self.first == 0:
                        # float cfl - float CFLWHISH, squared in double, summed
                        # into a float; the mean is a float division
                        d = float(np.float32(flat[k + nx]) - np.float32(self.cfl_wish))
                        self.cfl_std_tmp = np.float32(float(self.cfl_std_tmp) + d * d)
                        std = float(self.cfl_std_tmp / np.float32(frame + 1)) ** 0.5
                        f.write("%d %s %d %s %s\n" % (k + self.previous_n + 1, _g(flat[k]),
                                                      frame, _g(flat[k + nx]), _g(std)))
                    self.first = 0
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
