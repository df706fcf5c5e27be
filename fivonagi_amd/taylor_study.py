"""The taylor-angle-xy convergence study: the sweep of apps/taylor-angle-xy/
run.sh:4-81 (precision x resolution x CFL x angle, each run to finalTime with
the app's exports), the error extraction of plots/error.py and the
convergence-rate tables of plots/error_table_{1,2,3,4}.py, with the same
directory names, files and formats, so the reference's gnuplot scripts read
the results unchanged.

The reference compiles and runs one executable per combination, one after the
other.  Here every combination of one precision runs as an instance of a
single batched Context: each instance keeps its own grid, dt, T, window and
counters, and one launch per step advances all of them.

    python -m fivonagi_amd.taylor_study RESULTS_PATH DEPLOY_PATH \\
        [--etas 5,10,20,41,82] [--cfls 99,90,80,70,60] [--precisions 1,2] \\
        [--angles 00000,PIo32,PIo16,PIo08,PIo04] [--final-time 100] [--grids]

writes RESULTS_PATH/d-eta{E}-cfl{C}-p{P}-angle{A}/ (the runs' exports), then
DEPLOY_PATH/dr_error_etas_CPU/ (error.npy and one u1-error.dat per run) and
DEPLOY_PATH/table_error_{1,2,3,4}/ (the tables).  The tables need the whole
default sweep (they fit a line over all five resolutions).
"""
import argparse
import math
import os
import re
import sys
import time

import numpy as np

from .study import ERROR_TYPES, Layout
from .study import write_tables as _write_tables

# run.sh:4-8 and plots/error_table_*.py:9-14
ANGLES = {"00000": 0.0, "PIo32": math.pi / 32, "PIo16": math.pi / 16, "PIo08": math.pi / 8,
          "PIo04": math.pi / 4}
ANGLE_TEX = {"00000": "0", "PIo32": "\\pi/32", "PIo16": "\\pi/16", "PIo08": "\\pi/8",
             "PIo04": "\\pi/4"}
ETA_MULT = {5: 1, 10: 2, 20: 4, 41: 8, 82: 16}           # run.sh:28-43: eta -> ETAC
ETAS = [5, 10, 20, 41, 82]
CFLS = [60, 70, 80, 90, 99]
PRECISIONS = [1, 2]


def exec_name(eta, cfl, p, angle):
    """run.sh:57: the run's directory is d-<exec_name>."""
    return "eta{0}-cfl{1}-p{2}-angle{3}".format(eta, cfl, p, angle)


def run_sweep(results_path, etas=ETAS, cfls=CFLS, precisions=PRECISIONS, angles=tuple(ANGLES),
              final_time=100.0, grids=False, log=print):
    """run.sh: every (precision, eta, cfl, angle) to final_time with the
    app's exports.  One batched Context per precision.  Existing run
    directories are skipped, as run.sh does.  Returns per precision the GPU
    seconds (step_until calls only), the cell updates and the wall seconds."""
    import fivonagi_amd as fa
    from fivonagi_amd import workloads as wl
    from fivonagi_amd.export import taylor_frames
    stats = {}
    for p in precisions:
        probs, paths = [], []
        for eta in etas:
            for cfl in cfls:
                for a in angles:
                    path = os.path.join(results_path, "d-" + exec_name(eta, cfl, p, a))
                    if os.path.isdir(path):
                        log("warning: deploy destination already exists for: %s. Not overwritten"
                            % exec_name(eta, cfl, p, a))
                        continue
                    os.makedirs(path)
                    probs.append(wl.taylor(ETA_MULT[eta], theta=ANGLES[a], moving=True,
                                           cflw=cfl / 100.0))
                    paths.append(path)
        if not probs:
            continue
        t_wall = time.perf_counter()
        st = {}
        with fa.Context(probs, fp64=(p == 2)) as ctx:
            taylor_frames(ctx, paths, final_time=final_time, grids=grids, stats=st)
            cells = sum(pr.nx * pr.ny * ctx.info(k)["n_attempts"] for k, pr in enumerate(probs))
        stats[p] = dict(instances=len(probs), gpu_s=st["step_s"], cell_updates=cells,
                        wall_s=time.perf_counter() - t_wall)
        log("p%d: %d runs, %.3g cell updates, %.2f s stepping, %.2f s in all"
            % (p, len(probs), cells, st["step_s"], stats[p]["wall_s"]))
    return stats


def _cut_errors_py(xi_numeric_analytic):
    """plots/error.py:76-105 on the last cut file: relative L1, L2, Linf in
    Python floats from the printed values."""
    numeric, analytic = xi_numeric_analytic[:, 2], xi_numeric_analytic[:, 3]
    out = []
    for et in ERROR_TYPES:
        t1 = t2 = 0.0
        for i in range(numeric.size):
            if et == "L1":
                t1 += abs(analytic[i] - numeric[i])
                t2 += abs(analytic[i])
            elif et == "L2":
                t1 += (analytic[i] - numeric[i]) ** 2
                t2 += analytic[i] ** 2
            else:
                t1 = max(t1, abs(analytic[i] - numeric[i]))
                t2 = max(t2, abs(analytic[i]))
        out.append((t1 / t2) ** 0.5 if et == "L2" else t1 / t2)
    return out


def extract_errors(results_path, deploy_path, etas=ETAS, cfls=CFLS, precisions=PRECISIONS,
                   angles=tuple(ANGLES), final_time=100):
    """plots/error.py: the errors of every run's last cut (which must be at
    T = final_time), into DEPLOY/dr_error_etas_CPU/error.npy (raw float64,
    shape (eta, cfl, p, angle, error type)) and one u1-error.dat per run.
    Returns the array.  Does nothing (returns None) if the directory exists,
    as error.py does."""
    out = os.path.join(deploy_path, "dr_error_etas_CPU")
    if os.path.isdir(out):
        return None
    os.makedirs(out)
    error = np.ndarray((len(etas), len(cfls), len(precisions), len(angles), len(ERROR_TYPES)))
    for ai, a in enumerate(angles):
        for ci, cfl in enumerate(cfls):
            for pi, p in enumerate(precisions):
                for ei, eta in enumerate(etas):
                    name = exec_name(eta, cfl, p, a)
                    ref = os.path.join(results_path, "d-" + name)
                    cuts = sorted(f for f in os.listdir(ref) if re.match(r"u1-cut-[0-9]+\.dat$", f))
                    last = os.path.join(ref, cuts[-1])
                    T = None
                    with open(last) as f:
                        for line in f:
                            m = re.search(r"^#\s*T\s*=\s*([0-9\.\-\+eE]+).*", line)
                            if m is not None:
                                T = float(m.group(1))
                                break
                    if T is None or int(T + 0.5) != int(final_time):
                        raise RuntimeError("wrong time T = %s in %s" % (T, last))
                    frame = int(re.search(r"u1-cut-([0-9]+)\.dat", cuts[-1]).group(1))
                    errs = _cut_errors_py(np.loadtxt(last, ndmin=2))
                    error[ei, ci, pi, ai, :] = errs
                    d = os.path.join(out, "d-" + name)
                    os.mkdir(d)
                    with open(os.path.join(d, "u1-error.dat"), "w") as f:
                        f.write("# frame T error_L1 error_L2 error_Li\n")
                        f.write("%d %f %f %f %f\n" % (frame, T, errs[0], errs[1], errs[2]))
    error.tofile(os.path.join(out, "error.npy"))
    return error


LAYOUT = Layout("angle", list(ANGLES), [ANGLE_TEX[a] for a in ANGLES],
                "cfl{0}-p{1}-angle{2}", "conv_rates_stats_angles-p{0}.dat", "\\theta_T=",
                "{:.2e}")


def write_tables(deploy_path, error=None, etas=ETAS, cfls=CFLS, precisions=PRECISIONS,
                 angles=tuple(ANGLES)):
    """plots/error_table_1.py .. error_table_4.py of the Taylor app from
    error.npy (study.write_tables with the angle as the grouping column)."""
    if error is None:
        error = np.fromfile(os.path.join(deploy_path, "dr_error_etas_CPU", "error.npy"))
        error = error.reshape((len(etas), len(cfls), len(precisions), len(angles),
                               len(ERROR_TYPES)))
    layout = Layout("angle", list(angles), [ANGLE_TEX[a] for a in angles], LAYOUT.exec_fmt,
                    LAYOUT.file34, LAYOUT.tex_prefix, LAYOUT.finest_fmt)
    return _write_tables(deploy_path, error, etas, cfls, precisions, layout)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("results_path")
    ap.add_argument("deploy_path")
    ap.add_argument("--etas", default=",".join(map(str, ETAS)))
    ap.add_argument("--cfls", default="99,90,80,70,60")
    ap.add_argument("--precisions", default="1,2")
    ap.add_argument("--angles", default=",".join(ANGLES))
    ap.add_argument("--final-time", type=float, default=100.0)
    ap.add_argument("--grids", action="store_true", help="also write u1-NNNN.dat every frame")
    a = ap.parse_args(argv)
    etas = [int(v) for v in a.etas.split(",")]
    cfls = [int(v) for v in a.cfls.split(",")]
    ps = [int(v) for v in a.precisions.split(",")]
    angles = a.angles.split(",")
    os.makedirs(a.results_path, exist_ok=True)
    os.makedirs(a.deploy_path, exist_ok=True)
    run_sweep(a.results_path, etas, cfls, ps, angles, a.final_time, a.grids)
    cfl_sorted = sorted(cfls)
    err = extract_errors(a.results_path, a.deploy_path, etas, cfl_sorted, ps, angles,
                         a.final_time)
    if len(etas) >= 2:
        write_tables(a.deploy_path, err, etas, cfl_sorted, ps, angles)
    return 0


if __name__ == "__main__":
    sys.exit(main())
