"""The hifu-beam study: the sweep of apps/hifu-beam/run.sh:4-76 (precision x
resolution x CFL x aperture, each run to finalTime with the app's moving
window and data collection, keeping the run's last u1max.dat as
<exec_name>.dat, run.sh:66-70), the comparison of every run's on-axis maximum
with the digitized curves of Albin et al. (2012) that plots/error_collect.py
makes, and the convergence tables of plots/error_table_{1,2,3,4}.py, with the
same directory names, files and formats.

The reference compiles and runs one executable per combination.  Here every
combination of one precision is an instance of a single batched Context.

    python -m fivonagi_amd.hifu_study RESULTS_PATH DEPLOY_PATH [REFERENCE_DATA] \\
        [--etas 20,41,82] [--cfls 99,90,80,70,60] [--precisions 1,2] \\
        [--apertures 30,20,10] [--final-time 100] [--exports]

REFERENCE_DATA is the directory of the digitized curves (data-2-a{a}-d.dat,
x in focal lengths and u1/AMPL, as apps/hifu-beam/reference_data/extracted
holds them); without it only the runs are made.  The errors go to
DEPLOY_PATH/dr_error_collect/ (error.npy and u1-error-<exec_name>.dat), the
tables to DEPLOY_PATH/table_error_{1,2,3,4}/.
"""
import argparse
import os
import sys
import time

import numpy as np

from .study import ERROR_TYPES, Layout
from .study import write_tables as _write_tables

# run.sh:4-7, :28-39 and plots/error_collect.py:9-21
ETA_MULT = {20: 1, 41: 2, 82: 4}                # eta -> ETAC (NX = 512 ETAC)
ETAS = [20, 41, 82]
CFLS = [60, 70, 80, 90, 99]
PRECISIONS = [1, 2]
APERTURES = [10, 20, 30]
AMPL = 0.0004217
X_MIN_ERR, X_MAX_ERR = 0.1, 1.8                 # focal lengths, every aperture

LAYOUT = Layout("a", [str(a) for a in APERTURES], [str(a) for a in APERTURES],
                "cfl{0}-p{1}-a{2}", "conv_rates_stats_as-p{0}.dat", "a=", "{:.4f}")


def exec_name(eta, cfl, p, a):
    """run.sh:44: the run's directory is d-<exec_name>, its result
    <exec_name>.dat."""
    return "eta{0}-cfl{1}-p{2}-a{3}".format(eta, cfl, p, a)


def run_sweep(results_path, etas=ETAS, cfls=CFLS, precisions=PRECISIONS,
              apertures=(30, 20, 10), final_time=100.0, exports=False, log=print):
    """run.sh: every (precision, eta, cfl, aperture) to final_time, one
    batched Context per precision.  Each run's directory gets <exec_name>.dat,
    the u1max.dat of its last frame (run.sh renames it so); with exports the
    app's other files are written every frame too (without the whole-grid
    u1-NNNN.dat).  Existing run directories are skipped, as run.sh does.
    Returns per precision the stepping seconds, the cell updates and the wall
    seconds."""
    import fivonagi_amd as fa
    from fivonagi_amd import workloads as wl
    from fivonagi_amd.export import frame_loop, write_u1max
    stats = {}
    for p in precisions:
        probs, paths, names = [], [], []
        for eta in etas:
            for cfl in cfls:
                for a in apertures:
                    name = exec_name(eta, cfl, p, a)
                    path = os.path.join(results_path, "d-" + name)
                    if os.path.isdir(path):
                        log("warning: deploy destination already exists for: %s. Not overwritten"
                            % name)
                        continue
                    os.makedirs(path)
                    probs.append(wl.hifu(ETA_MULT[eta], ispam=float(a), cflw=cfl / 100.0))
                    paths.append(path)
                    names.append(name)
        if not probs:
            continue
        t_wall = time.perf_counter()
        with fa.Context(probs, fp64=(p == 2)) as ctx:
            exps = []
            for k, pr in enumerate(probs):
                ctx.set_collect(k, **wl.hifu_collect(pr))
                if exports:
                    exps.append(_FramesNoGrid(paths[k], pr))

            def on_frame(k, frame, info):
                if exports:
                    exps[k](frame, ctx.measure(k, rows=_u1max_rows(probs[k], info)), info["T"],
                            info["n_accepted"], info["md_x"], info["md_y"])

            step_s = frame_loop(ctx, final_time, 1.0, on_frame)
            cells = 0
            for k, pr in enumerate(probs):
                info = ctx.info(k)
                cells += pr.nx * pr.ny * info["n_attempts"]
                m = ctx.measure(k, rows=_u1max_rows(pr, info))
                write_u1max(os.path.join(paths[k], names[k] + ".dat"), m.reshape(-1), pr.nx,
                            info["md_x"], pr.hifu["f"], pr.hifu["ampl"])
        stats[p] = dict(instances=len(probs), gpu_s=step_s, cell_updates=cells,
                        wall_s=time.perf_counter() - t_wall)
        log("p%d: %d runs, %.3g cell updates, %.2f s stepping, %.2f s in all"
            % (p, len(probs), cells, step_s, stats[p]["wall_s"]))
    return stats


def _u1max_rows(prob, info):
    """Rows of the measurement grid up to the last u1max entry (flat index
    7 nx + MDX - 1)."""
    return min(prob.ny, 8 + info["md_x"] // prob.nx)


class _FramesNoGrid:
    """HifuExport's per-frame files without the two made from the state."""

    def __init__(self, path, prob):
        from fivonagi_amd.export import HifuExport
        from fivonagi_amd.workloads import hifu_collect
        jcut = hifu_collect(prob)["jcut"]
        self.ex = HifuExport(path, prob.hifu["f"], prob.hifu["ampl"], prob.cfl_wish, jcut)

    def __call__(self, frame, measure, T, n, mdx, mdy):
        self.ex(frame, None, measure, T, n, mdx, mdy)


def _errors_against(ref_x, ref_u, num_x, num_u):
    """plots/error_collect.py:54-104 for one run: the reference curve
    interpolated linearly at the run's abscissae (those up to the curve's
    last x; the script's loop then drops one more), relative L1, L2 and Linf
    errors over X_MIN_ERR <= x <= X_MAX_ERR."""
    from scipy import interpolate
    f = interpolate.interp1d(ref_x, ref_u, kind="slinear")
    i = 0
    for x in num_x:
        if x <= f.x[-1]:
            i += 1
        else:
            i -= 1
            break
    int_x = num_x[0:i]
    int_u = f(int_x)
    out = []
    for et in ERROR_TYPES:
        t1 = t2 = 0.0
        for j in range(int_x.size):
            if int_x[j] >= X_MIN_ERR and int_x[j] <= X_MAX_ERR:
                if et == "L1":
                    t1 += abs(int_u[j] - num_u[j])
                    t2 += abs(int_u[j])
                elif et == "L2":
                    t1 += (int_u[j] - num_u[j]) ** 2
                    t2 += int_u[j] ** 2
                else:
                    t1 = max(t1, abs(int_u[j] - num_u[j]))
                    t2 = max(t2, abs(int_u[j]))
        out.append((t1 / t2) ** 0.5 if et == "L2" else t1 / t2)
    return out


def error_collect(results_path, deploy_path, reference_path, etas=ETAS, cfls=CFLS,
                  precisions=PRECISIONS, apertures=APERTURES):
    """plots/error_collect.py: every run's <exec_name>.dat (u1max / AMPL)
    against data-2-a{a}-d.dat of reference_path, into
    DEPLOY/dr_error_collect/error.npy (raw float64, shape (eta, cfl, p, a,
    error type)) and u1-error-<exec_name>.dat.  Returns the array; does
    nothing (returns None) if the directory exists, as the script does."""
    out = os.path.join(deploy_path, "dr_error_collect")
    if os.path.isdir(out):
        return None
    os.makedirs(out)
    error = np.ndarray((len(etas), len(cfls), len(precisions), len(apertures), len(ERROR_TYPES)))
    for ai, a in enumerate(apertures):
        ref = np.loadtxt(os.path.join(reference_path, "data-2-a{0}-d.dat".format(a)))
        for ei, eta in enumerate(etas):
            for ci, cfl in enumerate(cfls):
                for pi, p in enumerate(precisions):
                    name = exec_name(eta, cfl, p, a)
                    num = np.loadtxt(os.path.join(results_path, "d-" + name, name + ".dat"))
                    errs = _errors_against(ref[:, 0], ref[:, 1], num[:, 0], num[:, 1] / AMPL)
                    error[ei, ci, pi, ai, :] = errs
                    with open(os.path.join(out, "u1-error-" + name + ".dat"), "w") as f:
                        f.write("# error_L1 error_L2 error_Li\n")
                        f.write("%f %f %f\n" % tuple(errs))
    error.tofile(os.path.join(out, "error.npy"))
    return error


def write_tables(deploy_path, error=None, etas=ETAS, cfls=CFLS, precisions=PRECISIONS,
                 apertures=APERTURES):
    """apps/hifu-beam/plots/error_table_1.py .. error_table_4.py from
    dr_error_collect/error.npy (study.write_tables with the aperture as the
    grouping column)."""
    if error is None:
        error = np.fromfile(os.path.join(deploy_path, "dr_error_collect", "error.npy"))
        error = error.reshape((len(etas), len(cfls), len(precisions), len(apertures),
                               len(ERROR_TYPES)))
    labels = [str(a) for a in apertures]
    layout = Layout("a", labels, labels, LAYOUT.exec_fmt, LAYOUT.file34, LAYOUT.tex_prefix,
                    LAYOUT.finest_fmt)
    return _write_tables(deploy_path, error, etas, cfls, precisions, layout)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("results_path")
    ap.add_argument("deploy_path")
    ap.add_argument("reference_data", nargs="?", default=None)
    ap.add_argument("--etas", default=",".join(map(str, ETAS)))
    ap.add_argument("--cfls", default="99,90,80,70,60")
    ap.add_argument("--precisions", default="1,2")
    ap.add_argument("--apertures", default="30,20,10")
    ap.add_argument("--final-time", type=float, default=100.0)
    ap.add_argument("--exports", action="store_true",
                    help="also write the per-frame files (cuts, microphones, cfl)")
    a = ap.parse_args(argv)
    etas = [int(v) for v in a.etas.split(",")]
    cfls = [int(v) for v in a.cfls.split(",")]
    ps = [int(v) for v in a.precisions.split(",")]
    aps = [int(v) for v in a.apertures.split(",")]
    os.makedirs(a.results_path, exist_ok=True)
    os.makedirs(a.deploy_path, exist_ok=True)
    run_sweep(a.results_path, etas, cfls, ps, aps, a.final_time, a.exports)
    if a.reference_data:
        err = error_collect(a.results_path, a.deploy_path, a.reference_data, etas, sorted(cfls),
                            ps, sorted(aps))
        if err is not None and len(etas) >= 2:
            write_tables(a.deploy_path, err, etas, sorted(cfls), ps, sorted(aps))
    return 0


if __name__ == "__main__":
    sys.exit(main())
