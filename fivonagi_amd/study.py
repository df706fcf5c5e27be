"""Convergence-rate tables shared by the two apps' studies: the four
plots/error_table_{1,2,3,4}.py scripts of apps/taylor-angle-xy and
apps/hifu-beam, which differ only in the parameter that groups the runs (the
angle, or the aperture a) and in a few names and number formats.

error[eta, cfl, p, group, type] holds the relative errors of every run
(type: L1, L2, Linf).  Per run, rate_last is log2 of the error ratio of the
last two resolutions, rate_reg the least-squares slope of log error against
-log eta (with its intercept and R^2).
"""
import os

import numpy as np

ERROR_TYPES = ["L1", "L2", "Li"]


class Layout:
    """What distinguishes the apps' tables.

    field: name of the grouping column (error_table_3.py's dtype field);
    labels: the groups as the tables print them; tex: their labels in table 4;
    exec_fmt: table 2's run name from (cfl, p, group); file34: file name of
    tables 3 and 4 for precision p; tex_prefix: table 4's group label prefix;
    finest_fmt: table 4's format of the finest errors."""

    def __init__(self, field, labels, tex, exec_fmt, file34, tex_prefix, finest_fmt):
        self.field, self.labels, self.tex = field, list(labels), list(tex)
        self.exec_fmt, self.file34 = exec_fmt, file34
        self.tex_prefix, self.finest_fmt = tex_prefix, finest_fmt


def rates(error, etas, ci, pi, gi, ei):
    """rate_last, rate_reg, intercept and R^2 of one run's errors over the
    resolutions (plots/error_table_1.py:62-74)."""
    from scipy import stats
    e = error[:, ci, pi, gi, ei]
    rate_last = (1 / np.log(2)) * np.log(e[-2] / e[-1])
    rate_reg, intercept, R, _, _ = stats.linregress(-np.log(etas), np.log(e))
    return rate_last, rate_reg, intercept, R ** 2


def write_tables(deploy_path, error, etas, cfls, precisions, layout):
    """error_table_1.py .. error_table_4.py: table 1 the rates over all groups
    and CFLs per precision and error type, table 2 every run's rates, table 3
    per group with the finest errors, table 4 the LaTeX rows of table 3.  A
    table whose directory exists is skipped, as the scripts do.  The "best"
    of a rate starts from 0 and the "worst" finest error from 0, as there."""
    groups = range(len(layout.labels))

    def table_dir(k):
        d = os.path.join(deploy_path, "table_error_%d" % k)
        if os.path.isdir(d):
            return None
        os.mkdir(d)
        return d

    def stats_over(pi, ei, group_list):
        s = dict(last=[], reg=[], R2=[], fin=[])
        for gi in group_list:
            for ci in range(len(cfls)):
                rl, rr, _, r2 = rates(error, etas, ci, pi, gi, ei)
                s["last"].append(rl)
                s["reg"].append(rr)
                s["R2"].append(r2)
                s["fin"].append(error[-1, ci, pi, gi, ei])
        return s

    d = table_dir(1)                                     # error_table_1.py
    if d:
        for pi, p in enumerate(precisions):
            with open(os.path.join(d, "conv_rates_stats%d.dat" % p), "w") as f:
                f.write("# error_type rate_last_best rate_last_worst rate_last_mean "
                        "rate_reg_best rate_reg_worst rate_reg_mean R2_reg_worst\n")
                for ei, et in enumerate(ERROR_TYPES):
                    s = stats_over(pi, ei, groups)
                    f.write("%s %f %f %f %f %f %f %f\n" % (
                        et, max(max(s["last"]), 0), min(s["last"]), np.mean(s["last"]),
                        max(max(s["reg"]), 0), min(s["reg"]), np.mean(s["reg"]), min(s["R2"])))
    d = table_dir(2)                                     # error_table_2.py
    if d:
        with open(os.path.join(d, "conv_rates_all.dat"), "w") as f:
            f.write("# error_type exec_name rate_last rate_reg intercept_reg R2_reg\n")
            for pi, p in enumerate(precisions):
                for ei, et in enumerate(ERROR_TYPES):
                    for gi, g in enumerate(layout.labels):
                        for ci, cfl in enumerate(cfls):
                            rl, rr, ic, r2 = rates(error, etas, ci, pi, gi, ei)
                            f.write("%s %s %f %f %f %f\n" % (
                                et, layout.exec_fmt.format(cfl, p, g), rl, rr, ic, r2))
    d = table_dir(3)                                     # error_table_3.py
    if d:
        for pi, p in enumerate(precisions):
            with open(os.path.join(d, layout.file34.format(p)), "w") as f:
                f.write("# error_type %s rate_last_best rate_last_worst rate_last_mean "
                        "rate_reg_best rate_reg_worst rate_reg_mean R2_reg_worst finest_best "
                        "finest_worst\n" % layout.field)
                for gi, g in enumerate(layout.labels):
                    f.write("# ------------------------ \n")
                    for ei, et in enumerate(ERROR_TYPES):
                        s = stats_over(pi, ei, [gi])
                        f.write("%s %s %f %f %f %f %f %f %f %f %f \n" % (
                            et, g, max(max(s["last"]), 0), min(s["last"]), np.mean(s["last"]),
                            max(max(s["reg"]), 0), min(s["reg"]), np.mean(s["reg"]),
                            min(s["R2"]), min(s["fin"]), max(max(s["fin"]), 0)))
    d = table_dir(4)                                     # error_table_4.py
    if d:
        fin = layout.finest_fmt
        for pi, p in enumerate(precisions):
            with open(os.path.join(d, layout.file34.format(p)), "w") as f:
                f.write("% rate_reg_best & rate_reg_worst & rate_reg_mean & R2_reg_worst & "
                        "finest_best & finest_worst \\\\\n")
                for gi, g in enumerate(layout.labels):
                    f.write("\\midrule\n")
                    f.write("\\multirow{2}{*}{$" + layout.tex_prefix + layout.tex[gi] +
                            "$} &\n")
                    for ei, et in enumerate(ERROR_TYPES):
                        if et == "L2":
                            continue
                        f.write("$E_1$ &\n" if et == "L1" else "& $E_\\infty$ &\n")
                        s = stats_over(pi, ei, [gi])
                        f.write(("{:.4f} & {:.4f} & {:.4f} & {:.4f} & " + fin + " & " + fin +
                                 " \\\\\n").format(
                            max(max(s["reg"]), 0), min(s["reg"]), np.mean(s["reg"]),
                            min(s["R2"]), min(s["fin"]), max(max(s["fin"]), 0)))
    return error
