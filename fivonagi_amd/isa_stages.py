"""Static instruction counts of the step kernel's row loop, per stage.

    python -m fivonagi_amd.isa_stages

Compiles fivonagi_amd/csrc/fv_step.hip for gfx950 to assembly, finds the row
loop of the FP32 single-instance edge-drive kernel (the one the throughput is
quoted on; the loop is unrolled three times, one row per third), and counts
its VALU, transcendental and SALU instructions per row.  Then the same for
copies of the source with one stage reduced -- each limiter replaced by the
first-order fluctuations, each sweep or the source removed -- so that the
differences are the stages' costs.  The copies are compiled to assembly in a
temporary directory and never run; nothing here builds or loads a library.
Prints a markdown table (issue slots: a transcendental occupies two).
"""
import os
import re
import subprocess
import sys
import tempfile

from .build import ARCH, HERE

SRC = os.path.join(HERE, "csrc", "fv_step.hip")
KERNEL = ("_ZN2fv14fv_step_kernelIfLb1ELb0EN4fvbc9EdgeDriveEEEvPKNS_8InstDescEPKNS_8WorkItem"
          "EPKmiiPKNS_3CtlEPKNS_9SyncSlotsE")
FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-mllvm", "-amdgpu-kernarg-preload-count=16", "--cuda-device-only",
         "-S"]

# first-order fluctuations only: P = max(s, 0) W, M = min(s, 0) W per family
_FIRST_ORDER = (
    "{{ const fvna::WavesT<R>& w_ = {f}.w;\n"
    "  const R al1 = fvna::fmax2({f}.L1, R(0)), be1 = {f}.L1 - al1;\n"
    "  const R al2 = fvna::fmax2({f}.L2, R(0)), be2 = {f}.L2 - al2;\n"
    "  const R al3 = fvna::fmax2({f}.L3, R(0)), be3 = {f}.L3 - al3;\n"
    "  {P}[0] = fvna::fma_(al1, w_.a1, al2 * w_.a2);\n"
    "  {P}[1] = fvna::fma_(al1, w_.w21, al2 * w_.w22);\n"
    "  {P}[2] = fvna::fma_(al3, w_.a3, fvna::fma_(al1, w_.w31, al2 * w_.w32));\n"
    "  {M}[0] = fvna::fma_(be1, w_.a1, be2 * w_.a2);\n"
    "  {M}[1] = fvna::fma_(be1, w_.w21, be2 * w_.w22);\n"
    "  {M}[2] = fvna::fma_(be3, w_.a3, fvna::fma_(be1, w_.w31, be2 * w_.w32)); }}")
_Y_LIMITER = ("    fvna::limited_flux(fc.f, fc.dl[0], fn.dl[0], fc.dl[1], fn.dl[1], fc.dl[2], fn.dl[2], dty,\n"
              "                       fc.P, M);")
_X_LIMITER = "    fvna::limited_flux(fx, dl1, dr1, dl2, dr2, dl3, dr3, dtx, PX, MX);"
_X_START = "    const R ir = fvna::rcp(R(1) + y1q);\n    const R un = y2q * ir, ut = y3q * ir;"
_Y_RIEMANN = ("    fvna::riemann(bm.q1, bm.q3, bm.q2, bm.un, bm.ut, bn.q1, bn.q3, bn.q2, bn.un, bn.ut,\n"
              "                  hbeta, fn.f);")
_SOURCE = ("    R r2 = fvna::apply_source(mid.q2, i2);\n"
           "    R r3 = fvna::apply_source(mid.q3, i3);")

VARIANTS = {
    "full step": [],
    "y limiter -> first order": [(_Y_LIMITER, _FIRST_ORDER.format(f="fc.f", P="fc.P", M="M"))],
    "x limiter -> first order": [(_X_LIMITER, _FIRST_ORDER.format(f="fx", P="PX", M="MX"))],
    "no x sweep": [(_X_START,
                    "    xn.q1 = y1q; xn.q2 = y2q; xn.q3 = y3q;\n"
                    "    { const R ir2 = fvna::rcp(R(1) + xn.q1); xn.u = xn.q2 * ir2;"
                    " xn.v = xn.q3 * ir2; }\n    return;\n" + _X_START)],
    "no y sweep": [(_Y_RIEMANN,
                    "    fn.f = fvna::IfaceT<R>{}; fn.f.L1 = R(1); fn.f.L2 = R(-1);"
                    " fn.f.smax = R(1);\n    pin(bm.q1); pin(bn.q1);"),
                   (_Y_LIMITER, "    fc.P[0] = fc.P[1] = fc.P[2] = M[0] = M[1] = M[2] = R(0);")],
    "no source": [(_SOURCE, "    R r2 = mid.q2; R r3 = mid.q3; (void)i2; (void)i3;")],
}


def row_loop_counts(asm, kernel=KERNEL):
    """(VALU, transcendental, SALU) per row of the innermost loop holding the
    three unrolled rows' stores."""
    lines = asm.split("\n")
    start = next(i for i, l in enumerate(lines) if l.startswith(kernel + ":"))
    end = next(i for i in range(start, len(lines)) if lines[i].strip().startswith(".Lfunc_end"))
    body = lines[start:end]
    labels = {}
    for i, l in enumerate(body):
        m = re.match(r"^(\.LBB\d+_\d+):", l)
        if m:
            labels[m.group(1)] = i
    best = None
    for i, l in enumerate(body):
        m = re.search(r"s_(?:cbranch_\w+|branch)\s+(\.LBB\d+_\d+)", l)
        if not m or labels.get(m.group(1), i + 1) >= i:
            continue
        seg = [x.strip() for x in body[labels[m.group(1)]:i]]
        stores = sum(1 for x in seg if x.startswith("global_store"))
        valu = sum(1 for x in seg if x.startswith("v_"))
        if stores >= 9 and (best is None or valu < best[0]):
            trans = sum(1 for x in seg if re.match(r"v_(rcp|rsq|sqrt|exp|log|sin|cos)_", x))
            salu = sum(1 for x in seg if x.startswith("s_"))
            best = (valu, trans, salu)
    if best is None:
        raise RuntimeError("row loop not found")
    return tuple(v / 3.0 for v in best)


def compile_asm(text, tmp, tag):
    src = os.path.join(tmp, f"{tag}.hip")
    # the copy includes the kernel's headers by their paths from csrc/
    text = text.replace('#include "../apps/', f'#include "{os.path.join(HERE, "apps")}/')
    text = text.replace('#include "fv_types.h"', f'#include "{os.path.join(HERE, "csrc", "fv_types.h")}"')
    with open(src, "w") as f:
        f.write(text)
    out = os.path.join(tmp, f"{tag}.s")
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    r = subprocess.run([hipcc, *FLAGS, "-o", out, src], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-3000:])
    with open(out) as f:
        return f.read()


def main():
    base = open(SRC).read()
    rows = []
    with tempfile.TemporaryDirectory(prefix="fv_isa_") as tmp:
        for k, (name, subs) in enumerate(VARIANTS.items()):
            text = base
            for a, b in subs:
                if text.count(a) != 1:
                    raise RuntimeError(f"{name}: source text not found")
                text = text.replace(a, b)
            v, t, s = row_loop_counts(compile_asm(text, tmp, f"v{k}"))
            rows.append((name, v, t, s))
            print(f"  {name}: {v:.1f} VALU, {t:.1f} transcendental, {s:.1f} SALU per row",
                  file=sys.stderr, flush=True)
    full = rows[0][1] + rows[0][2]
    print("| variant | VALU/row | transcendental/row | issue slots/row | difference | SALU/row |")
    print("|---|---|---|---|---|---|")
    for name, v, t, s in rows:
        slots = v + t
        diff = "" if name == "full step" else f"{full - slots:.1f}"
        print(f"| {name} | {v:.1f} | {t:.1f} | {slots:.1f} | {diff} | {s:.1f} |")


if __name__ == "__main__":
    main()
