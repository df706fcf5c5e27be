"""Build recipe for the HIP library (gfx950) and, for the tests, the CPU oracle.

    python -m fivonagi_amd.build          # build libfivonagi.so in-tree

The library is compiled with hipcc directly (no torch extension machinery):
it is a plain C-ABI shared object, include/fivonagi.h.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libfivonagi.so")
SOURCES = [os.path.join(HERE, "csrc", n) for n in ("fv_step.hip", "fv_adjoint_march.hip", "fv_scratch.hip", "fv_tables.cpp",
                                                     "fv_runtime.cpp")]
ARCH = os.environ.get("FIVONAGI_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed FP32 (v_pk_*_f32) anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # kernel arguments arrive preloaded in SGPRs (gfx950): a wave's first
    # loads need no round trip for the argument block
    "-mllvm", "-amdgpu-kernarg-preload-count=16",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, cwd=None):
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_lib(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    out = _run([hipcc, *HIPCC_FLAGS, *SOURCES, "-o", tmp])
    os.replace(tmp, LIB)
    if verbose and out.strip():
        print(out)
    return LIB


DRIVER_SRC = os.path.join(HERE, "apps", "hifu-beam", "driver.c")
DRIVER = os.path.join(HERE, "apps", "hifu-beam", "driver")


def build_driver():
    """The hifu-beam app's C host program, linked against libfivonagi.so."""
    _run(["gcc", "-std=c99", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), DRIVER_SRC,
          "-L", HERE, "-lfivonagi", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/../..", "-lm",
          "-o", DRIVER])
    return DRIVER


def build_oracle():
    """The CPU oracle is test infrastructure; building it is not using it."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all(verbose=False):
    build_lib(verbose)
    build_driver()
    build_oracle()


if __name__ == "__main__":
    build_all(verbose=True)
    print("built", LIB)
    sys.exit(0)
