"""The C-ABI library builds for gfx950, loads, and exports every entry point
include/fivonagi.h declares.  No GPU calls (runs on CPU-only hosts)."""
import ctypes
import os
import re

import pytest

import fivonagi_amd as fa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "fivonagi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fivonagi_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_abi():
    names = declared_functions()
    assert set(fa.EXPORTS) == set(names), names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(fa.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), f"{name} not exported by {fa.LIB_PATH}"


def test_library_has_gfx950_code_object():
    data = open(fa.LIB_PATH, "rb").read()
    assert b"gfx950" in data


STRUCTS = (("fivonagi_params", fa.Params), ("fivonagi_info_t", fa.Info),
           ("fivonagi_objective", fa.Objective), ("fivonagi_adjoint_opts", fa.AdjointOpts),
           ("fivonagi_collect_opts", fa.CollectOpts),
           ("fivonagi_noalloc_state", fa.NoallocState))


def test_struct_layout_matches_header(tmp_path):
    """ctypes mirrors of fivonagi_params / fivonagi_info_t agree with the C
    compiler's layout of include/fivonagi.h (sizes and every field offset)."""
    import subprocess
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "fivonagi.h"', "int main(void){"]
    for cname, cls in STRUCTS:
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                  check=True).stdout.splitlines())
    for cname, cls in STRUCTS:
        assert int(out[cname]) == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(out[f"{cname}.{f}"]) == getattr(cls, f).offset, f"{cname}.{f}"


def test_argument_errors_do_not_touch_the_gpu():
    lib = fa.load()
    h = ctypes.c_void_p()
    # null params: rejected before any HIP call
    assert lib.fivonagi_init(ctypes.byref(h), None, 1, None) == fa.ERR_ARG
    assert b"null" in lib.fivonagi_last_error()
    p = fa.Params()
    p.nx, p.ny = 4, 4   # below the 8x8 minimum
    arr = (ctypes.POINTER(ctypes.c_float) * 3)()
    assert lib.fivonagi_init(ctypes.byref(h), ctypes.byref(p), 1, arr) == fa.ERR_ARG
    assert lib.fivonagi_step(None, 1) == fa.ERR_ARG
    assert lib.fivonagi_sync(None) == fa.ERR_ARG
    # the no-allocation step checks its arguments before any HIP call
    st = fa.NoallocState()
    st.dt = 0.01
    qs = (ctypes.c_void_p * 6)()
    assert lib.fivonagi_step_noalloc(None, qs, 1, ctypes.byref(st), None) == fa.ERR_ARG
    assert lib.fivonagi_step_noalloc(ctypes.byref(p), qs, 1, ctypes.byref(st), None) == fa.ERR_ARG
    assert b"null state" in lib.fivonagi_last_error()
    qs = (ctypes.c_void_p * 6)(*[ctypes.c_void_p(4096)] * 6)
    assert lib.fivonagi_step_noalloc(ctypes.byref(p), qs, 1, ctypes.byref(st), None) == fa.ERR_ARG
    assert b">= 8" in lib.fivonagi_last_error()


@pytest.mark.parametrize("fp64,nx,ny,ok", [(False, 32768, 32768, True), (False, 32768, 32769, False),
                                           (True, 32768, 16384, True), (True, 32768, 16385, False)])
def test_grid_size_limit_follows_the_byte_offsets(fp64, nx, ny, ok):
    # one field must fit 32-bit byte offsets: 2^30 cells in FP32, 2^29 in FP64.
    # A grid at the limit fails later (on the null state pointers), never at
    # the size check; the check itself touches no GPU
    lib = fa.load()
    h = ctypes.c_void_p()
    p = fa.Params()
    p.nx, p.ny, p.dx, p.dy, p.dt0, p.cfl_wish = nx, ny, 1.0, 1.0, 0.1, 0.9
    T = ctypes.c_double if fp64 else ctypes.c_float
    arr = (ctypes.POINTER(T) * 3)()
    init = lib.fivonagi_init_f64 if fp64 else lib.fivonagi_init
    assert init(ctypes.byref(h), ctypes.byref(p), 1, arr) == fa.ERR_ARG
    msg = lib.fivonagi_last_error()
    assert (b"too large" in msg) != ok, msg
