"""fivonagi_step_noalloc (include/fivonagi.h) against fivonagi_step.

The entry point runs the same step kernels on a state the caller keeps on
the device, with its tables in the library's fixed device buffer, so its
results must equal a context's bit for bit: fields, T, dt, counts and the
moving-window offset.  Two calls in a row, the second on a smaller grid (the
buffer keeps the first call's tables; the second must not see them), a run
split over two calls, and a grid whose tables do not fit."""
import ctypes as C

import numpy as np
import pytest

import fivonagi_amd as fa
from fivonagi_amd import workloads as wl

pytestmark = pytest.mark.gpu


class Hip:
    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        for n in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize"):
            getattr(self.lib, n).restype = C.c_int

    def check(self, rc, what):
        assert rc == 0, f"{what}: hipError {rc}"

    def alloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.lib.hipMalloc(C.byref(p), C.c_size_t(nbytes)), "hipMalloc")
        return p.value

    def free(self, p):
        self.check(self.lib.hipFree(C.c_void_p(p)), "hipFree")

    def upload(self, dst, a):
        a = np.ascontiguousarray(a)
        self.check(self.lib.hipMemcpy(C.c_void_p(dst), a.ctypes.data_as(C.c_void_p),
                                      C.c_size_t(a.nbytes), 1), "hipMemcpy H2D")

    def download(self, src, shape):
        a = np.zeros(shape, np.float32)
        self.check(self.lib.hipMemcpy(a.ctypes.data_as(C.c_void_p), C.c_void_p(src),
                                      C.c_size_t(a.nbytes), 2), "hipMemcpy D2H")
        return a


@pytest.fixture(scope="module")
def hip():
    return Hip()


class DeviceState:
    """The caller's six device buffers for one instance (state + work)."""

    def __init__(self, hip, prob):
        self.hip, self.prob = hip, prob
        n = prob.nx * prob.ny * 4
        self.ptrs = [hip.alloc(n) for _ in range(6)]
        for f in range(3):
            hip.upload(self.ptrs[f], prob.q0[f])
            hip.upload(self.ptrs[3 + f], np.full((prob.ny, prob.nx), np.nan, np.float32))

    def fields(self):
        return np.stack([self.hip.download(self.ptrs[f], (self.prob.ny, self.prob.nx))
                         for f in range(3)])

    def close(self):
        for p in self.ptrs:
            self.hip.free(p)


def reference(prob, n):
    with fa.Context(prob) as ctx:
        ctx.step(n)
        q, st = ctx.collect(0)
        info = ctx.info(0)
    return q, st, info


def check_equal(prob, q, s, n):
    qr, sr, ir = reference(prob, n)
    assert s.n_accepted == sr["n_accepted"] and s.n_rejected == sr["n_rejected"]
    assert s.n_attempts == ir["n_attempts"]
    assert (s.md_x, s.md_y) == (ir["md_x"], ir["md_y"])
    assert s.T == sr["T"] and s.dt == sr["dt"] and s.error == 0
    assert np.array_equal(q, qr)


def test_two_calls_second_smaller(hip):
    """A 300x200 edge drive (fold subset as a bitmap), then a 96x64 pulse
    whose first attempts are rejected: each equals fivonagi_step bit for bit."""
    big = wl.edge_drive(300, 200, halfwidth_cells=25.0)
    small = wl.gauss_pulse(96, 64, sigma_cells=5.0, dt0=0.5)
    for prob, n in ((big, 40), (small, 25)):
        print(f"[noalloc] {prob.nx}x{prob.ny} bc={prob.bc} {n} steps", flush=True)
        ds = DeviceState(hip, prob)
        try:
            s = fa.step_noalloc(prob, ds.ptrs, n)
            q = ds.fields()
        finally:
            ds.close()
        check_equal(prob, q, s, n)
    assert s.n_rejected > 0


def test_split_over_calls_and_moving_window(hip):
    """The state carries over between calls: 17 + 23 steps equal 40, on a
    pulse whose window moves (the extended kernel)."""
    prob = wl.gauss_pulse(150, 90, sigma_cells=6.0, moving=(1.5, 0.5, 0.1))
    print("[noalloc] moving window 150x90, 17 + 23 steps", flush=True)
    ds = DeviceState(hip, prob)
    try:
        s = fa.step_noalloc(prob, ds.ptrs, 17)
        s = fa.step_noalloc(prob, ds.ptrs, 23, state=s)
        q = ds.fields()
    finally:
        ds.close()
    assert s.md_x > 3
    check_equal(prob, q, s, 40)


def test_tables_beyond_the_buffer_are_refused(hip):
    """A grid whose work tables exceed FIVONAGI_NOALLOC_SCRATCH_BYTES returns
    FIVONAGI_ERR_ARG before anything runs."""
    prob = wl.gauss_pulse(8, 8)
    prob.nx, prob.ny = 12000, 6000          # 207 strips x 6000 rows of lane masks
    print("[noalloc] 12000x6000 refused", flush=True)
    ptrs = [hip.alloc(prob.nx * prob.ny * 4) for _ in range(6)]
    try:
        with pytest.raises(fa.FivonagiError) as ei:
            fa.step_noalloc(prob, ptrs, 1)
        assert ei.value.code == fa.ERR_ARG
    finally:
        for p in ptrs:
            hip.free(p)
