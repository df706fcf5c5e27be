"""Host-runtime behaviour of the C ABI on the GPU: the prescribed dt sequence
at its exact length, the lifetime of the caller's edge arrays, the autotuner
beside data collection and a dt sequence, and a refused dt sequence leaving
the context as it was.  Each is checked against the FP64 oracle or against the
same run without the call under test."""
import numpy as np
import pytest

import fivonagi_amd as fa
from fivonagi_amd import workloads as wl
from tests.conftest import rel_linf
from tests.oracle_bridge import to_oracle

pytestmark = pytest.mark.gpu

TOL = 1e-5


def test_prescribed_dt_exact_length(oracle):
    """A sequence of exactly n entries runs n steps (attempt k uses dt[k]);
    the next attempt has no entry and the call that needs it returns
    FIVONAGI_ERR_DTSEQ (fivonagi_set_dt_sequence, include/fivonagi.h)."""
    prob = wl.gauss_pulse(64, 64, sigma_cells=5.0)
    n = 40
    dts = np.linspace(0.03, 0.045, n)
    with fa.Context(prob) as ctx:
        ctx.set_dt_sequence(0, dts)
        ctx.step(n)
        q, st = ctx.collect(0)
        assert ctx.sync() == 0
        with pytest.raises(fa.FivonagiError) as ei:
            ctx.step(1)
        assert ei.value.code == fa.ERR_DTSEQ
        _, st2 = ctx.collect(0)
    op = to_oracle(prob, dt_seq=dts)
    qo, sto, rc = oracle.run(op, prob.q0.astype(np.float64), n)
    assert rc == 0 and sto.n_accepted == n
    assert st["n_accepted"] == n and st["T"] == pytest.approx(sto.T, rel=1e-12)
    assert st2["n_accepted"] == n                  # the refused attempt did not run
    e = rel_linf(q, qo)
    print(f"[runtime] exact-length dt sequence: {e}", flush=True)
    assert max(e) < TOL


def test_edge_arrays_are_copied_at_init(oracle):
    """The caller may overwrite its edge_amp / edge_phase arrays after init:
    the context runs (and differentiates) the drive it was given."""
    prob = wl.edge_drive(200, 96, halfwidth_cells=20.0,
                         edge_phase=np.linspace(0.0, 0.3, 96))
    amp0, ph0 = prob.edge_amp.copy(), prob.edge_phase.copy()
    with fa.Context(prob) as ctx:
        prob.edge_amp[:] = 7.0                   # the caller reuses its arrays
        prob.edge_phase[:] = -1.0
        ctx.step(50)
        q, st = ctx.collect(0)
    prob.edge_amp[:] = amp0
    prob.edge_phase[:] = ph0
    op = to_oracle(prob)
    qo, sto, rc = oracle.run(op, prob.q0.astype(np.float64), 50)
    assert rc == 0
    e = rel_linf(q, qo)
    print(f"[runtime] edge arrays overwritten after init: {e}", flush=True)
    assert max(e) < TOL


def test_autotune_beside_collection_and_dt_sequence():
    """Autotuning a context that collects data and runs a prescribed dt
    sequence leaves the measurement grid, the overflow flag and the sequence
    position untouched: the frames after it equal those of an untuned run."""
    prob = wl.hifu(nx=128, ny=256, ispam=10.0)
    opts = wl.hifu_collect(prob)
    dts = np.full(4000, 0.035)

    def run(tune):
        with fa.Context(prob) as ctx:
            ctx.set_dt_sequence(0, dts)
            ctx.set_collect(0, **opts)
            T = prob.T0
            for _ in range(20):
                T += 1.0
                ctx.step_until(T)
            if tune:
                ctx.autotune(3)
            for _ in range(12):
                T += 1.0
                ctx.step_until(T)
            return ctx.collect(0), ctx.measure(0), ctx.info(0)
    (qa, sa), ma, ia = run(False)
    (qb, sb), mb, ib = run(True)
    assert sa == sb and ia["n_attempts"] == ib["n_attempts"]
    assert np.array_equal(qa, qb)
    assert np.array_equal(ma, mb)
    assert np.abs(ma[6:]).max() > 0


def test_refused_dt_sequence_changes_nothing():
    """A dt sequence shorter than the attempts already made is refused with
    FIVONAGI_ERR_ARG and leaves the CFL feedback running as before."""
    prob = wl.gauss_pulse(64, 64, sigma_cells=5.0)
    with fa.Context(prob) as ctx:
        ctx.step(10)
        with pytest.raises(fa.FivonagiError) as ei:
            ctx.set_dt_sequence(0, np.full(5, 0.01))
        assert ei.value.code == fa.ERR_ARG
        ctx.step(10)
        qa, sa = ctx.collect(0)
    with fa.Context(prob) as ctx:
        ctx.step(20)
        qb, sb = ctx.collect(0)
    assert sa == sb and np.array_equal(qa, qb)
