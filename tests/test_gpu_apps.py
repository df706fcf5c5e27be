"""The hifu-beam app's C host program (fivonagi_amd/apps/hifu-beam/driver.c,
built against libfivonagi.so) on the GPU: its frame loop with the moving
window and data collection must leave the same state and measurement grid,
bit for bit, as the same app set up from Python (workloads.hifu), and the
same state as the FP64 oracle within the FP32 tolerance."""
import os
import subprocess

import numpy as np
import pytest

import fivonagi_amd as fa
from fivonagi_amd import build as fb
from fivonagi_amd import workloads as wl
from tests.conftest import rel_linf
from tests.oracle_bridge import to_oracle

pytestmark = pytest.mark.gpu


def test_hifu_driver_matches_the_python_path(tmp_path, oracle):
    assert os.path.exists(fb.DRIVER), "build first: python -m fivonagi_amd.build"
    nx, ny, frames, ispam = 128, 256, 25, 10.0
    out = tmp_path / "hifu.bin"
    print(f"[apps] driver 1 {ispam} {frames} nx={nx} ny={ny}", flush=True)
    r = subprocess.run([fb.DRIVER, "1", str(ispam), str(frames), str(out), str(nx), str(ny)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert len(lines) == frames
    n = nx * ny
    raw = np.fromfile(out, np.float32)
    q, m = raw[:3 * n].reshape(3, ny, nx), raw[3 * n:].reshape(ny, nx)

    prob = wl.hifu(nx=nx, ny=ny, ispam=ispam)
    with fa.Context(prob) as ctx:
        ctx.set_collect(0, **wl.hifu_collect(prob))
        T = prob.T0
        for f in range(frames):
            T += 1.0
            ctx.step_until(T)
            info = ctx.info(0)
            want = (f"frame {f} T {info['T']:.6f} accepted {info['n_accepted']} rejected "
                    f"{info['n_rejected']} MDX {info['md_x']}")
            assert lines[f].startswith(want), (lines[f], want)
        qp, st = ctx.collect(0)
        mp = ctx.measure(0)
    assert np.array_equal(q, qp) and np.array_equal(m, mp)
    assert st["n_accepted"] > 0 and info["md_x"] > 0 and np.abs(m[6:]).max() > 0

    qo, sto, rc = oracle.run(to_oracle(prob, T_stop=T), prob.q0.astype(np.float64), 10 ** 6)
    assert rc == 0 and sto.n_accepted == st["n_accepted"]
    e = rel_linf(q, qo)
    print(f"[apps] C driver vs FP64 oracle after {frames} frames: " +
          " ".join(f"{v:.3e}" for v in e), flush=True)
    assert max(e) < 1e-5
