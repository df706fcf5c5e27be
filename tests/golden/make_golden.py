"""Writes the golden fixtures of this directory with the FP64 oracle.

    python tests/golden/make_golden.py

The reference ships no test vectors (it has no tests), so these fixtures are
oracle outputs on small seeded inputs: they pin the oracle against regressions
(bit-exact, tests/test_golden.py) and let the GPU tests check the HIP path
without rebuilding the oracle's answer.  Each .npz holds the problem
parameters, the initial state and the state/counters after `steps` accepted
steps; manifest.json lists their SHA-256.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fivonagi_amd import workloads as wl  # noqa: E402
from oracle import pyoracle as po  # noqa: E402
from tests.oracle_bridge import to_oracle  # noqa: E402


def cases():
    return {
        "pulse_32x32_20": (wl.gauss_pulse(32, 32, sigma_cells=3.0), 20),
        "pulse_48x40_fold_40": (wl.gauss_pulse(48, 40, sigma_cells=4.0, amp=0.02), 40),
        "edge_64x32_50": (wl.edge_drive(64, 32, halfwidth_cells=8.0), 50),
        "taylor_64_30": (wl.taylor(1, theta=0.3), 30),
        "reject_32x24_10": (wl.gauss_pulse(32, 24, sigma_cells=3.0, dt0=0.5), 10),
    }


def params_of(p):
    d = dict(nx=p.nx, ny=p.ny, dx=p.dx, dy=p.dy, beta=p.beta, delta=p.delta,
             cfl_max=p.cfl_max, cfl_wish=p.cfl_wish, dt0=p.dt0, T0=p.T0, bc=p.bc,
             cfl_mode=p.cfl_mode, omega=p.omega)
    if p.taylor:
        d["taylor"] = dict(p.taylor)
    return d


def main():
    manifest = {}
    for name, (p, steps) in cases().items():
        q, st, rc = po.run(to_oracle(p), p.q0.astype(np.float64), steps)
        assert rc == 0, name
        path = os.path.join(HERE, name + ".npz")
        np.savez(path, params=json.dumps(params_of(p)),
                 edge_amp=p.edge_amp if p.edge_amp is not None else np.zeros(0),
                 q0=p.q0, q=q, steps=steps, T=st.T, dt=st.dt,
                 n_accepted=st.n_accepted, n_rejected=st.n_rejected)
        manifest[name + ".npz"] = hashlib.sha256(open(path, "rb").read()).hexdigest()
        print(name, "T=%.6f rejected=%d" % (st.T, st.n_rejected))
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
