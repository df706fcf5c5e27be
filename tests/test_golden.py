"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py):
the oracle must reproduce them bit for bit, and the HIP path must match them
to the parity tolerance with the same accepted/rejected counts."""
import glob
import hashlib
import json
import os

import numpy as np
import pytest

import fivonagi_amd as fa
from tests.conftest import rel_linf

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(glob.glob(os.path.join(HERE, "*.npz")))


def load(path):
    z = np.load(path, allow_pickle=False)
    p = json.loads(str(z["params"]))
    edge = z["edge_amp"]
    prob = fa.Problem(p["nx"], p["ny"], p["dx"], p["dy"], p["beta"], p["delta"], z["q0"],
                      cfl_max=p["cfl_max"], cfl_wish=p["cfl_wish"], dt0=p["dt0"], T0=p["T0"],
                      bc=p["bc"], cfl_mode=p["cfl_mode"], omega=p["omega"],
                      edge_amp=edge if edge.size else None, taylor=p.get("taylor"))
    return prob, z


def test_manifest_hashes():
    man = json.load(open(os.path.join(HERE, "manifest.json")))
    assert sorted(man) == sorted(os.path.basename(f) for f in FIXTURES)
    for f in FIXTURES:
        assert hashlib.sha256(open(f, "rb").read()).hexdigest() == man[os.path.basename(f)]


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_reproduces_fixture(oracle, path):
    from tests.oracle_bridge import to_oracle
    prob, z = load(path)
    q, st, rc = oracle.run(to_oracle(prob), prob.q0.astype(np.float64), int(z["steps"]))
    assert rc == 0
    assert np.array_equal(q, z["q"])
    assert st.T == float(z["T"]) and st.n_rejected == int(z["n_rejected"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_hip_matches_fixture(path):
    prob, z = load(path)
    with fa.Context(prob) as ctx:
        ctx.step(int(z["steps"]))
        q, st = ctx.collect(0)
    e = rel_linf(q, z["q"])
    print(f"[golden] {os.path.basename(path)}: rel Linf {e}", flush=True)
    assert st["n_accepted"] == int(z["n_accepted"]) and st["n_rejected"] == int(z["n_rejected"])
    assert abs(st["T"] - float(z["T"])) <= 1e-6 * abs(float(z["T"])) + 1e-9
    # taylor: rho v is ~1/500 of rho' (see test_gpu_parity.test_taylor)
    tol = [1e-5, 1e-5, 3e-4 if prob.bc == fa.BC_TAYLOR else 1e-5]
    assert all(e[k] < tol[k] for k in range(3)), e
