"""The C++ drop-in class (include/frm_solvers.hpp) runs the LOD solver on the
GPU and matches the oracle (plain FP64, 1e-12)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from theses_code_amd.solvers import coefs_advection

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "tests", "cpp", "tg_dropin")


@pytest.mark.gpu
def test_cpp_tg_solver_dropin():
    assert os.path.exists(BIN), "build() compiles tests/cpp/tg_dropin"
    rng = np.random.default_rng(9)
    sN, M, steps = 30, 25, 40
    nn = (sN + 2) * (M + 2)
    coefs = coefs_advection(rng.uniform(-0.1, 0.1, nn), rng.uniform(-0.1, 0.1, nn), 1e-3, 1 / sN)
    C0 = rng.uniform(0, 1, nn)
    hdr = [sN, M, steps, 0.6, 1e-3, 1.0, 0.2, 0.01]
    with tempfile.TemporaryDirectory() as d:
        fi, fo = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        np.concatenate([np.array(hdr, float), *coefs, C0]).tofile(fi)
        r = subprocess.run([BIN, fi, fo], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert f"tstep {steps}" in r.stdout
        got = np.fromfile(fo)
    o = O.TGSolver(sN, M, 0.6, 1e-3, 1.0, 0.2, 0.01, *coefs, C0)
    o.step(steps)
    ref = o.state().reshape(-1)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
