"""The ia/ibk approximation plan (TG_solver dt_alg 4 as per-step operations
on per-node coefficient slots, frm_iaibk.cpp) executed on the CPU with the
GPU kernel's operation semantics, against the oracle's restatement of
calc_va (itself bit-exact against the reference's goldens): one node's
approximated sums over every step, rows and columns, at several orders.
Host code only (g++), no GPU."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "theses_code_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("iaibk") / "iaibk_check")
    # with the address and undefined-behaviour sanitizers: the planner indexes
    # its slot lists and value tables by hand
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-ffp-contract=off",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-I" + CSRC,
                    os.path.join(REPO, "tests", "cpp", "iaibk_check.cpp"),
                    os.path.join(CSRC, "frm_iaibk.cpp"), "-o", out], check=True)
    return out


@pytest.mark.parametrize("alpha", [0.3, 0.6, 0.9])
@pytest.mark.parametrize("direction", [0, 1])
def test_plan_matches_oracle(checker, alpha, direction):
    n = 700
    rng = np.random.default_rng(int(alpha * 10) + direction)
    # a smooth history with noise, increments of mixed sign
    F = 1e-3 * np.sin(0.05 * np.arange(n)) + 1e-5 * rng.normal(size=n)
    r = subprocess.run([checker, repr(alpha), str(direction)], input="\n".join("%.17g" % v for v in F),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    got = np.array([float(v) for v in r.stdout.split()])
    ref = O.dt4_series(alpha, direction, F)
    # the plan reproduces the reference's operations in its order; only the
    # evaluation order of the oracle's coefficient sums is the same too
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
