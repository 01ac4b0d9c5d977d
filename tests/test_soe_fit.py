"""The sum-of-exponentials fit (frm_soe.cpp) built for the host with
AddressSanitizer and UndefinedBehaviorSanitizer: histories shorter than the
number of modes (capacity - 1 < terms, where the least-squares system has
fewer rows than columns) and the long histories of the configurations.  The
fitted sum must reproduce the weights within the chosen forms' 1e-6 bar, and
the fit must stay inside its arrays.  Host code only (g++), no GPU."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "theses_code_amd", "csrc")


@pytest.fixture(scope="module")
def fitter(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("soefit") / "soe_fit_check")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
                    "-I" + os.path.join(REPO, "include"), "-I" + CSRC,
                    os.path.join(REPO, "tests", "cpp", "soe_fit_check.cpp"),
                    os.path.join(CSRC, "frm_soe.cpp"), os.path.join(CSRC, "frm_weights.cpp"),
                    "-o", out], check=True)
    return out


def fit(fitter, alpha, tau, kernel, pairs):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:exitcode=86")
    r = subprocess.run([fitter, repr(alpha), repr(tau), kernel] + [f"{K}:{L}" for K, L in pairs],
                       capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    rows = [ln.split() for ln in r.stdout.splitlines()]
    assert len(rows) == len(pairs)
    return {(int(a), int(b)): (float(c), float(d)) for a, b, c, d in rows}


@pytest.mark.parametrize("alpha", [0.6, 0.7])
def test_short_histories_caputo(fitter, alpha):
    # capacity K with K - 1 fitted lags (lag 1 is exact); 20:32 is the case
    # whose back-substitution used to read past the least-squares arrays
    pairs = [(K, L) for L in (32, 64) for K in (2, 3, 5, 8, 20, L, L + 1, L + 2)]
    res = fit(fitter, alpha, 1e-3, "caputo", pairs)
    for (K, L), (err, cmax) in res.items():
        assert err <= 1e-8, (K, L, err)
        assert cmax < 1e8, (K, L, cmax)


def test_short_and_long_histories_ab(fitter):
    res = fit(fitter, 0.5, 1e-3, "ab", [(20, 64), (63, 64), (64, 64), (300, 64)])
    for key, (err, _) in res.items():
        assert err <= 1e-10, (key, err)


def test_configuration_histories(fitter):
    # c1/c2 (32 terms over 5000 and 4000 steps) and c4 (64 terms, 2000 steps)
    res = fit(fitter, 0.6, 1e-3, "caputo", [(4000, 32), (5000, 32)])
    assert all(err <= 1e-8 for err, _ in res.values()), res
    res = fit(fitter, 0.3, 1e-3, "ab", [(2000, 64)])
    assert all(err <= 1e-10 for err, _ in res.values()), res
