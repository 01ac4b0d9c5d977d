"""The fitting error of solve_and_test (mt_minimal.cpp:1277-1330) reduced on
the GPU for a batch of particles (one problem each, parameters varied as a
PSO iteration varies them) against the oracle's restatement of the loop,
problem by problem.  The solver agrees with the oracle to ~1e-13, so the
sums of squares are held to 1e-11 relative."""
import numpy as np
import pytest

from oracle import oracle as O
from theses_code_amd import _native as N
from theses_code_amd import solvers as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(0)
    yield c
    c.close()


def _particles(rng, n):
    out = []
    for q in range(n):
        Nn = int(rng.choice([40, 64, 97]))
        x = np.arange(Nn + 2) / Nn
        U0 = -5.0 + np.sin(np.pi * x) + 0.05 * rng.normal(size=Nn + 2)
        U0[0] = -5.0
        prob = dict(N=Nn, L=1.0, alpha=float(rng.uniform(0.55, 0.95)), form=N.FORM_CONST,
                    D=float(rng.uniform(5e-4, 5e-3)), top=N.TOP_DIRICHLET, top_val=-5.0,
                    bottom_cond=1, Hbottom=-5.0)
        out.append((prob, U0))
    return out


def test_fit_error_batch_matches_oracle(ctx):
    rng = np.random.default_rng(2024)
    tau, steps = 0.01, 60
    parts = _particles(rng, 8)
    nobs = 80
    T = rng.uniform(0.0, steps * tau, nobs)
    Z = rng.uniform(0.02, 0.95, nobs)  # inside (0, L - dL] for every N used
    F = -5.0 + rng.normal(0.0, 0.3, nobs)
    b = S.MT1DBatch(ctx, [p for p, _ in parts], tau, steps, [u for _, u in parts])
    b.set_observations(T, Z, F)
    b.calc_step(7)
    b.calc_step(steps - 7)
    got = b.fit_error()
    for i, (p, U0) in enumerate(parts):
        o = O.MTSolver(p["N"], p["L"], p["alpha"], tau, U0, form=p["form"], D=p["D"],
                       top=p["top"], top_val=p["top_val"], bottom_cond=p["bottom_cond"],
                       Hbottom=p["Hbottom"])
        want = O.mt_fit_error(o, p["N"], p["L"], tau, steps, T, Z, F)
        print(f"particle {i}: N={p['N']} alpha={p['alpha']:.3f} err {got[i]:.9e} oracle {want:.9e}")
        assert got[i] == pytest.approx(want, rel=1e-11)
    # set_state restarts the clock and clears the errors
    b.set_state([u for _, u in parts])
    b.calc_step(steps)
    np.testing.assert_allclose(b.fit_error(), got, rtol=1e-14)
    b.close()


def test_fit_error_requires_observations(ctx):
    rng = np.random.default_rng(1)
    (p, U0), = _particles(rng, 1)
    b = S.MT1DBatch(ctx, [p], 0.01, 5, [U0])
    with pytest.raises(N.FrmError):
        b.fit_error()
    b.set_observations([], [], [])
    b.calc_step(5)
    assert b.fit_error()[0] == 0.0
    b.close()
