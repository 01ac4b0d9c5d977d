"""The oracle's restatement of solve_and_test's error loop (mt_minimal.cpp:
1288-1330) against an independent computation: the same observations
scored from the oracle's per-step states with numpy.interp in depth, the
step of each observation found by bucketing, and the time blend written
out.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O


def _problem(rng, N, alpha):
    x = np.arange(N + 2) / N
    U0 = -5.0 + np.sin(np.pi * x) + 0.05 * rng.normal(size=N + 2)
    U0[0] = -5.0
    kw = dict(form=O.FORM_CONST, D=2e-3, top=O.TOP_DIRICHLET, top_val=-5.0, bottom_cond=1,
              Hbottom=-5.0)
    return U0, kw


@pytest.mark.parametrize("N,alpha,steps", [(40, 0.8, 50), (64, 0.6, 37)])
def test_fit_error_matches_independent_scoring(N, alpha, steps):
    rng = np.random.default_rng(N + steps)
    L, tau = 1.0, 0.01
    U0, kw = _problem(rng, N, alpha)
    # observations inside (0, L - dL]: the reference reads an unset value
    # for depths in the last interval
    nobs = 60
    T = rng.uniform(0.0, steps * tau, nobs)
    Z = rng.uniform(0.01, L * (N - 1) / N, nobs)
    F = -5.0 + rng.normal(0.0, 0.3, nobs)
    got = O.mt_fit_error(O.MTSolver(N, L, alpha, tau, U0, **kw), N, L, tau, steps, T, Z, F)
    # independent: per-step states, bucketed observations, numpy.interp
    s = O.MTSolver(N, L, alpha, tau, U0, **kw)
    states = [s.state()[:N + 1].copy()]
    for _ in range(steps):
        s.step(1)
        states.append(s.state()[:N + 1].copy())
    xs = np.arange(N + 1) * (L / N)
    k = np.floor(T / tau).astype(int)  # step k + 1 starts at k tau
    want = 0.0
    for i in np.argsort(k, kind="stable"):
        if k[i] >= steps:
            continue
        before = np.interp(Z[i], xs, states[k[i]])
        after = np.interp(Z[i], xs, states[k[i] + 1])
        k1 = (T[i] - k[i] * tau) / tau
        v = min((1 - k1) * before + k1 * after, 0.0)
        want += (F[i] - v) ** 2
    assert got == pytest.approx(want, rel=1e-12)
    assert got > 0
