"""Pins the oracle's restatement of the Sum family of irrigation_fractional.cpp
(H_solver::time_summing, 4099-4170) without the reference: that file needs
libfftpack headers that are not part of the reference tree (includes at
265-267), so it cannot be built here and has no golden vectors.  The checks
are closed forms that a dropped term, a shifted weight index or a wrong power
would break:

* sum_alg 0 with the exact interval integrals as kb_3_cache and a history
  linear in time: the sum telescopes to the Caputo integral of a linear
  function, b ((T_{n+1})^{1-a} - (T_{n+1} - T_n)^{1-a}) / (1-a), on
  non-uniform steps (so tau[t] must pair with level pair t).
* sum_alg 1 equals sum_alg 0 over the levels newer than the first small
  weight (the crossing term excluded, as the CPU loop does), and equals
  sum_alg 0 when no weight is small.
* sum_alg 2 with the moments of x^j over each interval as coefficients
  converges to the same Caputo integral (the binomial series of
  (gtj - x)^-a), at a rate fixed by T_n / T_{n+1}."""
import numpy as np
import pytest

from oracle import oracle as O


def caputo_linear(b, T_end, T_n, a):
    return b * (T_end ** (1 - a) - (T_end - T_n) ** (1 - a)) / (1 - a)


def exact_kb(T, T_end, a):
    """kb(a, t, t+1, n+1) for g(t) = t: the integral of (T_end - x)^-a over
    [T[t], T[t+1]] (without the 1/Gamma factor the solver applies later)."""
    return ((T_end - T[:-1]) ** (1 - a) - (T_end - T[1:]) ** (1 - a)) / (1 - a)


@pytest.mark.parametrize("a,nlev", [(0.3, 2), (0.5, 17), (0.8, 64)])
def test_full_sum_is_caputo_of_linear(a, nlev):
    rng = np.random.default_rng(nlev)
    tau = rng.uniform(0.5e-3, 2e-3, nlev)  # tau[t]: step from T[t] to T[t+1]
    T = np.concatenate([[0.0], np.cumsum(tau)])[:nlev]
    T_end = T[-1] + tau[-1]  # Time[tstep+1]
    n = 5
    base, slope = rng.uniform(-1, 1, n), rng.uniform(-2, 2, n)
    old = base[None, :] + slope[None, :] * T[:, None]
    S = O.timesum01(0, 0.0, old, exact_kb(T, T_end, a), tau[:nlev - 1])
    ref = caputo_linear(slope, T_end, T[-1], a)
    assert np.max(np.abs(S - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_full_sum_pairs_tau_with_its_level_pair():
    # a shifted tau would not reproduce the slope on non-uniform steps
    tau = np.array([1e-3, 3e-3, 0.5e-3, 2e-3])
    T = np.concatenate([[0.0], np.cumsum(tau)])[:4]
    old = (2.0 * T)[:, None]
    kb = np.ones(3)
    S = O.timesum01(0, 0.0, old, kb, tau[:3])
    assert S[0] == pytest.approx(6.0, rel=1e-15)
    S_shift = O.timesum01(0, 0.0, old, kb, np.roll(tau[:3], 1))
    assert abs(S_shift[0] - 6.0) > 1.0


def test_restricted_sum_stops_before_the_first_small_weight():
    rng = np.random.default_rng(3)
    nlev, n = 30, 9
    old = rng.standard_normal((nlev, n))
    tau = rng.uniform(1e-3, 2e-3, nlev - 1)
    kb = np.linspace(0.1, 1.0, nlev - 1)  # increasing in t: newest weights largest
    eps = 0.5
    t0 = max(t for t in range(nlev - 1) if abs(kb[t]) < eps)
    S1 = O.timesum01(1, eps, old, kb, tau)
    S0_newer = O.timesum01(0, 0.0, old[t0 + 1:], kb[t0 + 1:], tau[t0 + 1:])
    assert np.array_equal(S1, S0_newer)
    # the crossing term itself is not added
    S0_with = O.timesum01(0, 0.0, old[t0:], kb[t0:], tau[t0:])
    assert not np.allclose(S1, S0_with)
    # no small weight: the full sum
    assert np.array_equal(O.timesum01(1, 0.05, old, kb, tau), O.timesum01(0, 0.0, old, kb, tau))


def test_restricted_sum_with_newest_weight_small_is_zero():
    old = np.arange(12.0).reshape(4, 3)
    S = O.timesum01(1, 1.0, old, np.array([5.0, 5.0, 0.5]), np.ones(3))
    assert np.array_equal(S, np.zeros(3))


@pytest.mark.parametrize("a", [0.2, 0.6, 0.9])
def test_series_sum_converges_to_caputo_of_linear(a):
    # uniform steps; at step k the newest interval is [T_{k-1}, T_k].  The
    # series is not normalised (v1_j ~ gtj^-j, moments ~ T^j), so, as in the
    # reference, times stay near 1 for 400 terms to stay finite.
    tau = 0.25
    n, nterms = 4, 400
    rng = np.random.default_rng(int(a * 10))
    slope = rng.uniform(-2, 2, n)
    ts = O.TimeSum2(n, nterms)
    for k in range(1, 9):
        T0, T1, T2 = (k - 1) * tau, k * tau, (k + 1) * tau
        j = np.arange(nterms)
        coefs = (1.0 - (T0 / T1) ** (j + 1)) / (j + 1)  # int x^j over [T0,T1] / T1^(j+1)
        diff = slope * tau
        S = ts(diff, coefs, gtj=T2, gt=T1, alpha=a, tau_n=tau)
        # sum_j v1_j H_j = sum_k diff_k int_{I_k} (gtj - x)^-a dx, diff_k = slope tau
        ref = caputo_linear(slope, T2, T1, a)
        # the tail of the series is below (T1/T2)^nterms
        assert np.max(np.abs(S - ref)) <= 1e-10 * np.max(np.abs(ref)), k


def test_series_sum_uses_only_the_coefficients_given():
    # fewer coefficients than terms: the missing moments stay zero
    ts = O.TimeSum2(1, 6)
    S = ts(np.array([2.0]), np.array([1.0, 1.0]), gtj=2.0, gt=1.0, alpha=0.5, tau_n=1.0)
    # H = [2, 2, 0, 0, 0, 0]; v1 = 2^-0.5 * [1, 0.25, ...]
    assert S[0] == pytest.approx(2 ** -0.5 * (2 + 2 * 0.25), rel=1e-15)
    assert np.array_equal(ts.H[2:, 0], np.zeros(4))
