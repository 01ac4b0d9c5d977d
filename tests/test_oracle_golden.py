"""The CPU oracle against the reference's own golden vectors and closed forms.

The golden vectors (tests/golden/*.npz) were written by the reference's CPU
classes, compiled from its sources (tests/golden/make_golden.py).  The oracle
must reproduce them to round-off; the closed-form checks pin what a plausible
restatement mistake (a dropped term, a shifted index) would break.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def _snap_steps(m):
    return [k for k in range(1, m["nsteps"] + 1) if k % m["every"] == 0 or k == m["nsteps"]]


@pytest.mark.parametrize("name", ["tg_a", "tg_b", "tg_psi"])
def test_tg_matches_reference(golden, name):
    m, z = golden(name)
    dL, dLx = O.tg_ref_spacing(m["sN"], m["M"])
    coefs = O.tg_coefs_ref(m["sN"], m["M"], z["V2"], m["dv"], dL, dLx)
    s = O.TGSolver(m["sN"], m["M"], m["alpha"], m["tau"], m["sigma"], m["C1"], m["S"], *coefs,
                   z["C0"], func=m["func"], fpow=m["fpow"], int_err=m["int_err"])
    t = 0
    for q, k in enumerate(_snap_steps(m)):
        s.step(k - t)
        t = k
        ref = z["snaps"][q]
        assert np.max(np.abs(s.state() - ref)) <= 1e-13 * np.max(np.abs(ref))
    kb = O.tg_kb_table(m["func"], m["fpow"], m["alpha"], m["tau"], m["nsteps"], m["int_err"])
    np.testing.assert_allclose(kb, z["kb"], rtol=1e-14, atol=0)


@pytest.mark.parametrize("name", ["tg_dt2_a", "tg_dt2_b"])
def test_tg_restricted_sum_matches_reference(golden, name):
    # dt_alg 2: the newest-first history sum stops at the first |kb| < dt_eps
    # (about 20 of 40-50 lags kept; the full history differs by 8-15%)
    m, z = golden(name)
    dL, dLx = O.tg_ref_spacing(m["sN"], m["M"])
    coefs = O.tg_coefs_ref(m["sN"], m["M"], z["V2"], m["dv"], dL, dLx)
    s = O.TGSolver(m["sN"], m["M"], m["alpha"], m["tau"], m["sigma"], m["C1"], m["S"], *coefs,
                   z["C0"], int_err=m["int_err"], dt2=m["dt_eps"])
    t = 0
    for q, k in enumerate(_snap_steps(m)):
        s.step(k - t)
        t = k
        ref = z["snaps"][q]
        assert np.max(np.abs(s.state() - ref)) <= 1e-14 * np.max(np.abs(ref))
    # the reference forms kb() newest first and stops where all eight
    # variants fall below dt_eps: the leading values agree, the rest are 0
    n = m["nsteps"]
    kb = O.tg_kb_table(m["func"], m["fpow"], m["alpha"], m["tau"], n, m["int_err"])
    kept = z["kb"][4] != 0
    np.testing.assert_allclose(kb[:, kept], z["kb"][:, kept], rtol=1e-14, atol=0)
    assert 0 < kept.sum() < n and np.all(kb[4, ~kept] < m["dt_eps"])


@pytest.mark.parametrize("name", ["tg_dt3_a", "tg_dt3_b"])
def test_tg_binomial_series_matches_reference(golden, name):
    # dt_alg 3 (Htmp1/Htmp2 moments, int_alg 6) with 25 and 10 terms; the
    # full-history trajectory differs from these by 1e-3..2e-2, and taking
    # every coefficient list at full length (ignoring where _kb_row2 stops)
    # by ~5e-8 in tg_dt3_a
    m, z = golden(name)
    dL, dLx = O.tg_ref_spacing(m["sN"], m["M"])
    coefs = O.tg_coefs_ref(m["sN"], m["M"], z["V2"], m["dv"], dL, dLx)
    s = O.TGSolver(m["sN"], m["M"], m["alpha"], m["tau"], m["sigma"], m["C1"], m["S"], *coefs,
                   z["C0"], func=m["func"], fpow=m["fpow"], int_err=m["int_err"], dt3=m["dt_eps"])
    t = 0
    for q, k in enumerate(_snap_steps(m)):
        s.step(k - t)
        t = k
        ref = z["snaps"][q]
        assert np.max(np.abs(s.state() - ref)) <= 1e-14 * np.max(np.abs(ref))
    # kb() of dt_alg 3 forms only the two newest full steps' integrals; the
    # step uses those at s = n (int_alg 6 takes _kb_row there, as int_alg 1;
    # at s = n - 1 it may take the truncated _kb_row2, which no step reads)
    n = m["nsteps"]
    kb = O.tg_kb_table(m["func"], m["fpow"], m["alpha"], m["tau"], n, m["int_err"])
    np.testing.assert_allclose(kb[:, n], z["kb"][:, n], rtol=1e-14, atol=0)
    assert not np.any(z["kb"][:, :n - 1])


@pytest.mark.parametrize("name,psi", [("mt_flux", 0), ("mt_dir_cv2", 0), ("psi_flux", 1),
                                      ("psi_quad_flux", 1)])
def test_mt_matches_reference(golden, name, psi):
    m, z = golden(name)
    s = O.MTSolver(m["N"], m["L"], m["alpha"], m["tau"], z["U0"], psi=psi,
                   beta=m["beta"] or 1.0, form=O.FORM_CV2, k=m["k"], Ev=m["Ev"],
                   bottom_cond=m["bottom_cond"], Hbottom=m["Hbottom"])
    t = 0
    for q, k in enumerate(_snap_steps(m)):
        s.step(k - t)
        t = k
        ref = z["snaps"][q]
        assert np.max(np.abs(s.state() - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("name,psi", [("mt_hist", 0), ("psi_hist", 1)])
def test_mt_history_matches_reference(golden, name, psi):
    m, z = golden(name)
    out = O.mt_history(m["N"], m["alpha"], m["tau"], z["snaps"], m["tstep"], psi=psi,
                       beta=m["beta"] or 1.0)
    np.testing.assert_allclose(out, z["Om"], rtol=1e-13, atol=1e-13 * np.max(np.abs(z["Om"])))


def test_psi_weights_match_reference(golden):
    m, z = golden("psi_w")
    for q, (a, b, tau) in enumerate(m["cases"]):
        w, off = z[f"w{q}"], 0
        for A in m["As"]:
            np.testing.assert_allclose(O.psi_weights(a, b, tau, A), w[off:off + A], rtol=1e-14)
            off += A


def test_psi_weights_with_quadrature_fallback_match_reference(golden):
    # g = t^80 with tau 1e-2 / 1e-4: some series are not finite and calc_int
    # takes _kb's adaptive quadrature (mt_psi_caputo.cpp:736-737)
    m, z = golden("psi_w_quad")
    for q, (a, b, tau) in enumerate(m["cases"]):
        w, off = z[f"w{q}"], 0
        for A in m["As"]:
            np.testing.assert_allclose(O.psi_weights(a, b, tau, A), w[off:off + A], rtol=1e-14)
            off += A


def test_psi_kb_quad_matches_reference(golden):
    m, z = golden("psi_kb_quad")
    got = np.array([O.psi_kb_quad(*r[:5], int(r[5])) for r in z["cases"]])
    np.testing.assert_array_equal(got, z["kb"])


def test_psi_kb_quad_closed_form():
    # beta = 1: the integral of (gtj - t)^-alpha is
    # ((gtj - t0)^(1-alpha) - (gtj - t1)^(1-alpha)) / (1 - alpha).  The
    # reference's rule stops where the 1- and 2-panel midpoint values differ
    # by 1e-5, so the error is that size per leaf, larger on the interval
    # ending at the singularity (sing) at high alpha.
    for alpha in (0.05, 0.3, 0.6, 0.9):
        for t0, t1, tj, sing in ((0.0, 0.01, 0.02, 0), (0.2, 0.25, 0.9, 0), (0.3, 0.4, 0.4, 1),
                                 (0.0, 1.0, 1.0, 1), (1.5, 2.0, 2.0, 1)):
            ex = ((tj - t0) ** (1 - alpha) - (tj - t1) ** (1 - alpha)) / (1 - alpha)
            got = O.psi_kb_quad(alpha, 1.0, t0, t1, tj, sing)
            bound = 3e-2 if (sing and alpha > 0.8) else 3e-4
            assert abs(got - ex) <= bound * ex, (alpha, t0, t1, sing, got, ex)


def test_gamma_against_math():
    # Cody's rational form is good to a few ulp below 12; above 12 the A&S
    # series is exponentiated, which costs ~1e-14 relative by x = 40.
    for x in [0.05, 0.3, 0.5, 0.999, 1.0, 1.2, 1.5, 2.0, 3.7, 11.9]:
        assert O.gamma(x) == pytest.approx(math.gamma(x), rel=2e-15)
    for x in [12.5, 40.0]:
        assert O.gamma(x) == pytest.approx(math.gamma(x), rel=5e-14)


def test_tg_kb_closed_form_caputo():
    # g = t: kb(s,ja,sa,l) = Da tau^(1-a)/(1-a) [(n+ja-s-sa)^(1-a) - (n+ja-s-sa-l)^(1-a)]
    alpha, tau, n = 0.7, 1e-3, 37
    Da = 1.0 / math.gamma(1 - alpha)
    for s in [0, 5, 36, 37]:
        for ja, sa, l in [(0.5, 0, 1), (1, 0, 1), (0.5, 0, 0.5), (1, 0.5, 0.5), (1, 0, 0.5)]:
            d = n + ja - s - sa
            want = 0.0 if d - l < 0 else Da * tau ** (1 - alpha) / (1 - alpha) * (
                d ** (1 - alpha) - (d - l) ** (1 - alpha))
            got = O.tg_kb(0, 1.0, alpha, tau, n, s, ja, sa, l)
            assert got == pytest.approx(want, rel=1e-11, abs=1e-300)


def test_psi_weights_identity_kernel_closed_form():
    # beta = 1: calc_int(a, t, t+1, A) = tau^(1-a)/(1-a) [(A-t)^(1-a) - (A-t-1)^(1-a)]
    alpha, tau, A = 0.6, 1e-2, 40
    w = O.psi_weights(alpha, 1.0, tau, A)
    t = np.arange(A)
    want = tau ** (1 - alpha) / (1 - alpha) * ((A - t) ** (1 - alpha) - (A - t - 1.0) ** (1 - alpha))
    np.testing.assert_allclose(w, want, rtol=1e-11)


def test_psi_weights_quadrature_beta():
    # beta != 1 against a fine midpoint-free Gauss-Legendre quadrature after
    # the substitution that removes the endpoint singularity; the reference's
    # series stops at absolute term size 1e-5, so agreement is to ~1e-4 rel.
    alpha, beta, tau, A = 0.6, 1.4, 1e-2, 30
    w = O.psi_weights(alpha, beta, tau, A)
    G = (A * tau) ** beta
    xg, wg = np.polynomial.legendre.leggauss(200)
    for t in [0, 3, 10, 20, 29]:
        a0, a1 = t * tau, (t + 1) * tau
        # u = (G - s^beta)^(1-alpha) is smooth; integrate in s on [a0, a1]
        if t < A - 1:
            s = 0.5 * (a1 - a0) * xg + 0.5 * (a1 + a0)
            q = 0.5 * (a1 - a0) * np.sum(wg * (G - s ** beta) ** (-alpha))
        else:
            # last interval: substitute s = a1 - r^(1/(1-alpha))
            p = 1.0 / (1 - alpha)
            rmax = (a1 - a0) ** (1 - alpha)
            r = 0.5 * rmax * (xg + 1)
            s = a1 - r ** p
            q = 0.5 * rmax * np.sum(wg * p * r ** (p - 1) * (G - s ** beta) ** (-alpha))
        assert w[t] == pytest.approx(q, rel=2e-3)


def test_mt_kappa_weights_telescope():
    # with every increment equal to 1 the L1 history sum telescopes:
    # sum_{t=1}^{os-2} kappa(os-t) = (os-1)^(1-a) - 1
    N, alpha, tau, os_ = 20, 0.7, 0.01, 15
    snaps = np.arange(os_, dtype=float)[:, None] * np.ones((1, N + 2))
    out = O.mt_history(N, alpha, tau, snaps, os_ + 1)
    Da = tau ** (-alpha) / math.gamma(2 - alpha)
    np.testing.assert_allclose(out, Da * ((os_ - 1) ** (1 - alpha) - 1.0), rtol=1e-13)


# ---------------------------------------------------------------- AB path
def test_ab_weights_history_solve_match_reference(golden):
    m, z = golden("mf")
    for q, (a, tau, maxl, maxk, eps, D) in enumerate(m["weights"]):
        np.testing.assert_allclose(O.ab_weights(a, tau, maxl, maxk, eps, D), z[f"w{q}"],
                                   rtol=1e-14)
    for q, (a, tau, maxl, maxk, eps, nmax) in enumerate(m["hist"]):
        got = O.ab_history(a, tau, maxl, maxk, eps, z[f"fvs{q}"])
        np.testing.assert_allclose(got, z[f"h{q}"], rtol=1e-13,
                                   atol=1e-13 * np.max(np.abs(z[f"h{q}"])))
    for q, (a, T, n, mm, simple, maxl, maxk, eps, F) in enumerate(m["solve"]):
        err, _ = O.ab_solve(a, T, n, mm, maxl, maxk, eps, F)
        assert err == pytest.approx(z["solve_err"][q], rel=1e-12)


def test_ab_series_history_matches_reference(golden):
    # MF_intF_choice on one node's accumulators as solve_and_check (simple = 0)
    # calls it: series only, series then integral form, and maxl = 1000
    m, z = golden("mf_series")
    for q, (a, tau, maxl, maxk, eps, nmax) in enumerate(m["hist"]):
        got, choice, recounts = O.ab_series_history(a, tau, maxl, maxk, eps, z[f"fvs{q}"])
        np.testing.assert_array_equal(choice, z[f"choice{q}"])
        np.testing.assert_allclose(got, z[f"h{q}"], rtol=1e-15, atol=0)
        assert recounts == 0
    assert set(np.concatenate([z[f"choice{q}"] for q in range(len(m["hist"]))])) == {1.0, 2.0}


def test_ab_series_recount_matches_reference(golden):
    # T = 4 and 600 terms: (t tau)^(l+1) overflows, and MF_intF restarts its
    # sum from the sorted pyramid of the finite terms (MF_int_series.cpp:600-609)
    m, z = golden("mf_recount")
    for q, (a, tau, maxl, maxk, eps, nmax) in enumerate(m["hist"]):
        got, choice, recounts = O.ab_series_history(a, tau, maxl, maxk, eps, z[f"fvs{q}"])
        np.testing.assert_array_equal(choice, z[f"choice{q}"])
        np.testing.assert_array_equal(got, z[f"h{q}"])
        assert recounts > 0
    for q, (a, T, n, mm, simple, maxl, maxk, eps, F) in enumerate(m["solve"]):
        err, _ = O.ab_solve_series(a, T, n, mm, maxl, maxk, eps, F)
        assert err == z["solve_err"][q]


@pytest.mark.parametrize("q", [1, 3, 4])
def test_ab_series_solve_matches_reference(golden, q):
    # the reference's own configurations (convection_diffusion_test_time:
    # 6000 steps, 100 intervals, maxl 50; convection_diffusion_test: maxl 1000)
    m, z = golden("mf_series")
    a, T, n, mm, simple, maxl, maxk, eps, F = m["solve"][q]
    err, _ = O.ab_solve_series(a, T, n, mm, maxl, maxk, eps, F)
    assert err == pytest.approx(z["solve_err"][q], rel=1e-14)


def test_ab_series_first_step_is_the_kernel_integral():
    # at i = 2 the history is one interval [0, tau] at tj = 2 tau: the series
    # sum_l (tau^(l+1) - 0) Kc(l) F equals F I(2), I(2) the same series as
    # MF_int_choice forms it for the full-history form
    for a, tau in [(0.3, 0.02), (0.6, 0.01), (0.5, 1e-3)]:
        F = 0.731
        got, _, _ = O.ab_series_history(a, tau, 50, 1000, 1e-10, np.array([F]))
        full = O.ab_history(a, tau, 50, 1000, 1e-10, np.array([F, 0.0]))
        assert got[0] == pytest.approx(full[0], rel=1e-13)


def test_mittag_leffler_matches_reference(golden):
    _, z = golden("mf")
    got = np.array([O.mittag_leffler(*row) for row in z["ml_in"]])
    np.testing.assert_allclose(got, z["ml_out"], rtol=1e-14, atol=1e-300)


def test_mittag_leffler_special_cases():
    from math import cos, cosh, erfc, exp, sqrt
    for zz in [0.3, 0.9, 2.0, -0.5, -2.0, -4.0]:
        assert O.mittag_leffler(1.0, 1.0, zz, 1e-15) == pytest.approx(exp(zz), rel=1e-12)
        want = cosh(sqrt(zz)) if zz > 0 else cos(sqrt(-zz))
        assert O.mittag_leffler(2.0, 1.0, zz, 1e-15) == pytest.approx(want, rel=1e-11)
    for zz in [0.2, 0.7, -0.3, -1.0, -2.5]:
        want = exp(zz * zz) * erfc(-zz)
        assert O.mittag_leffler(0.5, 1.0, zz, 1e-15) == pytest.approx(want, rel=1e-8)


def test_ab_weights_against_quadrature():
    # I(d) = integral over [(d-1) tau, d tau] of E_a(-a/(1-a) u^a) du; the
    # integrand is bounded, so a substitution u = (d-1)tau + tau v^2 and
    # Gauss-Legendre in v resolve the u^a behaviour at u = 0.  The
    # reference's methods truncate: the l-series at d = 1 (tau/tj = 1, its
    # terms do not decay) and the maxl-node quadrature of the integral form
    # (d = 30 here) are good to ~2e-4; the other lags to ~1e-8.
    a, tau = 0.6, 0.02
    lam = a / (1 - a)
    w = O.ab_weights(a, tau, 60, 1000, 1e-12, 30)
    xg, wg = np.polynomial.legendre.leggauss(200)
    v = 0.5 * (xg + 1)
    for d, rel in [(1, 5e-4), (2, 1e-7), (5, 1e-7), (12, 1e-7), (30, 5e-4)]:
        u = (d - 1) * tau + tau * v * v
        f = np.array([O.mittag_leffler(a, 1.0, -lam * ui ** a, 1e-15) for ui in u])
        q = 0.5 * np.sum(wg * f * 2 * tau * v)
        assert w[d - 1] == pytest.approx(q, rel=rel), d


@pytest.mark.parametrize("name", ["tg_dt4_a", "tg_dt4_b"])
def test_tg_ia_ibk_approximation_matches_reference(golden, name):
    # dt_alg 4: per node and direction the "long" (ia) and "short" (ibk)
    # series approximations of the history sum; the trajectories differ from
    # the full history by 3e-7 (alpha 0.7) to 7e-5 (alpha 0.4), so
    # agreement to 1e-14 pins the approximation structure itself
    m, z = golden(name)
    dL, dLx = O.tg_ref_spacing(m["sN"], m["M"])
    coefs = O.tg_coefs_ref(m["sN"], m["M"], z["V2"], m["dv"], dL, dLx)
    s = O.TGSolver(m["sN"], m["M"], m["alpha"], m["tau"], m["sigma"], m["C1"], m["S"], *coefs,
                   z["C0"], int_err=m["int_err"], dt4=True)
    full = O.TGSolver(m["sN"], m["M"], m["alpha"], m["tau"], m["sigma"], m["C1"], m["S"], *coefs,
                      z["C0"], int_err=m["int_err"])
    t = 0
    for q, k in enumerate(_snap_steps(m)):
        s.step(k - t)
        full.step(k - t)
        t = k
        ref = z["snaps"][q]
        assert np.max(np.abs(s.state() - ref)) <= 1e-14 * np.max(np.abs(ref))
        d = np.max(np.abs(full.state() - ref)) / np.max(np.abs(ref))
        assert 1e-9 < d < 1e-3, d
