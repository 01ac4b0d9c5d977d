"""Batched 1D Atangana-Baleanu solve (solve_and_check of MF_int_series.cpp)
on the GPU against the reference's own outputs (the `mf` and `mf_series`
goldens: its error measure for the three manufactured solutions, simple = 1
full history and simple = 0 series history) and the oracle's final fields.  Plain FP64: 1e-12 norm-wise relative on the field; the error
measure sum (C - CC)^2 is a difference-of-nearly-equal quantity (C - CC is
~1e-2..1e-4 of C), so it is held to 1e-9 relative."""
import numpy as np
import pytest

from oracle import oracle as O
from theses_code_amd import _native as N
from theses_code_amd import solvers as S

pytestmark = pytest.mark.gpu
TOL = 1e-12
TOL_ERR = 1e-9


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(0)
    yield c
    c.close()


def relerr(a, b):
    return np.max(np.abs(np.asarray(a) - b)) / max(np.max(np.abs(b)), 1e-300)


def _batch(ctx, probs, tau, steps):
    U0 = [np.zeros(p["N"] + 2) for p in probs]
    return S.MT1DBatch(ctx, probs, tau, steps, U0, kind=N.MT_AB)


def test_ab1d_matches_reference_golden(ctx, golden):
    m, z = golden("mf")
    for q, (a, T, n, mm, simple, maxl, maxk, eps, F) in enumerate(m["solve"]):
        prob = dict(N=mm, alpha=a, src=F, ab_maxl=maxl, ab_maxkiter=maxk, ab_eps=eps)
        b = _batch(ctx, [prob], T / n, n)
        b.calc_step(n)
        err = b.ab_error()[0]
        C = b.state()[0][:mm + 1]
        _, Cref = O.ab_solve(a, T, n, mm, maxl, maxk, eps, F)
        b.close()
        print(f"case {q}: err {err:.6e} golden {z['solve_err'][q]:.6e} field {relerr(C, Cref):.2e}")
        assert err == pytest.approx(z["solve_err"][q], rel=TOL_ERR)
        assert relerr(C, Cref) <= TOL


# An uneven batch sharing tau: orders 0.35..0.9, all three manufactured
# solutions, 12 to 1024 unknowns (one to sixteen per lane), each problem
# against its own oracle run.
def test_ab1d_uneven_batch_matches_oracle(ctx):
    tau, steps = 1.0 / 40, 40
    spec = [(0.35, 12, 0), (0.6, 64, 1), (0.85, 200, 2), (0.5, 1025, 0), (0.72, 129, 1),
            (0.45, 513, 2), (0.9, 3, 0)]
    probs = [dict(N=mm, alpha=a, src=F, ab_maxl=50, ab_maxkiter=1000, ab_eps=1e-10)
             for a, mm, F in spec]
    b = _batch(ctx, probs, tau, steps)
    b.calc_step(17)
    b.calc_step(steps - 17)
    errs = b.ab_error()
    states = b.state()
    b.close()
    for i, (a, mm, F) in enumerate(spec):
        e_ref, Cref = O.ab_solve(a, steps * tau, steps, mm, 50, 1000, 1e-10, F)
        assert relerr(states[i][:mm + 1], Cref) <= TOL, (i, relerr(states[i][:mm + 1], Cref))
        assert errs[i] == pytest.approx(e_ref, rel=TOL_ERR), i


# FP64-accurate kernel integrals instead of MF_int_choice: the same scheme
# with weights that are not truncated; its error against the manufactured
# solution stays at the discretisation level of the MF_int_choice run.
def test_ab1d_exact_weights_run(ctx):
    tau, steps = 1.0 / 60, 60
    base = dict(N=20, alpha=0.5, src=0, ab_maxl=50, ab_maxkiter=1000, ab_eps=1e-10)
    b1 = _batch(ctx, [base], tau, steps)
    b2 = _batch(ctx, [dict(base, ab_exact=1)], tau, steps)
    b1.calc_step(steps)
    b2.calc_step(steps)
    e1, e2 = b1.ab_error()[0], b2.ab_error()[0]
    b1.close()
    b2.close()
    assert np.isfinite(e2) and 0.2 < e2 / e1 < 5.0, (e1, e2)


def test_ab1d_restart_and_capacity(ctx):
    prob = dict(N=16, alpha=0.7, src=1, ab_maxl=50)
    b = _batch(ctx, [prob], 1.0 / 40, 40)
    b.calc_step(40)
    first = b.state()[0].copy()
    e1 = b.ab_error()[0]
    with pytest.raises(N.FrmError):
        b.calc_step(1)
    b.set_state([np.zeros(18)])
    b.calc_step(40)
    assert np.array_equal(b.state()[0], first)
    assert b.ab_error()[0] == e1
    b.close()


# ---------------------------------------------------------------- series history (simple = 0)
def _series_prob(a, mm, maxl, F, maxk=1000, eps=1e-10):
    return dict(N=mm, alpha=a, src=F, ab_maxl=maxl, ab_maxkiter=maxk, ab_eps=eps, ab_series=1)


@pytest.mark.parametrize("q", range(5))
def test_ab1d_series_matches_reference_golden(ctx, golden, q):
    # the reference's own configurations: convection_diffusion_test_time
    # (6000 steps, 100 intervals, maxl 50, orders 0.9 / 0.5 / 0.1) and
    # convection_diffusion_test (maxl 1000), plus F_func 2
    m, z = golden("mf_series")
    a, T, n, mm, simple, maxl, maxk, eps, F = m["solve"][q]
    print(f"series case {q}: a={a} n={n} m={mm} maxl={maxl} F={F}", flush=True)
    b = _batch(ctx, [_series_prob(a, mm, maxl, F, maxk, eps)], T / n, n)
    b.calc_step(n)
    err = b.ab_error()[0]
    C = b.state()[0][:mm + 1]
    b.close()
    _, Cref = O.ab_solve_series(a, T, n, mm, maxl, maxk, eps, F)
    print(f"  err {err:.6e} golden {z['solve_err'][q]:.6e} field {relerr(C, Cref):.2e}")
    assert relerr(C, Cref) <= TOL
    assert err == pytest.approx(z["solve_err"][q], rel=TOL_ERR)


# Series and full-history problems in one batch (the full-history sums skip
# the series problems' nodes), orders on both sides of the series/integral
# switch, against the oracle problem by problem.
def test_ab1d_series_mixed_batch_matches_oracle(ctx):
    tau, steps = 1.0 / 50, 50
    spec = [(0.35, 30, 0, 1), (0.6, 64, 1, 1), (0.85, 129, 2, 0), (0.9, 40, 0, 1),
            (0.5, 200, 2, 1), (0.72, 17, 1, 0)]
    probs = [dict(N=mm, alpha=a, src=F, ab_maxl=50, ab_maxkiter=1000, ab_eps=1e-10, ab_series=sr)
             for a, mm, F, sr in spec]
    b = _batch(ctx, probs, tau, steps)
    b.calc_step(23)
    b.calc_step(steps - 23)
    errs, states = b.ab_error(), b.state()
    b.close()
    for i, (a, mm, F, sr) in enumerate(spec):
        solve = O.ab_solve_series if sr else O.ab_solve
        e_ref, Cref = solve(a, steps * tau, steps, mm, 50, 1000, 1e-10, F)
        assert relerr(states[i][:mm + 1], Cref) <= TOL, (i, relerr(states[i][:mm + 1], Cref))
        assert errs[i] == pytest.approx(e_ref, rel=TOL_ERR), i


@pytest.mark.parametrize("graph", [False, True])
def test_ab1d_series_recount_matches_reference(ctx, golden, graph):
    # tau 0.05 to T = 3 and 4 with 600-700 terms: the power-series
    # accumulators overflow and MF_intF's sum restarts from the sorted pyramid
    # of the finite terms (MF_int_series.cpp:600-609) at the nodes and steps
    # after that; one problem of the batch (maxl 50) never overflows.  The
    # 60-step case is read at step 60, the others at step 80.
    m, z = golden("mf_recount")
    cases = m["solve"]
    tau = cases[0][1] / cases[0][2]
    probs = [_series_prob(a, mm, maxl, F, maxk, eps)
             for a, T, n, mm, simple, maxl, maxk, eps, F in cases]
    probs.append(_series_prob(0.3, 25, 50, 1))
    nmax = max(c[2] for c in cases)
    b = S.MT1DBatch(ctx, probs, tau, nmax, [np.zeros(p["N"] + 2) for p in probs], kind=N.MT_AB,
                    graph=graph)
    done = 0
    for n in sorted({c[2] for c in cases}):
        b.calc_step((n - done) // 2)
        b.calc_step(n - done - (n - done) // 2)
        done = n
        errs, states = b.ab_error(), b.state()
        for i, (a, T, nn, mm, simple, maxl, maxk, eps, F) in enumerate(cases):
            if nn != n:
                continue
            print(f"recount case {i}: a={a} T={T} n={nn} maxl={maxl} F={F}", flush=True)
            _, Cref = O.ab_solve_series(a, T, nn, mm, maxl, maxk, eps, F)
            assert relerr(states[i][:mm + 1], Cref) <= TOL, i
            assert errs[i] == pytest.approx(z["solve_err"][i], rel=TOL_ERR), i
    b.close()
    _, Cref = O.ab_solve_series(0.3, tau * nmax, nmax, 25, 50, 1000, 1e-10, 1)
    assert relerr(states[-1][:26], Cref) <= TOL


def test_ab1d_series_restart_clears_accumulators(ctx):
    prob = _series_prob(0.8, 20, 50, 2)
    b = _batch(ctx, [prob], 1.0 / 30, 30)
    b.calc_step(30)
    first, e1 = b.state()[0].copy(), b.ab_error()[0]
    b.set_state([np.zeros(22)])
    b.calc_step(30)
    assert np.array_equal(b.state()[0], first)
    assert b.ab_error()[0] == e1
    b.close()


def test_ab1d_series_rejects_exact_weights(ctx):
    with pytest.raises(N.FrmError):
        _batch(ctx, [dict(_series_prob(0.5, 10, 50, 0), ab_exact=1)], 0.1, 10)
