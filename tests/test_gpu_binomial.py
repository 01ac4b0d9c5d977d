"""The reference's single-series (binomial) history of the 2D LOD solver
(TG_solver with dt_alg = 3, fr_d_t_g_2d.cpp:3385-3428) on the GPU
(FRM_PREC_BINOMIAL) against the reference's own trajectories (the tg_dt3
goldens) and the oracle.  This form is the reference's method, not an
approximation of this build, so it is held to the plain FP64 bar: 1e-12
norm-wise relative, max|u - u_ref| / max|u_ref|."""
import numpy as np
import pytest

from oracle import oracle as O
from theses_code_amd import _native as N
from theses_code_amd import solvers as S

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(0)
    yield c
    c.close()


def relerr(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("name", ["tg_dt3_a", "tg_dt3_b"])
def test_binomial_matches_reference_golden(ctx, golden, name):
    m, z = golden(name)
    print(f"{name}: {m['sN']}x{m['M']} alpha {m['alpha']} terms {m['dt_eps']}", flush=True)
    dL, dLx = S.tg_ref_spacing(m["sN"], m["M"])
    coefs = S.coefs_ref(z["V2"], m["dv"], dL, dLx)
    s = S.TGSolver(ctx, m["sN"], m["M"], m["alpha"], m["tau"], m["sigma"], coefs, z["C0"],
                   capacity=m["nsteps"], C1=m["C1"], S=m["S"], precision=N.PREC_BINOMIAL,
                   series_terms=int(m["dt_eps"]), series_tol=m["int_err"])
    ks = [k for k in range(1, m["nsteps"] + 1) if k % m["every"] == 0 or k == m["nsteps"]]
    t = 0
    for q, k in enumerate(ks):
        s.calc_step(k - t)
        t = k
        assert relerr(s.b_C, z["snaps"][q]) <= TOL, (name, k)
    s.close()


# Larger grids and runs against the oracle's restatement (which emulates the
# reference's _kb_row2 calls and their coefficient cache), with term counts
# from 1 to 60 and tolerances where _kb_row2 stops early.
@pytest.mark.parametrize("sN,M,alpha,steps,terms,tol", [
    (40, 33, 0.6, 80, 25, 1e-6), (97, 64, 0.85, 60, 8, 1e-4), (33, 70, 0.45, 50, 60, 1e-5),
    (20, 20, 0.7, 30, 1, 1e-6)])
def test_binomial_matches_oracle(ctx, sN, M, alpha, steps, terms, tol):
    rng = np.random.default_rng(sN * M + terms)
    nn = (sN + 2) * (M + 2)
    h = 1.0 / (sN + 1)
    coefs = S.coefs_advection(rng.uniform(-0.1, 0.1, nn), rng.uniform(-0.1, 0.1, nn), 1e-3, h)
    C0 = rng.uniform(0, 1, nn)
    tau, sigma, C1, Ssrc = 1e-3, 1.0, 0.0, 0.05
    o = O.TGSolver(sN, M, alpha, tau, sigma, C1, Ssrc, *coefs, C0, int_err=tol, dt3=terms)
    g = S.TGSolver(ctx, sN, M, alpha, tau, sigma, coefs, C0, capacity=steps, C1=C1, S=Ssrc,
                   precision=N.PREC_BINOMIAL, series_terms=terms, series_tol=tol)
    for chunk in (1, steps // 2 - 1, steps - steps // 2):
        o.step(chunk)
        g.calc_step(chunk)
        assert relerr(g.b_C, o.state().reshape(M + 2, sN + 2)) <= TOL
    g.close()


def test_binomial_restart_clears_moments(ctx):
    rng = np.random.default_rng(5)
    sN = M = 24
    nn = (sN + 2) * (M + 2)
    coefs = S.coefs_advection(rng.uniform(-0.1, 0.1, nn), rng.uniform(-0.1, 0.1, nn), 1e-3,
                              1.0 / (sN + 1))
    C0 = rng.uniform(0, 1, nn)
    g = S.TGSolver(ctx, sN, M, 0.7, 1e-3, 1.0, coefs, C0, capacity=40, precision=N.PREC_BINOMIAL)
    g.calc_step(40)
    first = g.b_C.copy()
    g.set_state(C0)
    g.calc_step(40)
    assert np.array_equal(g.b_C, first)
    g.close()


def test_binomial_rejects_ab_kernel(ctx):
    nn = 12 * 12
    z = np.zeros(nn)
    with pytest.raises(N.FrmError):
        S.TGSolver(ctx, 10, 10, 0.5, 1e-3, 1.0, (z, z, z, z), z, capacity=4,
                   precision=N.PREC_BINOMIAL, kernel=N.KERNEL_AB)
