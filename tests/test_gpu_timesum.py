"""The Sum family of irrigation_fractional.cpp (H_solver::time_summing,
4099-4170; the OpenCL/CUDA Sum kernel and fr_ocl_call_sum, 1254-1320,
1416-1489, 1828-1905) on the GPU through frm_timesum_*, against the oracle's
restatement (tests/test_oracle_timesum.py pins it by closed forms; the
reference file needs libfftpack, which the reference tree lacks, so it has no
golden vectors).

Tolerances: sum_alg 0 and 1 in FP64 sum in another order than the
reference's newest-first loop (time chunks), so they are held to 1e-12
norm-wise relative, max|S - S_ref| / max|S_ref|; sum_alg 2 keeps the
reference's operation order and is bit-exact; the compensated FP32 form
(FRM_PREC_F32C) to 1e-6."""
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


def levels(rng, nlev, n):
    """A smooth-in-time field with noise, like successive solver states."""
    base = rng.uniform(-1, 1, n)
    rate = rng.uniform(-1, 1, n)
    t = np.arange(nlev)[:, None]
    return base[None, :] + rate[None, :] * np.sqrt(t + 1.0) + 1e-3 * rng.standard_normal((nlev, n))


def caputo_kb(T, T_end, a):
    return ((T_end - T[:-1]) ** (1 - a) - (T_end - T[1:]) ** (1 - a)) / (1 - a)


@pytest.mark.parametrize("sum_alg,n,nlev,eps", [
    (0, 1000, 40, 0.0), (0, 102 ** 3, 24, 0.0), (0, 4099, 300, 0.0),
    (1, 20000, 120, 0.1), (1, 777, 60, 10.0)])
def test_sum01_matches_oracle(ctx, sum_alg, n, nlev, eps):
    print(f"sum_alg {sum_alg} n {n} levels {nlev} eps {eps}", flush=True)
    rng = np.random.default_rng(n + nlev)
    old = levels(rng, nlev, n)
    tau = rng.uniform(0.5, 1.5, nlev)
    T = np.concatenate([[0.0], np.cumsum(tau)])
    g = S.TimeSum(ctx, n, sum_alg=sum_alg, sum_param=eps, capacity=nlev)
    checked = 0
    for k in range(nlev):
        g.push(old[k])
        assert g.old_size == k + 1
        if k % 7 == 0 or k == nlev - 1:
            kb = caputo_kb(T[:k + 1], T[k + 1], 0.6)  # the weights at tstep = k
            got = g.sums(kb_3_cache=kb, tau=tau[:k])
            ref = O.timesum01(sum_alg, eps, old[:k + 1], kb, tau[:k])
            if k == 0:
                assert np.array_equal(got, np.zeros(n))
            else:
                assert relerr(got, ref) <= TOL, k
            checked += 1
    assert checked >= 2
    g.close()


def test_sum01_linear_history_is_caputo_of_linear_at_full_size(ctx):
    # size-independent property: exact interval weights on a linear history
    n, nlev, a = 102 ** 3, 50, 0.7
    rng = np.random.default_rng(1)
    slope = rng.uniform(-2, 2, n)
    tau = rng.uniform(0.5, 1.5, nlev)
    T = np.concatenate([[0.0], np.cumsum(tau)])
    g = S.TimeSum(ctx, n, sum_alg=0, capacity=nlev)
    for k in range(nlev):
        g.push(slope * T[k])
    got = g.sums(kb_3_cache=caputo_kb(T[:nlev], T[nlev], a), tau=tau[:nlev - 1])
    ref = slope * (T[nlev] ** (1 - a) - (T[nlev] - T[nlev - 1]) ** (1 - a)) / (1 - a)
    assert relerr(got, ref) <= 1e-12
    g.close()


# 50000 nodes take the 2-nodes-per-lane FP32 kernel, 102^3 + 5 (a store
# stride of 1037 x 1024) the 4-nodes-per-lane one with a half-idle last
# workgroup
@pytest.mark.parametrize("sum_alg,n,nlev", [(0, 50000, 200), (1, 50000, 200),
                                            (0, 102 ** 3 + 5, 40)])
def test_sum01_f32c_within_1e6(ctx, sum_alg, n, nlev):
    rng = np.random.default_rng(9 + sum_alg)
    old = levels(rng, nlev, n)
    tau = rng.uniform(0.5, 1.5, nlev)
    T = np.concatenate([[0.0], np.cumsum(tau)])
    eps = 0.0 if sum_alg == 0 else 0.01
    g = S.TimeSum(ctx, n, sum_alg=sum_alg, sum_param=eps, capacity=nlev,
                  precision=N.PREC_F32C)
    for k in range(nlev):
        g.push(old[k])
    kb = caputo_kb(T[:nlev], T[nlev], 0.4)
    got = g.sums(kb_3_cache=kb, tau=tau[:nlev - 1])
    ref = O.timesum01(sum_alg, eps, old, kb, tau[:nlev - 1])
    err = relerr(got, ref)
    print(f"F32C sum_alg {sum_alg}: max rel {err:.2e}, "
          f"rms {np.sqrt(np.mean((got - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)):.2e}")
    assert err <= 1e-6
    g.close()


@pytest.mark.parametrize("n,nterms,ncoefs", [(1000, 16, 16), (4097, 21, 13), (102 ** 3, 8, 8),
                                            (300, 3, 40)])
def test_sum2_bit_exact_vs_oracle(ctx, n, nterms, ncoefs):
    print(f"sum_alg 2 n {n} terms {nterms} coefs {ncoefs}", flush=True)
    rng = np.random.default_rng(n + nterms)
    nlev = 12
    old = levels(rng, nlev, n)
    tau = 0.3
    a = 0.55
    g = S.TimeSum(ctx, n, sum_alg=2, sum_param=nterms)
    o = O.TimeSum2(n, nterms)
    for k in range(nlev):
        g.push(old[k])
        if k == 0:
            assert np.array_equal(g.sums(coefs=np.ones(ncoefs), gtj=tau, gt=0.0, alpha=a,
                                         tau_n=tau), np.zeros(n))
            continue
        T1, T2 = (k + 1) * tau, (k + 2) * tau
        coefs = rng.uniform(0.1, 1.0, ncoefs) / (np.arange(ncoefs) + 1.0)
        got = g.sums(coefs=coefs, gtj=T2, gt=T1, alpha=a, tau_n=tau)
        ref = o(old[k] - old[k - 1], coefs, T2, T1, a, tau)
        assert np.array_equal(got, ref), k
        # a second evaluation before the next push does not fold again
        again = g.sums(coefs=coefs, gtj=T2, gt=T1, alpha=a, tau_n=tau)
        assert np.array_equal(again, ref), k
    g.close()


def test_clear_restarts(ctx):
    n, nlev = 3000, 20
    rng = np.random.default_rng(4)
    old = levels(rng, nlev, n)
    tau = np.ones(nlev)
    kb = rng.uniform(0.1, 1, nlev)
    for alg in (0, 2):
        g = S.TimeSum(ctx, n, sum_alg=alg, sum_param=6, capacity=nlev)
        outs = []
        for rep in range(2):
            for k in range(nlev):
                g.push(old[k])
                if alg == 2 and k > 0:
                    r = g.sums(coefs=np.full(6, 0.5), gtj=1.5, gt=1.2, alpha=0.5, tau_n=1.0)
            if alg == 0:
                r = g.sums(kb_3_cache=kb, tau=tau)
            outs.append(r)
            g.clear()
            assert g.old_size == 0
        assert np.array_equal(outs[0], outs[1])
        g.close()


def test_capacity_and_arguments(ctx):
    g = S.TimeSum(ctx, 100, sum_alg=0, capacity=3)
    for k in range(3):
        g.push(np.full(100, float(k)))
    with pytest.raises(N.FrmError):
        g.push(np.zeros(100))
    with pytest.raises(N.FrmError):
        g.sums()  # no kb_3_cache
    g.close()
    with pytest.raises(N.FrmError):
        S.TimeSum(ctx, 100, sum_alg=3)
    with pytest.raises(N.FrmError):
        S.TimeSum(ctx, 100, sum_alg=2, sum_param=0.0)
    with pytest.raises(N.FrmError):
        S.TimeSum(ctx, 100, sum_alg=2, sum_param=8, precision=N.PREC_F32C)
    with pytest.raises(N.FrmError):
        S.TimeSum(ctx, 100, sum_alg=0, capacity=4, precision=N.PREC_MFMA_BF16)
