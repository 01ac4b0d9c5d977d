"""frm_soe_* (the standalone sum-of-exponentials history) on the GPU.

The recurrence is this build's form, not a reference algorithm, so it is
pinned by what the mathematics fixes rather than by the oracle:
  * with fixed decays, folding increments D_s into H_l = q_l H_l + D_s gives
    sum_s w_t(k - s) D_s with w_t(j) = sum_l c[t, l] q_l^j, evaluated here as
    a direct history sum over an explicit weight table;
  * with decays that change every step, H_l = sum_s prod_{r>s} q_l(r) D_s;
  * a constant increment gives the geometric sum (1 - q^K) / (1 - q);
  * weights fitted by frm_host_soe_fit (L1 Caputo weights) reproduce the full
    history sum to within the fit's own relative error.
Tolerances: FP64 states 1e-12 norm-wise relative, FP32 states 1e-6."""
import numpy as np
import pytest

from theses_code_amd import _native as N
from theses_code_amd import solvers as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(0)
    yield c
    c.close()


def relerr(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("n,L,nt,f32", [(5000, 24, 2, False), (777, 64, 1, False),
                                        (4099, 16, 2, True)])
def test_fixed_decays_equal_direct_history_sum(ctx, n, L, nt, f32):
    print(f"soe n {n} L {L} targets {nt} f32 {f32}", flush=True)
    rng = np.random.default_rng(n + L)
    steps = 40
    q = rng.uniform(0.3, 0.999, L)
    c = rng.uniform(0.0, 1.0, (nt, L))
    D = rng.uniform(0.0, 1.0, (steps, n))  # positive: no cancellation in the bound
    w = c @ (q[:, None] ** np.arange(steps)[None, :])  # w[t, j]
    s = S.SoeHistory(ctx, n, L, nt, f32=f32)
    scale = 0.75
    for k in range(steps):
        got = s.step(D[k], q, c, scale=scale)
        if k % 9 == 0 or k == steps - 1:
            ref = scale * (w[:, k::-1] @ D[:k + 1])  # lag k - s for step s
            err = relerr(got, ref)
            assert err <= (1e-6 if f32 else 1e-12), (k, err)
    s.close()


def test_varying_decays_products(ctx):
    rng = np.random.default_rng(7)
    n, L, steps = 300, 8, 20
    Q = rng.uniform(0.2, 1.0, (steps, L))
    c = rng.uniform(-1.0, 1.0, (1, L))
    D = rng.standard_normal((steps, n))
    s = S.SoeHistory(ctx, n, L)
    for k in range(steps):
        got = s.step(D[k], Q[k], c)[0]
    # after the last step: H_l = sum_s prod_{r=s+1}^{steps-1} Q[r, l] D_s
    P = np.array([np.prod(Q[s_ + 1:], axis=0) for s_ in range(steps)])  # [s, l]
    ref = (P @ c[0]) @ D
    assert relerr(got, ref) <= 1e-12
    s.close()


@pytest.mark.parametrize("f32,K", [(False, 30), (True, 8)])
def test_constant_increment_geometric_sum_at_full_size(ctx, f32, K):
    # 4M + 7 nodes (a ragged last workgroup), 64 modes, two targets
    n, L = (1 << 22) + 7, 64
    rng = np.random.default_rng(3)
    q = np.exp(-np.logspace(-3, 1.5, L))
    c = rng.uniform(0.1, 1.0, (2, L))
    s = S.SoeHistory(ctx, n, L, 2, f32=f32)
    one = N.DeviceArray.from_host(ctx, np.ones(n))
    for _ in range(K):
        out = s.step_device(one, q, c)
    got = out.to_host().reshape(2, n)
    ref = c @ ((1.0 - q ** K) / (1.0 - q))
    err = max(relerr(got[t], np.full(n, ref[t])) for t in range(2))
    assert err <= (1e-6 if f32 else 1e-12), err
    one.free()
    s.close()


def test_fitted_caputo_weights_reproduce_full_history(ctx):
    # L1 weights kappa(m) = m^(1-a) - (m-1)^(1-a) for lags 1..K fitted by the
    # host SOE fit; the newest increment has lag 1
    a, K, n, L = 0.6, 300, 2000, 32
    m = np.arange(1, K + 1, dtype=np.float64)
    kappa = m ** (1 - a) - (m - 1) ** (1 - a)
    q, c, max_rel = N.host_soe_fit(kappa, 1, L, 1e-3)
    assert max_rel < 1e-7
    rng = np.random.default_rng(11)
    D = rng.uniform(0.0, 1.0, (K, n))
    s = S.SoeHistory(ctx, n, L)
    for k in range(K):
        got = s.step(D[k], q, c)[0]
    ref = kappa[::-1] @ D  # step s has lag K - s
    err = relerr(got, ref)
    print(f"fit max rel {max_rel:.2e}, history sum rel err {err:.2e}")
    assert err <= max_rel * 1.01 + 1e-13
    s.close()


def test_evaluate_without_fold_accumulate_and_clear(ctx):
    rng = np.random.default_rng(5)
    n, L = 1000, 12
    q = rng.uniform(0.5, 0.99, L)
    c = rng.uniform(0, 1, (1, L))
    c2 = rng.uniform(0, 1, (1, L))
    D = rng.uniform(0, 1, (6, n))
    s = S.SoeHistory(ctx, n, L)
    for k in range(5):
        s.step(D[k], q, c)
    e1 = s.step(None, c=c2)[0]
    e2 = s.step(None, c=c2)[0]  # no fold: the states did not move
    assert np.array_equal(e1, e2)
    acc = s.step_device(None, c=c2, scale=0.5, accumulate=True).to_host()
    assert relerr(acc, 1.5 * e1) <= 1e-15
    last = s.step(D[5], q, c)[0]
    s.clear()
    fresh = S.SoeHistory(ctx, n, L)
    for k in range(6):
        r1 = s.step(D[k], q, c)[0]
        r2 = fresh.step(D[k], q, c)[0]
        assert np.array_equal(r1, r2)
    assert np.array_equal(r1, last)
    s.close()
    fresh.close()


def test_arguments(ctx):
    for bad in (dict(nterms=0), dict(nterms=65), dict(ntargets=3)):
        kw = dict(nterms=4, ntargets=1)
        kw.update(bad)
        with pytest.raises(N.FrmError):
            S.SoeHistory(ctx, 100, kw["nterms"], kw["ntargets"])
    s = S.SoeHistory(ctx, 100, 4)
    with pytest.raises(N.FrmError):
        s.step(np.ones(100), None, np.ones(4))  # a fold needs q
    s.close()
