"""Low-level HIP kernels through the C ABI against brute-force NumPy."""
import numpy as np
import pytest

from theses_code_amd import _native as N

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(0)
    yield c
    c.close()


def _dev(ctx, a):
    return N.DeviceArray.from_host(ctx, a)


@pytest.mark.parametrize("n,nsys", [(1, 3), (5, 7), (64, 9), (100, 33), (512, 12), (1000, 5),
                                    (2048, 3)])
def test_tridiag_matches_dense_solve(ctx, n, nsys):
    rng = np.random.default_rng(n * 7 + nsys)
    a = rng.uniform(-1, 1, (nsys, n))
    c = rng.uniform(-1, 1, (nsys, n))
    b = 2.5 + np.abs(a) + np.abs(c) + rng.uniform(0, 1, (nsys, n))
    d = rng.normal(0, 1, (nsys, n))
    bufs = [_dev(ctx, v) for v in (a, b, c, d)]
    x = N.DeviceArray(ctx, nsys * n)
    desc = N.TridiagDesc(nsys, n, n, 1)
    N._check(N.lib().frm_tridiag(ctx.h, desc, *[q.ptr for q in bufs], x.ptr), ctx.h)
    got = x.to_host().reshape(nsys, n)
    for s in range(nsys):
        Amat = np.diag(b[s]) + np.diag(a[s, 1:], -1) + np.diag(c[s, :-1], 1)
        want = np.linalg.solve(Amat, d[s])
        np.testing.assert_allclose(got[s], want, rtol=1e-12, atol=1e-12 * np.max(np.abs(want)))
    for q in bufs + [x]:
        q.free()


def test_tridiag_strided_layout(ctx):
    # element stride nsys, system stride 1 (interleaved layout)
    rng = np.random.default_rng(5)
    n, nsys = 300, 17
    a, c = rng.uniform(-1, 1, (2, n, nsys))
    b = 3.0 + rng.uniform(0, 1, (n, nsys))
    d = rng.normal(0, 1, (n, nsys))
    bufs = [_dev(ctx, v) for v in (a, b, c, d)]
    x = N.DeviceArray(ctx, n * nsys)
    N._check(N.lib().frm_tridiag(ctx.h, N.TridiagDesc(nsys, n, 1, nsys),
                                 *[q.ptr for q in bufs], x.ptr), ctx.h)
    got = x.to_host().reshape(n, nsys)
    for s in range(nsys):
        Amat = np.diag(b[:, s]) + np.diag(a[1:, s], -1) + np.diag(c[:-1, s], 1)
        np.testing.assert_allclose(got[:, s], np.linalg.solve(Amat, d[:, s]), rtol=1e-12,
                                   atol=1e-13)


@pytest.mark.parametrize("depth,nodes,nt", [(0, 100, 1), (1, 1000, 2), (300, 1000, 2),
                                            (777, 5000, 1), (513, 2, 2)])
def test_memterm_matches_numpy(ctx, depth, nodes, nt):
    rng = np.random.default_rng(depth + nodes)
    cap = max(depth, 1) + 3
    h = N._V()
    N._check(N.lib().frm_hist_create(ctx.h, nodes, cap, N.PREC_F64, N.ct.byref(h)), ctx.h)
    snaps = np.cumsum(rng.normal(0, 1, (depth + 1, nodes)), axis=0)
    dev = [_dev(ctx, s) for s in snaps]
    for s in range(depth):
        N._check(N.lib().frm_hist_push(ctx.h, h, dev[s + 1].ptr, dev[s].ptr), ctx.h)
    w = rng.uniform(0, 1, (nt, cap))
    dw = _dev(ctx, w)
    out = _dev(ctx, np.full(nt * nodes, 0.25))
    desc = N.MemtermDesc(nt, cap, dw.ptr, 1.5, 1)
    N._check(N.lib().frm_memterm(ctx.h, h, desc, out.ptr), ctx.h)
    got = out.to_host().reshape(nt, nodes)
    inc = np.diff(snaps, axis=0)  # Delta_s, s = 0..depth-1
    for t in range(nt):
        want = 0.25 + 1.5 * sum(w[t, depth - 1 - s] * inc[s] for s in range(depth)) \
            if depth else np.full(nodes, 0.25)
        np.testing.assert_allclose(got[t], want, rtol=1e-12, atol=1e-12 * (1 + np.max(np.abs(want))))
    d = N.ct.c_int32()
    N.lib().frm_hist_depth(h, N.ct.byref(d))
    assert d.value == depth
    N.lib().frm_hist_destroy(h)
    for q in dev + [dw, out]:
        q.free()


def test_hist_capacity_error(ctx):
    h = N._V()
    N._check(N.lib().frm_hist_create(ctx.h, 10, 1, N.PREC_F64, N.ct.byref(h)), ctx.h)
    a = _dev(ctx, np.ones(10))
    N._check(N.lib().frm_hist_push(ctx.h, h, a.ptr, a.ptr), ctx.h)
    assert N.lib().frm_hist_push(ctx.h, h, a.ptr, a.ptr) == N.FRM_E_RANGE
    N.lib().frm_hist_destroy(h)
    a.free()


def test_diag_probes_run_and_check_sizes(ctx):
    # the bandwidth probes behind the bench's measured figures: they run and
    # return a positive rate; the read stream takes whole 8 MiB rows only
    assert ctx.read_bandwidth(64 << 20, reps=2) > 0.0
    assert ctx.copy_bandwidth(16 << 20, reps=2) > 0.0
    src = ctx.malloc(12 << 20)
    ms = N.ct.c_double()
    rc = N.lib().frm_diag_read(ctx.h, src, 12 << 20, 1, N.ct.byref(ms))
    ctx.free(src)
    assert rc == N.FRM_E_INVALID
