"""frm_memterm_static (the allocation-free memory term: partial sums in a
static per-GPU array) through the C ABI against frm_memterm on the same
store, weights and output: bit-identical results on two calls in a row, the
second on a smaller input (fewer nodes, more time chunks, one target),
and FRM_E_RANGE for a store whose partials do not fit the array."""
import numpy as np
import pytest

from theses_code_amd import _native as N

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(0)
    yield c
    c.close()


def _store(ctx, rng, nodes, depth):
    h = N._V()
    N._check(N.lib().frm_hist_create(ctx.h, nodes, depth, N.PREC_F64, N.ct.byref(h)), ctx.h)
    a = N.DeviceArray.from_host(ctx, rng.normal(0, 1, nodes))
    b = N.DeviceArray.from_host(ctx, rng.normal(0, 1, nodes))
    for s in range(depth):
        # alternate the order so the increments change sign from slot to slot
        x, y = (a, b) if s % 2 else (b, a)
        N._check(N.lib().frm_hist_push(ctx.h, h, x.ptr, y.ptr), ctx.h)
    return h, [a, b]


def _both(ctx, h, nodes, nt, depth, rng):
    w = N.DeviceArray.from_host(ctx, rng.uniform(0, 1, nt * depth))
    init = rng.normal(0, 1, nt * nodes)
    outs = []
    for fn in (N.lib().frm_memterm, N.lib().frm_memterm_static):
        out = N.DeviceArray.from_host(ctx, init)
        desc = N.MemtermDesc(nt, depth, w.ptr, 0.75, 1)
        N._check(fn(ctx.h, h, desc, out.ptr), ctx.h)
        outs.append(out.to_host())
        out.free()
    w.free()
    return outs


def test_static_memterm_matches_memterm_twice(ctx):
    rng = np.random.default_rng(11)
    # first call: 300000 nodes (one time chunk), two targets, depth 96
    # second call: 5000 nodes (many time chunks), one target, depth 700
    for nodes, nt, depth in ((300_000, 2, 96), (5_000, 1, 700)):
        print(f"frm_memterm_static: nodes {nodes}, targets {nt}, depth {depth}", flush=True)
        h, keep = _store(ctx, rng, nodes, depth)
        ref, got = _both(ctx, h, nodes, nt, depth, rng)
        assert np.all(np.isfinite(ref))
        assert np.array_equal(got, ref)
        N.lib().frm_hist_destroy(h)
        for a in keep:
            a.free()


def test_static_memterm_rejects_an_oversized_store(ctx):
    # 1.5M nodes with two targets need 3M partial sums > 2^21
    h, keep = _store(ctx, np.random.default_rng(12), 1_500_000, 2)
    w = N.DeviceArray.from_host(ctx, np.ones(4))
    out = N.DeviceArray(ctx, 2 * 1_500_000)
    desc = N.MemtermDesc(2, 2, w.ptr, 1.0, 0)
    rc = N.lib().frm_memterm_static(ctx.h, h, desc, out.ptr)
    assert rc == N.FRM_E_RANGE
    assert b"static workspace" in N.lib().frm_last_error(ctx.h)
    # the allocating form takes the same store
    N._check(N.lib().frm_memterm(ctx.h, h, desc, out.ptr), ctx.h)
    N.lib().frm_hist_destroy(h)
    for a in keep + [w, out]:
        a.free()
