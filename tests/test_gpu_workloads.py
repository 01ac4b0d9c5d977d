"""The chosen forms at the full size and length of the configurations that
quote them, against the plain FP64 GPU path on the same seeded inputs:

* 2D Caputo LOD, 1024 x 1024 interior nodes, alpha 0.6, 4000 steps: the
  blocked Toeplitz MFMA form (two BF16 parts, FP32 increments, far products
  over up to 3968 history slots) and the 32-term sum-of-exponentials form
  (FP64 mode planes of 268 MB, above the 200 MiB threshold where they are
  streamed with non-temporal accesses);
* 4096 psi-Caputo problems (64..1024 nodes, k0 exp(b u) diffusivity, three
  Picard passes), 3000 steps: compensated FP32 history against the batched
  FP64 history, and the batched FP64 run against the oracle's H_solver
  restatement on whole runs of two problems;
* the largest Atangana-Baleanu parameter set, 2048 x 2048, 2000 steps: the
  64-term exponential sum with FP32 mode states (1 GiB of planes, streamed)
  and the three-part MFMA form;
* the 512 x 512 workload, alpha 0.7, over its whole 5000 steps: the MFMA
  and SOE forms and the compensated FP32 history.

Bar: max-norm and RMS relative error <= 1e-6 against FP64 for the chosen
forms; <= 1e-12 for FP64 against the oracle."""
import numpy as np
import pytest

from theses_code_amd import _native as N
from theses_code_amd import workloads as W
from theses_code_amd.solvers import MT1DBatch, TGSolver

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(0)
    yield c
    c.close()


def errs(a, b):
    d = a - b
    return (float(np.max(np.abs(d)) / np.max(np.abs(b))),
            float(np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(b * b))))


def _run(ctx, case, alpha, cfg, steps, **kw):
    sN, M, coefs, C0 = case
    s = TGSolver(ctx, sN, M, alpha, cfg["tau"], cfg["sigma"], coefs, C0, capacity=steps, C1=0.0,
                 S=0.0, **kw)
    s.calc_step(steps)
    out = s.b_C
    fit = s.fit_error if kw.get("precision") == N.PREC_SOE else None
    s.close()
    return out, fit


def test_c2_chosen_forms_full_run(ctx):
    cfg = W.C2_WORKLOAD
    case = W.lod2d_case(cfg["n_interior"], cfg["seed"], vmax=cfg["vmax"], D=cfg["D"])
    K, alpha = cfg["steps"], cfg["alpha"]
    print(f"c2: 1024^2, alpha {alpha}, {K} steps, plain FP64", flush=True)
    ref, _ = _run(ctx, case, alpha, cfg, K)
    assert np.all(np.isfinite(ref))
    print("c2: MFMA, two BF16 parts", flush=True)
    mf, _ = _run(ctx, case, alpha, cfg, K, precision=N.PREC_MFMA_BF16, mfma_parts=2)
    mx, rms = errs(mf, ref)
    print(f"c2 MFMA: max {mx:.3e} rms {rms:.3e}", flush=True)
    assert mx <= 1e-6 and rms <= 1e-6
    print("c2: SOE, 32 terms, FP64 planes", flush=True)
    so, fit = _run(ctx, case, alpha, cfg, K, precision=N.PREC_SOE, soe_terms=cfg["soe_terms"])
    mx, rms = errs(so, ref)
    print(f"c2 SOE: fit {fit:.2e} max {mx:.3e} rms {rms:.3e}", flush=True)
    assert fit <= 1e-8 and mx <= 1e-6 and rms <= 1e-6


def test_c4_largest_set_full_run(ctx):
    cfg = W.C4_WORKLOAD
    sets = W.c4_sets()
    p = max(sets, key=lambda q: q["n_interior"])
    # the largest side drawn is below 2048: the set is run at 2048^2, with
    # that set's order and seed
    case = W.ab_diffusion_case(2048, p["seed"], D=cfg["D"])
    K, alpha = cfg["steps"], p["alpha"]
    kw = dict(kernel=N.KERNEL_AB, ab_exact=True)
    print(f"c4: 2048^2, alpha {alpha:.3f}, {K} steps, plain FP64", flush=True)
    ref, _ = _run(ctx, case, alpha, cfg, K, **kw)
    assert np.all(np.isfinite(ref))
    print("c4: SOE, 64 terms, FP32 planes", flush=True)
    so, fit = _run(ctx, case, alpha, cfg, K, precision=N.PREC_SOE, soe_terms=cfg["soe_terms"],
                   soe_f32=True, **kw)
    mx, rms = errs(so, ref)
    print(f"c4 SOE: fit {fit:.2e} max {mx:.3e} rms {rms:.3e}", flush=True)
    assert fit <= 1e-10 and mx <= 1e-6 and rms <= 1e-6
    print("c4: MFMA, three BF16 parts", flush=True)
    mf, _ = _run(ctx, case, alpha, cfg, K, precision=N.PREC_MFMA_BF16, mfma_parts=3, **kw)
    mx, rms = errs(mf, ref)
    print(f"c4 MFMA: max {mx:.3e} rms {rms:.3e}", flush=True)
    assert mx <= 1e-6 and rms <= 1e-6


def test_c3_batch_full_run(ctx):
    from oracle import oracle as O

    cfg = W.C3_WORKLOAD
    probs, U0 = W.mt1d_c3_case()
    K, picard = cfg["steps"], cfg["picard"]

    def run(prec):
        b = MT1DBatch(ctx, probs, cfg["tau"], K, U0, kind=N.MT_PSI, precision=prec,
                      picard=picard)
        b.calc_step(K)
        U = b.state()
        b.close()
        return U

    print(f"c3: {len(probs)} psi problems, {K} steps, batched FP64", flush=True)
    U64 = run(N.PREC_F64)
    print("c3: compensated FP32 history", flush=True)
    U32 = run(N.PREC_F32C)
    a64, a32 = np.concatenate(U64), np.concatenate(U32)
    assert np.all(np.isfinite(a64)) and np.all(np.isfinite(a32))
    mx, rms = errs(a32, a64)
    print(f"c3 F32C: max {mx:.3e} rms {rms:.3e}", flush=True)
    assert mx <= 1e-6 and rms <= 1e-6
    # whole runs of two problems (a small and a large one) by the oracle
    sizes = [p["N"] for p in probs]
    for i in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        p = probs[i]
        print(f"c3: oracle, problem {i} (N = {p['N']})", flush=True)
        o = O.MTSolver(p["N"], p["L"], p["alpha"], cfg["tau"], U0[i], psi=N.MT_PSI,
                       beta=p["beta"], form=p["form"], k0=p["k0"], b=p["b"], top=p["top"],
                       top_val=p["top_val"], bottom_cond=p["bottom_cond"], picard=picard)
        o.step(K)
        ref = o.state()
        err = float(np.max(np.abs(U64[i] - ref)) / np.max(np.abs(ref)))
        print(f"c3 problem {i}: FP64 vs oracle {err:.3e}", flush=True)
        assert err <= 1e-12


def test_c1_chosen_forms_full_depth(ctx):
    # the 512^2, alpha 0.7 workload over its whole 5000 steps (history depth
    # up to 4999): both chosen forms against the plain FP64 path
    cfg = W.C1_WORKLOAD
    case = W.lod2d_case(cfg["n_interior"], cfg["seed"], vmax=cfg["vmax"], D=cfg["D"])
    K, alpha = cfg["steps"], cfg["alpha"]
    print(f"c1: 512^2, alpha {alpha}, {K} steps, plain FP64", flush=True)
    ref, _ = _run(ctx, case, alpha, cfg, K)
    for name, kw in (("MFMA", dict(precision=N.PREC_MFMA_BF16)),
                     ("SOE", dict(precision=N.PREC_SOE, soe_terms=32)),
                     ("F32C", dict(precision=N.PREC_F32C))):
        got, _ = _run(ctx, case, alpha, cfg, K, **kw)
        mx, rms = errs(got, ref)
        print(f"c1 {name}: max {mx:.3e} rms {rms:.3e}", flush=True)
        assert mx <= 1e-6 and rms <= 1e-6
