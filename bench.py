#!/usr/bin/env python3
"""Benchmark: node-steps/s of the fractional solvers' hot path on one MI355X.

Default workload (--config c2): the 2D Caputo convection-diffusion LOD model
at 1024 x 1024 interior nodes, alpha 0.6, tau 1e-3, 4000 steps, seeded
divergence-free velocity (max|v| 0.1), D 1e-3, 8 Gaussian blobs.  The plain
FP64 full-history path (the drop-in) and the chosen forms (blocked Toeplitz
on MFMA with split BF16 products; the 32-term exponential sum) each run the
whole 4000 steps from t = 0 on the same inputs; their fields at step 4000
give the max / RMS relative error against FP64.  The K = --steps timed
steps (default 64) are a window at history depth T/2 = 2000 inside those
runs, starting on a 64-step block boundary (1984); the MFMA form forms the
next block's far product in node slices of about one round of workgroups
(16 at this size on a 256-CU MI355X, one every 4 steps), so a window of K
steps pays for the slices that start in it.  The W = --warmup steps just before it
are the last untimed ones.  `value` is the window rate of the fastest form
within 1e-6 of FP64; the whole-run rates stand beside it.

    python bench.py [--gpus 1] [--steps 64] [--warmup 5] [--config c2]

Prints one JSON line.  `roofline` describes the chosen form's dominant
kernel -- for the MFMA form the far product k_block_far: algorithmic bytes
per launch = 4 B x settled slots x slice nodes (read) + 4 B x 128 rows x
slice nodes (written), over its HIP-event-timed duration on the solver's
stream inside the window.  The plain path's memory-term kernel has its own
roofline object.
`cpu_baseline` times the oracle (a port of the reference's CPU path) on
LOD steps of the same workload at the window's history depth.  The other
configurations (--config c0|c1|c3|c4|ab1d|ab1d_series|timesum3d|soe_step)
print their own lines.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

from theses_code_amd.rng import Xoshiro  # noqa: E402  (seeded inputs; no GPU state)

METRIC = "node-steps/s at ≤1e-6 rel err vs FP64: chosen form vs plain FP64, 1 GPU"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="c2/c4: timed steps, a window at history depth T/2 of whole runs "
                         "(default 64, one block); other configurations: the run length "
                         "(default: the configuration's)")
    ap.add_argument("--warmup", type=int, default=5,
                    help="untimed steps just before the timed ones (c1, c2, c4)")
    ap.add_argument("--total-steps", type=int, default=None,
                    help="c2/c4: run length T (default: the configuration's, 4000 / 2000)")
    ap.add_argument("--cpu-1thread", action="store_true",
                    help="c2/c4: also time one oracle step on one thread")
    ap.add_argument("--breakdown-soe", action="store_true",
                    help="c2/c4: per-kernel times of the SOE form (a second whole run)")
    ap.add_argument("--cpu-steps", type=int, default=4,
                    help="oracle steps timed for the CPU baseline (0 skips it)")
    ap.add_argument("--parity-steps", type=int, default=12)
    ap.add_argument("--no-copy-probe", action="store_true",
                    help="skip the copy / read bandwidth probes")
    ap.add_argument("--config", default="c2",
                    choices=["c1", "c0", "c2", "c3", "c4", "ab1d", "ab1d_series", "timesum3d",
                             "soe_step"],
                    help="c2 (default): 2D LOD 1024x1024, alpha 0.6, 4000 steps, plain FP64 vs "
                         "the chosen forms; c1: 2D LOD 512x512 plain FP64 (5000 steps); c0: 1D "
                         "mt_minimal N=200; c3: 4096 "
                         "psi-Caputo problems, FP64 vs compensated FP32; c4: 16 AB sets, plain "
                         "FP64 vs the chosen forms; ab1d: 1024 AB test problems, full history; "
                         "ab1d_series: 1024 AB test problems, the reference's series history; "
                         "timesum3d: time sums of the 3D irrigation solver, 128x128x64 nodes; "
                         "soe_step: the standalone SOE history, 39.2M nodes, 64 modes")
    ap.add_argument("--nprob", type=int, default=None, help="c3: number of problems")
    ap.add_argument("--sets", type=int, default=None, help="c4: number of parameter sets")
    ap.add_argument("--plain-sample", type=int, default=64,
                    help="c3: problems timed through the one-solver-per-problem plain path")
    ap.add_argument("--soe-f32", action="store_true", help="c2: FP32 mode states")
    ap.add_argument("--soe-f64", action="store_true",
                    help="c4: FP64 mode states (the configuration's own are FP32)")
    ap.add_argument("--sweep-w", type=int, default=0,
                    help="wavefronts per sweep line for the chosen forms (0: autotuned)")
    ap.add_argument("--breakdown-plain", action="store_true",
                    help="c2/c4: also time plain FP64 per kernel (a second full run)")
    ap.add_argument("--no-patterns", action="store_true",
                    help="c0/c3: skip the launch-pattern runs (fixed history block length 64, "
                         "steps replayed as HIP graphs)")
    ap.add_argument("--forms", default="f32c,soe,mfma",
                    help="comma-separated forms timed against plain FP64 on the same workload "
                         "(soe, mfma; binomial = the reference's own series history, c1/c2; "
                         "empty for none)")
    return ap.parse_args()


def cpu_baseline(case, cfg, depth, nsteps, ab=None, steps_1thread=1):
    """The oracle (a port of the reference's CPU path) on a window of LOD steps
    at history depth `depth` (prefilled), with all OpenMP threads and then
    with one."""
    from oracle import oracle as O

    sN, M, coefs, C0 = case
    o = O.TGSolver(sN, M, cfg["alpha"], cfg["tau"], cfg["sigma"], 0.0, 0.0, *coefs, C0, ab=ab)
    o.prefill(depth, seed=12345)
    o.step(1)  # first step at this depth also touches every snapshot (and makes the weights)
    threads = O.get_threads()
    t0 = time.perf_counter()
    o.step(nsteps)
    dt = time.perf_counter() - t0
    v1 = None
    if steps_1thread > 0:
        O.set_threads(1)
        t0 = time.perf_counter()
        o.step(steps_1thread)
        v1 = (sN - 1) * (M - 1) * steps_1thread / (time.perf_counter() - t0)
        O.set_threads(threads)
    nodes = (sN - 1) * (M - 1)
    del o
    return {
        "value": nodes * nsteps / dt,
        "unit": "node-steps/s",
        "cores": threads,
        "value_1thread": v1,
        "kind": "port",
        "sample": (f"{nsteps} LOD step(s) ({steps_1thread} more on one thread) of the "
                   f"{sN - 1}x{M - 1} workload from history depth {depth + 1} "
                   f"(synthetic prefilled history; "
                   f"per-step cost is linear in depth, so depth T/2 is the run average); "
                   f"oracle/frm_oracle.c, OpenMP over the nodes of each line as "
                   f"fr_d_t_g_2d.cpp:3341"),
        "seconds": dt,
    }


def run_mt1d(args, cfg_name):
    """c0 / c3: the batched 1D H_solver; c3 times the FP64 and the compensated
    FP32 forms on the same inputs and reports the FP32 form's error."""
    from theses_code_amd import _native as N
    from theses_code_amd.solvers import MT1DBatch
    from theses_code_amd import workloads as Wl

    if cfg_name == "c0":
        cfg = Wl.C0_WORKLOAD
        prob, u0 = Wl.mt1d_c0_case()
        probs, U0, kind, picard = [prob], [u0], N.MT_L1, 1
    else:
        cfg = Wl.C3_WORKLOAD
        probs, U0 = Wl.mt1d_c3_case(nprob=args.nprob or cfg["nprob"])
        kind, picard = N.MT_PSI, cfg["picard"]
    K = args.steps if args.steps is not None else cfg["steps"]
    nodes = sum(p["N"] + 1 for p in probs)  # nodes 0..N of every problem
    ctx = N.Context(0)
    results = {}

    def one_run(prec, timing, graph=False, hist_block=0):
        b = MT1DBatch(ctx, probs, cfg["tau"], K, U0, kind=kind, precision=prec, picard=picard,
                      graph=graph, hist_block=hist_block)
        hb = b.hist_block
        if timing:
            ctx.timing_reset()
            ctx.timing_enable((1 << N.K_MEMTERM) | (1 << N.K_OTHER) | (1 << N.K_SWEEP1D))
        ctx.sync()
        t0 = time.perf_counter()
        b.calc_step(K)
        ctx.sync()
        dt = time.perf_counter() - t0
        kt = None
        if timing:
            ctx.timing_enable(0)
            kt = {name: ctx.timing(k) for name, k in (("memterm", N.K_MEMTERM),
                                                       ("weights", N.K_OTHER),
                                                       ("sweep", N.K_SWEEP1D))}
        U = np.concatenate(b.state())
        b.close()
        one_run.hist_block = hb
        return dt, U, kt

    # The wall clock comes from a run without per-launch timing events (c0's
    # steps are tens of microseconds, and with the events on the psi weights
    # are no longer formed ahead on their own stream); the kernel times come
    # from a second, event-timed run.
    two_pass = True
    for prec in ([N.PREC_F64, N.PREC_F32C] if cfg_name == "c3" else [N.PREC_F64]):
        dt, U, kt = one_run(prec, not two_pass)
        if two_pass:
            kt = one_run(prec, True)[2]
        results[prec] = (dt, U, kt)
    dt, U64, kt = results[N.PREC_F64]
    hist_block = one_run.hist_block
    patterns = None
    if not args.no_patterns:
        # the launch patterns of the chosen form on the same run: eager
        # launches with the autotuned history block length (the line's
        # value), the same with the former fixed length 64, and every step
        # replayed as a HIP graph (one graph per launch sequence, built on
        # first use inside the timed region, node arguments set per step)
        cp = N.PREC_F32C if N.PREC_F32C in results else N.PREC_F64
        dte, Ue, _ = results[cp]
        patterns = {"precision": "f32c" if cp == N.PREC_F32C else "f64",
                    "eager": {"value": nodes * K / dte, "ms_per_step": dte * 1e3 / K,
                              "hist_block": hist_block}}
        if hist_block != 64:
            dtf, Uf, _ = one_run(cp, False, hist_block=64)
            patterns["eager_hist_block_64"] = {
                "value": nodes * K / dtf, "ms_per_step": dtf * 1e3 / K,
                "max_rel_diff": float(np.max(np.abs(Uf - Ue)) / np.max(np.abs(Ue)))}
        dtg, Ug, _ = one_run(cp, False, graph=True)
        patterns["graph_replay"] = {"value": nodes * K / dtg, "ms_per_step": dtg * 1e3 / K,
                                    "identical_to_eager": bool(np.array_equal(Ug, Ue))}
    per_problem = None
    if cfg_name == "c3" and args.plain_sample > 0:
        # The plain FP64 drop-in as the reference's callers use it: one solver
        # object per problem (H_solver per problem), stepped one after the
        # other on the context's stream -- timed on the first `plain_sample`
        # problems of the batch (bounded), every step of the run.
        P = min(args.plain_sample, len(probs))
        sample_nodes = sum(p["N"] + 1 for p in probs[:P])
        offs = np.concatenate([[0], np.cumsum([p["N"] + 2 for p in probs])])
        singles = [MT1DBatch(ctx, [probs[i]], cfg["tau"], K, [U0[i]], kind=kind,
                             precision=N.PREC_F64, picard=picard) for i in range(P)]
        ctx.sync()
        t0 = time.perf_counter()
        for b in singles:
            b.calc_step(K)
        ctx.sync()
        dts = time.perf_counter() - t0
        dmax = 0.0
        for i, b in enumerate(singles):
            u = b.state()[0]
            ub = U64[offs[i]:offs[i + 1]]
            dmax = max(dmax, float(np.max(np.abs(u - ub)) / np.max(np.abs(ub))))
            b.close()
        per_problem = {"form": "plain FP64 drop-in, one solver per problem, sequential launches",
                       "value": sample_nodes * K / dts, "seconds": dts, "problems_timed": P,
                       "sample_nodes": sample_nodes,
                       "max_rel_diff_vs_batched_fp64": dmax}
    line = {"metric": METRIC, "value": nodes * K / dt, "unit": "node-steps/s", "n_gpus": 1,
            "steps": K, "warmup": 0, "ms_per_step": dt * 1e3 / K, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": cfg["name"], "problems": len(probs), "nodes": nodes,
                       "tau": cfg["tau"], "picard": picard, "parallelism": "1 GPU"},
            "kernels_ms": {k: v[1] for k, v in kt.items()}}
    ms, nb = kt["memterm"][1], kt["memterm"][2]
    traffic = None
    tf = os.path.join(REPO, "profiles", "mt1d_psi_traffic.json")
    if cfg_name == "c3" and os.path.exists(tf) and kt["memterm"][0]:
        # PMC-measured ratio of HBM to algorithmic bytes of this kernel
        with open(tf) as f:
            traffic = json.load(f)["traffic_over_algorithmic"] * nb / kt["memterm"][0]
    line["roofline"] = {"bound": "hbm", "achieved": nb / (ms * 1e-3) / 1e9 if ms else 0.0,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": (nb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if ms else 0.0,
                        "traffic": traffic, "kernel": "k_mt_memterm_f64"}
    line["config"]["hist_block"] = hist_block
    if patterns:
        line["launch_patterns"] = patterns
    if cfg_name == "c3" and args.cpu_steps > 0:
        # the oracle (H_solver restatement with the psi weights rebuilt every
        # step, as the reference does) on whole runs of the first problems,
        # all threads, then one problem on one thread; also parity of the
        # batched FP64 GPU run on those problems
        from oracle import oracle as O
        offs = np.concatenate([[0], np.cumsum([p["N"] + 2 for p in probs])])

        def oracle_run(i):
            p = probs[i]
            o = O.MTSolver(p["N"], p.get("L", 1.0), p["alpha"], cfg["tau"], U0[i], psi=kind,
                           beta=p["beta"], form=p["form"], k0=p["k0"], b=p["b"], top=p["top"],
                           top_val=p["top_val"], bottom_cond=p["bottom_cond"], picard=picard)
            o.step(K)
            return o.state()

        threads = O.get_threads()
        t0 = time.perf_counter()
        par = 0.0
        for i in range(2):
            ref = oracle_run(i)
            ub = U64[offs[i]:offs[i + 1]]
            par = max(par, float(np.max(np.abs(ub - ref)) / np.max(np.abs(ref))))
        dcpu = time.perf_counter() - t0
        n2 = sum(p["N"] + 1 for p in probs[:2])
        O.set_threads(1)
        t0 = time.perf_counter()
        oracle_run(2)
        d1 = time.perf_counter() - t0
        O.set_threads(threads)
        line["cpu_baseline"] = {"value": n2 * K / dcpu, "unit": "node-steps/s", "cores": threads,
                                "value_1thread": (probs[2]["N"] + 1) * K / d1, "kind": "port",
                                "seconds": dcpu,
                                "sample": f"whole {K}-step runs of problems 0 and 1 (problem 2 on "
                                          "one thread) by oracle/frm_oracle.c, psi weights rebuilt "
                                          "every step as mt_psi_caputo.cpp:1461-1491"}
        line["parity"] = {"vs": "oracle H_solver restatement, whole runs of problems 0 and 1",
                          "max_rel_err": par}
    if cfg_name == "c0" and args.cpu_steps > 0:
        # the reference's own CPU case: the oracle's H_solver restatement on
        # the whole run (all steps), the field compared with the GPU's
        from oracle import oracle as O
        p = probs[0]
        o = O.MTSolver(p["N"], p.get("L", 1.0), p["alpha"], cfg["tau"], U0[0], form=p["form"],
                       D=p["D"], top=p["top"], top_val=p["top_val"],
                       bottom_cond=p["bottom_cond"], Hbottom=p["Hbottom"])
        t0 = time.perf_counter()
        o.step(K)
        dcpu = time.perf_counter() - t0
        ref = o.state()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        line["cpu_baseline"] = {"value": nodes * K / dcpu, "unit": "node-steps/s",
                                "cores": threads, "kind": "port", "seconds": dcpu,
                                "sample": f"the whole run ({K} steps) by oracle/frm_oracle.c "
                                          "(H_solver::calc_step restatement; OpenMP over the "
                                          "history sum once it exceeds 2e5 terms, as "
                                          "mt_minimal.cpp:909)"}
        line["parity"] = {"vs": "oracle H_solver restatement, whole run",
                          "max_rel_err": float(np.max(np.abs(U64 - ref)) / np.max(np.abs(ref)))}
    if N.PREC_F32C in results:
        dt32, U32, kt32 = results[N.PREC_F32C]
        d = U32 - U64
        line["chosen_form"] = {
            "form": "batched launch of all problems, FP32 history, compensated FP32 sums, "
                    "FP64 weights/Thomas/state",
            "value": nodes * K / dt32, "speedup_vs_batched_fp64": dt / dt32,
            "max_rel_err_vs_fp64": float(np.max(np.abs(d)) / np.max(np.abs(U64))),
            "rms_rel_err_vs_fp64": float(np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(U64 * U64))),
            "kernels_ms": {k: v[1] for k, v in kt32.items()}}
        line["batched_fp64"] = {"value": nodes * K / dt, "seconds": dt,
                                "kernels_ms": line["kernels_ms"], "roofline": line["roofline"]}
        # the value's form: the compensated FP32 memory term (4 B per slot and
        # node of the problems' own nodes; four nodes per lane at this size)
        ms32, nb32 = kt32["memterm"][1], kt32["memterm"][2]
        a32 = nb32 / (ms32 * 1e-3) / 1e9 if ms32 else 0.0
        line["roofline"] = {"bound": "hbm", "achieved": a32, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": a32 / HBM_PEAK_GBS, "traffic": None,
                            "traffic_source": "not profiled (no counter pass over this command)",
                            "kernel": "k_mt_memterm_f32c",
                            "launches": kt32["memterm"][0],
                            "avg_launch_ms": ms32 / kt32["memterm"][0] if kt32["memterm"][0] else None,
                            "bytes_model": "4 B x depth x the problems' own nodes (padding lanes "
                                           "load nothing)"}
        line["value"] = nodes * K / dt32
        line["ms_per_step"] = dt32 * 1e3 / K
        line["dtype"] = "f64 weights/Thomas/state, f32 compensated history"
        if per_problem:
            line["plain_fp64_per_problem"] = per_problem
            line["speedup_vs_plain_fp64"] = line["value"] / per_problem["value"]
    print(json.dumps(line))
    ctx.close()


def run_ab1d(args, series=False):
    """Batched 1D Atangana-Baleanu solves (solve_and_check form; `series`:
    the series history of simple = 0, else the full history of simple = 1):
    the batch, the one-solver-per-problem drop-in on a sample, and the
    oracle's CPU solve on a sample, which is also the parity check."""
    from oracle import oracle as O
    from theses_code_amd import _native as N
    from theses_code_amd import workloads as Wl
    from theses_code_amd.solvers import MT1DBatch

    if series:
        cfg = Wl.AB1D_SERIES_WORKLOAD
        probs = Wl.ab1d_series_case(nprob=args.nprob or cfg["nprob"])
        oracle_solve, kernel, n_cpu = O.ab_solve_series, "k_ab_series", 2
        plain_cap = 16
    else:
        cfg = Wl.AB1D_WORKLOAD
        probs = Wl.ab1d_case(nprob=args.nprob or cfg["nprob"])
        oracle_solve, kernel, n_cpu = O.ab_solve, "k_mt_memterm_f64", 4
        plain_cap = args.plain_sample
    K = args.steps if args.steps is not None else cfg["steps"]
    tau = cfg["tau"]
    nodes = sum(p["N"] + 1 for p in probs)
    ctx = N.Context(0)
    U0 = [np.zeros(p["N"] + 2) for p in probs]
    t0 = time.perf_counter()
    b = MT1DBatch(ctx, probs, tau, K, U0, kind=N.MT_AB)
    setup = time.perf_counter() - t0
    # wall clock without per-launch timing events (the steps are tens of
    # microseconds); the kernel times come from a second, event-timed run
    ctx.sync()
    t0 = time.perf_counter()
    b.calc_step(K)
    ctx.sync()
    dt = time.perf_counter() - t0
    states = b.state()
    b.close()
    b = MT1DBatch(ctx, probs, tau, K, U0, kind=N.MT_AB)
    ctx.timing_reset()
    ctx.timing_enable((1 << N.K_MEMTERM) | (1 << N.K_SWEEP1D))
    b.calc_step(K)
    ctx.timing_enable(0)
    kt = {name: ctx.timing(k) for name, k in (("memterm", N.K_MEMTERM), ("sweep", N.K_SWEEP1D))}
    b.close()
    # the plain drop-in: one solver per problem, on a sample
    P = min(plain_cap, len(probs))
    singles = [MT1DBatch(ctx, [probs[i]], tau, K, [U0[i]], kind=N.MT_AB) for i in range(P)]
    ctx.sync()
    t0 = time.perf_counter()
    for s1 in singles:
        s1.calc_step(K)
    ctx.sync()
    dts = time.perf_counter() - t0
    for s1 in singles:
        s1.close()
    sample_nodes = sum(p["N"] + 1 for p in probs[:P])
    # CPU: the oracle's whole solve of the first problems, one thread (the
    # reference's solve_and_check has no parallel loop); parity against them
    t0 = time.perf_counter()
    cn, par = 0, 0.0
    for i, p in enumerate(probs[:n_cpu]):
        _, Cref = oracle_solve(p["alpha"], K * tau, K, p["N"], cfg["maxl"], cfg["maxkiter"],
                               cfg["eps"], p["src"])
        cn += p["N"] + 1
        par = max(par, float(np.max(np.abs(states[i][:p["N"] + 1] - Cref)) / np.max(np.abs(Cref))))
    dcpu = time.perf_counter() - t0
    what = "series history (simple = 0)" if series else "full history (simple = 1)"
    cpu = {"value": cn * K / dcpu, "unit": "node-steps/s", "cores": 1, "kind": "port",
           "sample": f"oracle solve_and_check restatement, {what}, first {n_cpu} problems, all "
                     f"{K} steps (weights included, as the reference computes them in the run)",
           "seconds": dcpu}
    ms, nb = kt["memterm"][1], kt["memterm"][2]
    line = {"metric": METRIC, "value": nodes * K / dt, "unit": "node-steps/s", "n_gpus": 1,
            "steps": K, "warmup": 0, "ms_per_step": dt * 1e3 / K, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": cfg["name"], "problems": len(probs), "nodes": nodes,
                       "tau": tau, "history": what, "parallelism": "1 GPU"},
            "setup_seconds": setup,
            "kernels_ms": {k: v[1] for k, v in kt.items()},
            "roofline": {"bound": "hbm", "achieved": nb / (ms * 1e-3) / 1e9 if ms else 0.0,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": (nb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if ms else 0.0,
                         "traffic": None, "kernel": kernel},
            "parity": {"vs": f"oracle solve_and_check restatement, {what}", "problems": n_cpu,
                       "max_rel_err": par},
            "plain_fp64_per_problem": {"value": sample_nodes * K / dts, "seconds": dts,
                                       "problems_timed": P},
            "speedup_vs_plain_fp64": (nodes * K / dt) / (sample_nodes * K / dts),
            "cpu_baseline": cpu}
    print(json.dumps(line))
    ctx.close()


def _lod_run(ctx, N, case, alpha, tau, sigma, T, precision, window=None, warmup=0,
             mask=0, timing=False, stop_after_window=False, **kw):
    """One run of T steps from t = 0.

    `window` = (start, K): the steps start+1 .. start+K are timed on their
    own between two synchronisations, with HIP events on the launches of the
    kernel classes in `mask` (the rest of the run carries no events); the
    `warmup` steps just before the window are the last untimed ones.
    `stop_after_window`: the run ends with the window (the whole-run time
    and the field returned are then those of step start + K).
    `timing`: events on every launch of the run instead (a breakdown run;
    its wall time then includes the events).  Returns the whole run's wall
    time, the final field, the per-class event totals and the window's data."""
    from theses_code_amd.solvers import TGSolver

    sN, M, coefs, C0 = case
    s = TGSolver(ctx, sN, M, alpha, tau, sigma, coefs, C0, capacity=T, C1=0.0, S=0.0,
                 precision=precision, **kw)
    classes = (("memterm", N.K_MEMTERM), ("rows", N.K_ROWS), ("cols", N.K_COLS),
               ("blockfar", N.K_BLOCKFAR))
    ctx.timing_reset()
    ctx.timing_enable(sum(1 << k for _, k in classes) if timing else 0)
    ctx.sync()
    t0 = time.perf_counter()
    win = {}
    if window:
        start, K = window
        pre = max(start - warmup, 0)
        s.calc_step(pre)
        s.calc_step(start - pre)
        ctx.sync()
        ctx.timing_enable(mask)
        tw = time.perf_counter()
        s.calc_step(K)
        ctx.sync()
        win["seconds"] = time.perf_counter() - tw
        ctx.timing_enable(0)
        win["kernels"] = {name: ctx.timing(k) for name, k in classes if mask >> k & 1}
        if not stop_after_window:
            s.calc_step(T - start - K)
    else:
        s.calc_step(T)
    ctx.sync()
    dt = time.perf_counter() - t0
    ctx.timing_enable(0)
    kt = {}
    for name, k in classes if timing else ():
        n, ms, _ = ctx.timing(k)
        if n:
            kt[name] = ms
    out = s.b_C
    fit = s.fit_error
    win["sweep_shape"] = s.sweep_shape
    win["far_slices"] = s.far_slices
    s.close()
    return dt, out, kt, fit, win


FORMS = {"f32c": "FP32 increments and weights, Kahan-compensated FP32 history sums, FP64 "
                 "sweeps",
         "soe": "sum-of-exponentials history (FP64 mode states unless stated)",
         "mfma": "blocked Toeplitz on MFMA: FP32 increments, split hi+lo BF16, FP32 acc, "
                 "64-step blocks, FP64 near part",
         "binomial": "the reference's single-series history (TG_solver dt_alg 3, 25 terms, "
                     "int_err 1e-5): its method, not a chosen form"}
BINOMIAL_TERMS = 25
BLOCK_STEPS = 64


def _stored_ratio(name):
    """PMC-measured ratio of HBM bytes (FETCH_SIZE + WRITE_SIZE, calibrated on
    the copy kernel) to algorithmic bytes of a kernel, from profiles/."""
    tf = os.path.join(REPO, "profiles", name)
    if not os.path.exists(tf):
        return None, None
    with open(tf) as f:
        j = json.load(f)
    return j["traffic_over_algorithmic"], "profiles/" + name


def run_lod_forms(args, cfg_name):
    """2D LOD at 1024^2 (c2) or the 16 AB sets (c4): plain FP64 full history
    against the chosen forms on the same inputs.  Every run goes the whole
    configured length T from t = 0; the K = --steps timed steps are a window
    at history depth T/2 (the run average, since per-step cost is linear in
    depth), started on a 64-step block boundary (the MFMA form's far
    product is formed in node slices, one every 64 / slices steps).  The line's value is the
    window rate of the fastest form within 1e-6 of FP64 (error taken at step
    T); the whole-run rates are reported beside it.  Each window is timed
    twice on the same steps: in the whole run, with HIP events on the form's
    dominant kernel (its launch durations give the roofline), and in a run
    that stops after the window and records no events (the rate: two event
    records per far-product launch cost about 3.5% of the window)."""
    from theses_code_amd import _native as N
    from theses_code_amd import workloads as Wl

    ctx = N.Context(0)
    if cfg_name == "c2":
        cfg = Wl.C2_WORKLOAD
        runs = [dict(alpha=cfg["alpha"], case=Wl.lod2d_case(cfg["n_interior"], cfg["seed"],
                                                            vmax=cfg["vmax"], D=cfg["D"]))]
        kw = dict(kernel=N.KERNEL_CAPUTO)
    else:
        cfg = Wl.C4_WORKLOAD
        sets = Wl.c4_sets(nsets=args.sets or cfg["nsets"])
        runs = [dict(alpha=p["alpha"], case=Wl.ab_diffusion_case(p["n_interior"], p["seed"],
                                                                 D=cfg["D"])) for p in sets]
        kw = dict(kernel=N.KERNEL_AB, ab_exact=True)
    T = args.total_steps or cfg["steps"]
    K = args.steps if args.steps is not None else BLOCK_STEPS
    if not 1 <= K <= T:
        raise SystemExit(f"bench.py: --steps {K} must lie in 1..{T}")
    # window start: depth T/2 rounded down to a block boundary, early enough
    # for K steps
    start = min(T // 2, T - K) // BLOCK_STEPS * BLOCK_STEPS
    W = min(args.warmup, start)
    forms = [f for f in args.forms.split(",") if f]
    # the AB sets run their exponential-series recurrences in FP32, as the
    # configuration states; the 1024^2 Caputo run keeps FP64 states unless asked
    soe_f32 = args.soe_f32 or (cfg_name == "c4" and not args.soe_f64)
    fkw = {"soe": dict(precision=N.PREC_SOE, soe_terms=cfg["soe_terms"], soe_f32=soe_f32,
                       sweep_w=args.sweep_w),
           "mfma": dict(precision=N.PREC_MFMA_BF16, sweep_w=args.sweep_w),
           "f32c": dict(precision=N.PREC_F32C, sweep_w=args.sweep_w),
           "binomial": dict(precision=N.PREC_BINOMIAL, series_terms=BINOMIAL_TERMS,
                            sweep_w=args.sweep_w)}
    if cfg_name == "c4" and "binomial" in forms:
        print("# the binomial-series history is the Caputo solver's (dt_alg 3); skipped on the AB "
              "sets", file=sys.stderr)
        forms.remove("binomial")
    # events in the window on the dominant kernel of each form only: the
    # far product (one launch per block) and the plain memory term (ms long)
    fmask = {"mfma": 1 << N.K_BLOCKFAR, "soe": 1 << N.K_MEMTERM, "f32c": 1 << N.K_MEMTERM,
             "binomial": 0}
    tot_nodes, t_plain, tw_plain, tw_plain_ev, kt_plain = 0, 0.0, 0.0, 0.0, {}
    plain_win = {"launches": 0, "ms": 0.0, "bytes": 0.0}
    res = {f: dict(seconds=0.0, window=0.0, window_events=0.0, errs=[], rms=[], fits=[], kernels_ms={},
                   far={"launches": 0, "ms": 0.0, "bytes": 0.0, "flops": 0.0, "slices": []},
                   mem={"launches": 0, "ms": 0.0, "bytes": 0.0}, shapes=[]) for f in forms}
    for r in runs:
        sN, M = r["case"][0], r["case"][1]
        nodes = (sN - 1) * (M - 1)
        tot_nodes += nodes
        common = (ctx, N, r["case"], r["alpha"], cfg["tau"], cfg["sigma"], T)
        dp, up, _, _, wp = _lod_run(*common, precision=N.PREC_F64, window=(start, K), warmup=W,
                                    mask=1 << N.K_MEMTERM, **kw)
        t_plain += dp
        tw_plain_ev += wp["seconds"]
        tw_plain += _lod_run(*common, precision=N.PREC_F64, window=(start, K), warmup=W,
                             stop_after_window=True, **kw)[4]["seconds"]
        n_, ms_, by_ = wp["kernels"]["memterm"]
        plain_win["launches"] += n_
        plain_win["ms"] += ms_
        plain_win["bytes"] += by_
        if args.breakdown_plain:
            for k, v in _lod_run(*common, precision=N.PREC_F64, timing=True, **kw)[2].items():
                kt_plain[k] = kt_plain.get(k, 0.0) + v
        msg = f"# set alpha={r['alpha']:.3f} grid={sN - 1}^2 plain {dp:.3f}s"
        for f in forms:
            ds, us, _, fit, w0 = _lod_run(*common, window=(start, K), warmup=W, mask=fmask[f],
                                          **kw, **fkw[f])
            q = res[f]
            q["seconds"] += ds
            q["window_events"] += w0["seconds"]
            if fmask[f]:
                # the same window again, without events, for the rate
                w0n = _lod_run(*common, window=(start, K), warmup=W, stop_after_window=True,
                               **kw, **fkw[f])[4]
                q["window"] += w0n["seconds"]
            else:
                q["window"] += w0["seconds"]
            q["shapes"].append(w0["sweep_shape"])
            if f == "mfma":
                # the far slices in the window: algorithmic bytes from the
                # library (every settled slot read once, 4 B per node; the
                # 128 F rows written once, 4 B per node and row); a slice
                # covers 1/slices of the nodes, so the useful flops (2 x 128
                # per slot and node) are 64 x (bytes - 512 B x nodes / slices)
                n_, ms_, by_ = w0["kernels"]["blockfar"]
                sl = max(1, w0["far_slices"])
                q["far"]["launches"] += n_
                q["far"]["ms"] += ms_
                q["far"]["bytes"] += by_
                q["far"]["flops"] += 64.0 * (by_ - 512.0 * nodes * n_ / sl)
                q["far"]["slices"].append(sl)
            if f in ("soe", "f32c"):
                n_, ms_, by_ = w0["kernels"]["memterm"]
                q["mem"]["launches"] += n_
                q["mem"]["ms"] += ms_
                q["mem"]["bytes"] += by_
            # per-kernel breakdown from a second, event-timed run
            if f not in ("soe", "f32c") or args.breakdown_soe:
                for k, v in _lod_run(*common, timing=True, **kw, **fkw[f])[2].items():
                    q["kernels_ms"][k] = q["kernels_ms"].get(k, 0.0) + v
            d = us - up
            q["errs"].append(float(np.max(np.abs(d)) / np.max(np.abs(up))))
            q["rms"].append(float(np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(up * up))))
            q["fits"].append(fit)
            msg += f" {f} {ds:.3f}s err {q['errs'][-1]:.2e}"
        print(msg, file=sys.stderr, flush=True)
    out_forms = {}
    for f in forms:
        q = res[f]
        o = {"form": FORMS[f], "value": tot_nodes * K / q["window"], "window_seconds": q["window"],
             "window_seconds_with_events": q["window_events"],
             "value_whole_run": tot_nodes * T / q["seconds"], "seconds_whole_run": q["seconds"],
             "speedup_vs_plain_fp64": tw_plain / q["window"],
             "speedup_vs_plain_fp64_whole_run": t_plain / q["seconds"],
             "max_rel_err_vs_fp64": max(q["errs"]), "rms_rel_err_vs_fp64": max(q["rms"]),
             "error_at_step": T, "sweep_shapes_rows_cols": q["shapes"]}
        if q["kernels_ms"]:
            o["kernels_ms_whole_run"] = q["kernels_ms"]
        if f == "soe":
            if soe_f32:
                o["form"] = "sum-of-exponentials history (FP32 mode states)"
            o["terms"] = cfg["soe_terms"]
            o["states"] = "fp32" if soe_f32 else "fp64"
            o["max_fit_rel_err"] = max(q["fits"])
        out_forms[f] = o
    ok = [f for f in forms if np.isfinite(out_forms[f]["max_rel_err_vs_fp64"])
          and out_forms[f]["max_rel_err_vs_fp64"] <= 1e-6
          and out_forms[f]["rms_rel_err_vs_fp64"] <= 1e-6]
    read_gbs = None if args.no_copy_probe else ctx.read_bandwidth()
    pw_gbs = plain_win["bytes"] / (plain_win["ms"] * 1e-3) / 1e9 if plain_win["ms"] else 0.0
    ratio, src = _stored_ratio(f"memterm_lod2d_{'1024' if cfg_name == 'c2' else 'ab'}_traffic.json")
    plain = {"form": "plain FP64 full history (the drop-in)",
             "value": tot_nodes * K / tw_plain, "window_seconds": tw_plain,
             "window_seconds_with_events": tw_plain_ev,
             "value_whole_run": tot_nodes * T / t_plain, "seconds_whole_run": t_plain,
             "kernels_ms_whole_run": kt_plain or None,
             "roofline": {"bound": "hbm", "achieved": pw_gbs, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": pw_gbs / HBM_PEAK_GBS,
                          "frac_of_measured_read": pw_gbs / read_gbs if read_gbs else None,
                          "kernel": "k_memterm_f64<2>", "launches": plain_win["launches"],
                          "avg_launch_ms": plain_win["ms"] / max(plain_win["launches"], 1),
                          "bytes_per_launch": plain_win["bytes"] / max(plain_win["launches"], 1),
                          "traffic": ratio * plain_win["bytes"] / plain_win["launches"]
                          if ratio and plain_win["launches"] else None,
                          "traffic_source": f"stored PMC ratio ({src}) x algorithmic bytes"
                          if ratio else None}}
    if not ok:
        # no form meets the bar: the line is the plain path
        best, chosen = "plain_fp64", plain
    else:
        best = max(ok, key=lambda f: out_forms[f]["value"])
        chosen = out_forms[best]
    line = {"metric": METRIC, "value": chosen["value"], "unit": "node-steps/s",
            "n_gpus": 1, "steps": K, "warmup": W,
            "ms_per_step": chosen["window_seconds"] * 1e3 / K, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None,
            "dtype": {"mfma": "bf16x2 split MFMA (fp32 acc) far history, fp64 near/sweeps",
                      "soe": "fp64 sweeps; " + ("fp32" if soe_f32 else "fp64") + " SOE states",
                      "f32c": "f32 compensated history, fp64 sweeps",
                      "binomial": "f64", "plain_fp64": "f64"}[best],
            "data": "synthetic",
            "config": {"workload": cfg["name"], "form": best, "sets": len(runs),
                       "nodes": tot_nodes, "tau": cfg["tau"], "steps_total": T,
                       "window": {"start_step": start, "steps": K, "warmup_steps": W,
                                  "history_depth": [start, start + K - 1],
                                  "note": "the K timed steps lie at history depth T/2 of "
                                          "whole runs from t = 0, starting on a 64-step block "
                                          "boundary; the MFMA form's far product is formed "
                                          "in node slices of about one round of workgroups "
                                          "(two per CU), one every 64 / slices steps; the "
                                          "rate is timed on a run without events, the "
                                          "roofline's launch durations on the same window of "
                                          "the whole run, with events",
                                  "mfma_far_slices": res["mfma"]["far"]["slices"]
                                  if "mfma" in res else None},
                       "parallelism": "1 GPU"},
            "chosen_form": dict(chosen, name=best),
            "forms": out_forms,
            "plain_fp64": plain,
            "speedup_vs_plain_fp64": tw_plain / chosen["window_seconds"],
            "measured_read_GBs": read_gbs}
    if best == "mfma" and res["mfma"]["far"]["launches"]:
        far = res["mfma"]["far"]
        a = far["bytes"] / (far["ms"] * 1e-3) / 1e9
        ratio, src = _stored_ratio(f"blockfar_lod2d_{'1024' if cfg_name == 'c2' else 'ab'}"
                                   "_traffic.json")
        nprod = 6 if cfg_name == "c4" else 3
        line["roofline"] = {
            "bound": "hbm", "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": a / HBM_PEAK_GBS,
            "frac_of_measured_read": a / read_gbs if read_gbs else None,
            "traffic": ratio * far["bytes"] / far["launches"] if ratio else None,
            "traffic_source": f"stored PMC ratio ({src}) x algorithmic bytes" if ratio else None,
            "kernel": f"k_block_far<{nprod}> (far product of the MFMA form, one node slice per "
                      f"launch: {sorted(set(far['slices']))} slices per block, one every 64 / "
                      "slices steps)",
            "launches": far["launches"], "avg_launch_ms": far["ms"] / far["launches"],
            "bytes_per_launch": far["bytes"] / far["launches"],
            "bytes_model": "4 B x settled slots x slice nodes read + 4 B x 128 rows x slice "
                           "nodes written",
            "useful_tflops": far["flops"] / (far["ms"] * 1e-3) / 1e12,
            "split_products": nprod,
            "mfma_tflops_bf16": nprod * far["flops"] / (far["ms"] * 1e-3) / 1e12}
    elif best == "soe" and res["soe"]["mem"]["launches"]:
        mem = res["soe"]["mem"]
        a = mem["bytes"] / (mem["ms"] * 1e-3) / 1e9
        line["roofline"] = {
            "bound": "hbm", "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": a / HBM_PEAK_GBS, "frac_of_measured_read": a / read_gbs if read_gbs else None,
            "traffic": None, "kernel": "k_soe_memterm (mode recurrences of the SOE form)",
            "launches": mem["launches"], "avg_launch_ms": mem["ms"] / mem["launches"],
            "bytes_per_launch": mem["bytes"] / mem["launches"],
            "bytes_model": "mode planes read and written (8 or 4 B x modes x 2) + ring and "
                           "partials, per node"}
    else:
        line["roofline"] = plain["roofline"]
    if args.parity_steps > 0:
        # the plain FP64 path against the oracle on the first steps of the
        # first set (the AB sets with the reference's MF_int_choice weights)
        from oracle import oracle as O
        from theses_code_amd.solvers import TGSolver

        r0 = runs[0]
        sN, M, coefs, C0 = r0["case"]
        P = args.parity_steps
        ab = (50, 1000, 1e-10) if cfg_name == "c4" else None
        o = O.TGSolver(sN, M, r0["alpha"], cfg["tau"], cfg["sigma"], 0.0, 0.0, *coefs, C0, ab=ab)
        o.step(P)
        ref = o.state()
        del o
        s = TGSolver(ctx, sN, M, r0["alpha"], cfg["tau"], cfg["sigma"], coefs, C0, capacity=P,
                     C1=0.0, S=0.0, kernel=kw["kernel"])
        s.calc_step(P)
        d = s.b_C - ref
        s.close()
        line["parity"] = {"vs": "oracle (CPU port of fr_d_t_g_2d.cpp, pinned to reference "
                                "goldens), plain FP64 path, first set",
                          "steps": P, "max_rel_err": float(np.max(np.abs(d)) / np.max(np.abs(ref))),
                          "rms_rel_err": float(np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(ref * ref)))}
    if args.cpu_steps > 0:
        # the oracle on a window at depth T/2 of the first set (the AB sets
        # with the reference's MF_int_choice weights)
        r0 = runs[0]
        c = dict(alpha=r0["alpha"], tau=cfg["tau"], sigma=cfg["sigma"])
        line["cpu_baseline"] = cpu_baseline(r0["case"], c, depth=start,
                                            nsteps=2 if cfg_name == "c2" else 1,
                                            ab=(50, 1000, 1e-10) if cfg_name == "c4" else None,
                                            steps_1thread=1 if args.cpu_1thread else 0)
    print(json.dumps(line))
    ctx.close()


def run_timesum(args):
    """Time sums of the 3D irrigation solver (the Sum family): plain FP64
    full-history sums (sum_alg 0), the compensated FP32 store as the chosen
    form, and the reference's series form (sum_alg 2).  A step is one
    push of the newest level plus the sums of every node, with the weights
    (kb_3_cache, tau) handed over from the host as the reference's caller
    does; the levels are device-resident (a pool of distinct levels pushed in
    turn)."""
    from oracle import oracle as O
    from theses_code_amd import _native as N
    from theses_code_amd import workloads as Wl
    from theses_code_amd.solvers import TimeSum

    cfg = Wl.TIMESUM3D_WORKLOAD
    n = (cfg["sN"] + 2) * (cfg["M"] + 2) * (cfg["K"] + 2)
    K = args.steps if args.steps is not None else cfg["steps"]
    tau, alpha = cfg["tau"], cfg["alpha"]
    T = np.arange(K + 2) * tau
    taus = np.full(K + 1, tau)
    kbs = [Wl.caputo_interval_weights(T[:k + 1], T[k + 1], alpha) for k in range(K)]
    # series form: the moments of x^j over the newest interval [T_{k-1}, T_k],
    # over T_k^(j+1), as its coefficients (the caller's, prepared like kbs)
    j = np.arange(cfg["series_terms"])
    ser_coefs = [np.zeros(j.size)] + [(1.0 - (T[k - 1] / T[k]) ** (j + 1)) / (j + 1)
                                      for k in range(1, K)]
    ctx = N.Context(0)
    pool_h = Wl.timesum3d_levels(n, cfg["pool"])
    pool = [N.DeviceArray.from_host(ctx, a) for a in pool_h]

    def run(precision, sum_alg=0, timing=False):
        """Wall clock without per-launch events; with `timing`, a second,
        event-timed run gives the kernel times."""
        dt, _, out = run1(precision, sum_alg, False)
        kt = run1(precision, sum_alg, True)[1] if timing else None
        return dt, kt, out

    def run1(precision, sum_alg, timing):
        ts = TimeSum(ctx, n, sum_alg=sum_alg, capacity=K + 1, precision=precision,
                     sum_param=cfg["series_terms"] if sum_alg == 2 else 0.0)
        # warm up the kernels once, then restart from an empty store
        ts.push(pool[0])
        ts.push(pool[1])
        if sum_alg == 2:
            ts.eval_device(coefs=np.ones(cfg["series_terms"]), gtj=2 * tau, gt=tau,
                           alpha=alpha, tau_n=tau)
        else:
            ts.eval_device(kb_3_cache=kbs[1], tau=taus)
        ts.clear()
        if timing:
            ctx.timing_reset()
            ctx.timing_enable(1 << N.K_MEMTERM)
        ctx.sync()
        t0 = time.perf_counter()
        for k in range(K):
            ts.push(pool[k % len(pool)])
            if sum_alg == 2:
                ts.eval_device(coefs=ser_coefs[k], gtj=T[k + 1], gt=T[k], alpha=alpha,
                               tau_n=tau)
            else:
                ts.eval_device(kb_3_cache=kbs[k], tau=taus)
        ctx.sync()
        dt = time.perf_counter() - t0
        kt = ctx.timing(N.K_MEMTERM) if timing else None
        if timing:
            ctx.timing_enable(0)
        S = ts.eval_device(kb_3_cache=kbs[K - 1], tau=taus) if sum_alg != 2 else ts._S
        out = S.to_host()
        ts.close()
        return dt, kt, out

    dt, kt, S64 = run(N.PREC_F64, timing=True)
    launches, ms, nbytes = kt
    # HBM bytes per launch: the PMC-measured ratio of counted to algorithmic
    # bytes of this kernel (profiles/memterm_timesum3d_traffic.json) times
    # this run's algorithmic bytes per launch
    traffic = None
    tf = os.path.join(REPO, "profiles", "memterm_timesum3d_traffic.json")
    if os.path.exists(tf) and launches:
        with open(tf) as f:
            traffic = json.load(f)["traffic_over_algorithmic"] * nbytes / launches
    value = n * K / dt
    achieved = nbytes / (ms * 1e-3) / 1e9 if ms else 0.0
    # parity of the last step against the oracle on a node sample, all levels
    rng = Xoshiro(cfg["seed"])
    sub = np.sort(rng.choice(n, 4096))
    old_sub = np.stack([pool_h[k % len(pool_h)][sub] for k in range(K)])
    ref = O.timesum01(0, 0.0, old_sub, kbs[K - 1], taus)
    par = float(np.max(np.abs(S64[sub] - ref)) / np.max(np.abs(ref)))
    dt32, kt32, S32 = run(N.PREC_F32C, timing=True)
    d = S32 - S64
    chosen = {"form": "FP32 level differences, Kahan-compensated FP32 sums (FRM_PREC_F32C)",
              "value": n * K / dt32, "seconds": dt32, "speedup_vs_plain_fp64": dt / dt32,
              "max_rel_err_vs_fp64": float(np.max(np.abs(d)) / np.max(np.abs(S64))),
              "rms_rel_err_vs_fp64": float(np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(S64 ** 2))),
              "kernel_ms": kt32[1],
              "kernel_GBs": kt32[2] / (kt32[1] * 1e-3) / 1e9 if kt32[1] else 0.0}
    dts, kts, _ = run(N.PREC_F64, sum_alg=2, timing=True)
    series = {"form": f"the reference's series sums (sum_alg 2), {cfg['series_terms']} moments "
                      "per node; its accuracy is the method's and the caller's coefficients'",
              "value": n * K / dts, "seconds": dts, "kernel_ms": kts[1],
              "kernel_GBs": kts[2] / (kts[1] * 1e-3) / 1e9 if kts[1] else 0.0}
    for a in pool:
        a.free()
    del old_sub
    # CPU: the oracle's time_summing on a node sample at the run's mean depth
    cpu = None
    if args.cpu_steps > 0:
        ns, depth = 65536, max(K // 2, 1)
        old_cpu = np.stack([pool_h[k % len(pool_h)][:ns] for k in range(depth + 1)])
        kb = Wl.caputo_interval_weights(T[:depth + 1], T[depth + 1], alpha)
        O.timesum01(0, 0.0, old_cpu[:8], kb, taus)
        reps, t0 = 0, time.perf_counter()
        while reps < 3 or time.perf_counter() - t0 < 5.0:
            O.timesum01(0, 0.0, old_cpu, kb, taus)
            reps += 1
        dcpu = (time.perf_counter() - t0) / reps
        lib = O.lib()
        cpu = {"value": ns / dcpu, "unit": "node-steps/s",
               "cores": lib.orc_get_threads() if hasattr(lib, "orc_get_threads") else 1,
               "kind": "port",
               "sample": f"oracle time_summing (sum_alg 0) on {ns} nodes at depth {depth} "
                         f"(the run's mean), {reps} repetitions; OpenMP over nodes"}
    line = {"metric": METRIC, "value": value, "unit": "node-steps/s", "n_gpus": 1, "steps": K,
            "warmup": 1, "ms_per_step": dt * 1e3 / K, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": cfg["name"], "nodes": n, "tau": tau, "alpha": alpha,
                       "sum_alg": 0, "parallelism": "1 GPU"},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                         "kernel": "k_memterm_f64<1>", "launches": launches,
                         "avg_launch_ms": ms / launches if launches else None},
            "parity": {"vs": "oracle time_summing restatement, last step, 4096-node sample",
                       "max_rel_err": par},
            "chosen_form": chosen, "series_form": series, "cpu_baseline": cpu}
    print(json.dumps(line))
    ctx.close()


def run_soe_step(args):
    """The standalone sum-of-exponentials history (frm_soe_step) at the node
    count of the 16 AB parameter sets (16 x 2.45e6 nodes): 64 modes fitted to
    the L1 Caputo weights, one target; a step folds one
    device-resident increment and evaluates the sum at every node.  The line
    is the FP64-state form; FP32 states on the same inputs are reported with
    their error (random signed increments cancel against the slow modes'
    large states, so that error is well above the FP32 rounding unit); a node
    sample is checked against the direct history sum with the fitted
    weights."""
    from theses_code_amd import _native as N
    from theses_code_amd.solvers import SoeHistory

    n, L, alpha = 16 * 2_450_000, 64, 0.6
    K = args.steps if args.steps is not None else 200
    lags = np.arange(1, 4001, dtype=np.float64)
    q, c, fit_rel = N.host_soe_fit(lags ** (1 - alpha) - (lags - 1) ** (1 - alpha), 1, L, 1e-3)
    ctx = N.Context(0)
    rng = Xoshiro(0x5EED0004)
    pool_h = [rng.uniform(-1.0, 1.0, n) for _ in range(4)]
    pool = [N.DeviceArray.from_host(ctx, a) for a in pool_h]

    def run(f32, timing):
        s = SoeHistory(ctx, n, L, 1, f32=f32)
        s.step_device(pool[0], q, c)  # warm-up, then a fresh run
        s.clear()
        if timing:
            ctx.timing_reset()
            ctx.timing_enable(1 << N.K_MEMTERM)
        ctx.sync()
        t0 = time.perf_counter()
        for k in range(K):
            out = s.step_device(pool[k % len(pool)], q, c)
        ctx.sync()
        dt = time.perf_counter() - t0
        kt = None
        if timing:
            kt = ctx.timing(N.K_MEMTERM)
            ctx.timing_enable(0)
        res = out.to_host()
        s.close()
        return dt, kt, res

    dt64, _, o64 = run(False, False)
    _, kt, _ = run(False, True)
    dt32, _, o32 = run(True, False)
    launches, ms, nbytes = kt
    achieved = nbytes / (ms * 1e-3) / 1e9 if ms else 0.0
    traffic = None
    tf = os.path.join(REPO, "profiles", "soe_step_traffic.json")
    if os.path.exists(tf) and launches:
        # PMC-measured ratio of HBM to algorithmic bytes of this kernel
        with open(tf) as f:
            traffic = json.load(f)["traffic_over_algorithmic"] * nbytes / launches
    d = o32 - o64
    sub = np.sort(rng.choice(n, 4096))
    w = c[0] @ (q[:, None] ** np.arange(K)[None, :])
    Dsub = np.stack([pool_h[k % len(pool_h)][sub] for k in range(K)])
    ref = w[::-1] @ Dsub
    par = float(np.max(np.abs(o64[sub] - ref)) / np.max(np.abs(ref)))
    for a in pool:
        a.free()
    line = {"metric": METRIC, "value": n * K / dt64, "unit": "node-steps/s", "n_gpus": 1,
            "steps": K, "warmup": 1, "ms_per_step": dt64 * 1e3 / K, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": "soe_history_39.2M_nodes_64_modes", "nodes": n, "modes": L,
                       "alpha": alpha, "fit_max_rel": fit_rel, "parallelism": "1 GPU"},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                         "kernel": "k_soe_step<double, 1, true>", "launches": launches,
                         "avg_launch_ms": ms / launches if launches else None,
                         "bytes_per_launch": nbytes / launches if launches else None,
                         "measured_copy_GBs": None if args.no_copy_probe else ctx.copy_bandwidth()},
            "f32_states": {"value": n * K / dt32, "seconds": dt32,
                           "max_rel_err_vs_fp64": float(np.max(np.abs(d)) / np.max(np.abs(o64))),
                           "rms_rel_err_vs_fp64": float(np.sqrt(np.mean(d * d)) /
                                                        np.sqrt(np.mean(o64 ** 2)))},
            "parity": {"vs": "direct history sum with the fitted weights, 4096-node sample, "
                             "FP64 states", "max_rel_err": par},
            "cpu_baseline": None}
    print(json.dumps(line))
    ctx.close()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the history sums shard by node with no exchange)", file=sys.stderr)
        sys.exit(2)
    if args.config in ("c0", "c3"):
        run_mt1d(args, args.config)
        return
    if args.config in ("c2", "c4"):
        run_lod_forms(args, args.config)
        return
    if args.config in ("ab1d", "ab1d_series"):
        run_ab1d(args, series=args.config == "ab1d_series")
        return
    if args.config == "timesum3d":
        run_timesum(args)
        return
    if args.config == "soe_step":
        run_soe_step(args)
        return

    from theses_code_amd import _native as N
    from theses_code_amd.solvers import TGSolver
    from theses_code_amd.workloads import C1_WORKLOAD, lod2d_case

    cfg = C1_WORKLOAD
    case = lod2d_case(cfg["n_interior"], cfg["seed"], vmax=cfg["vmax"], D=cfg["D"])
    sN, M, coefs, C0 = case
    nodes = (sN - 1) * (M - 1)
    K = args.steps if args.steps is not None else cfg["steps"]
    W = args.warmup
    ctx = N.Context(0)
    cap = max(K, W, args.parity_steps, 1)
    s = TGSolver(ctx, sN, M, cfg["alpha"], cfg["tau"], cfg["sigma"], coefs, C0, capacity=cap,
                 C1=0.0, S=0.0)

    # parity against the oracle on the first steps of the same problem
    parity = None
    if args.parity_steps > 0:
        from oracle import oracle as O
        o = O.TGSolver(sN, M, cfg["alpha"], cfg["tau"], cfg["sigma"], 0.0, 0.0, *coefs, C0)
        o.step(args.parity_steps)
        s.calc_step(args.parity_steps)
        ref = o.state()
        got = s.b_C
        d = got - ref
        parity = {"vs": "oracle (CPU port of fr_d_t_g_2d.cpp, pinned to reference goldens)",
                  "steps": args.parity_steps,
                  "max_rel_err": float(np.max(np.abs(d)) / np.max(np.abs(ref))),
                  "rms_rel_err": float(np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(ref * ref)))}
        del o

    # warmup, then reset to t = 0
    s.set_state(C0)
    if W:
        s.calc_step(W)
        ctx.sync()
    s.set_state(C0)

    ctx.timing_reset()
    ctx.timing_enable(1 << N.K_MEMTERM)
    ctx.sync()
    t0 = time.perf_counter()
    s.calc_step(K)
    ctx.sync()
    t1 = time.perf_counter()
    ctx.timing_enable(0)
    launches, ms, nbytes = ctx.timing(N.K_MEMTERM)
    final = s.b_C
    dt = t1 - t0
    value = nodes * K / dt

    achieved = (nbytes / (ms * 1e-3) / 1e9) if ms > 0 else 0.0
    # HBM bytes per launch: the PMC-measured ratio of counted to algorithmic
    # bytes of this kernel (profiles/memterm_lod2d_512_traffic.md) times this
    # run's algorithmic bytes per launch
    traffic = None
    tf = os.path.join(REPO, "profiles", "memterm_lod2d_512_traffic.json")
    if os.path.exists(tf) and launches:
        with open(tf) as f:
            traffic = json.load(f)["traffic_over_algorithmic"] * nbytes / launches
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "kernel": "k_memterm_f64<2>", "launches": launches,
                "avg_launch_ms": ms / launches if launches else None,
                "bytes_per_launch_avg": nbytes / launches if launches else None}

    # the chosen forms on the same workload and inputs, against this plain run
    chosen = {}
    if args.forms:
        s.close()
        for f in args.forms.split(","):
            fk = {"mfma": dict(precision=N.PREC_MFMA_BF16), "soe": dict(precision=N.PREC_SOE),
                  "f32c": dict(precision=N.PREC_F32C),
                  "binomial": dict(precision=N.PREC_BINOMIAL, series_terms=BINOMIAL_TERMS)}[f]
            sf = TGSolver(ctx, sN, M, cfg["alpha"], cfg["tau"], cfg["sigma"], coefs, C0,
                          capacity=cap, C1=0.0, S=0.0, **fk)
            if W:
                sf.calc_step(W)
                ctx.sync()
                sf.set_state(C0)
            ctx.sync()
            t0 = time.perf_counter()
            sf.calc_step(K)
            ctx.sync()
            dtf = time.perf_counter() - t0
            d = sf.b_C - final
            chosen[f] = {"form": FORMS[f], "value": nodes * K / dtf, "seconds": dtf,
                         "speedup_vs_plain_fp64": dt / dtf,
                         "max_rel_err_vs_fp64": float(np.max(np.abs(d)) / np.max(np.abs(final))),
                         "rms_rel_err_vs_fp64": float(np.sqrt(np.mean(d * d)) /
                                                      np.sqrt(np.mean(final * final)))}
            sf.close()
    if not args.no_copy_probe:
        roofline["measured_copy_GBs"] = ctx.copy_bandwidth()
    cpu = None
    if args.cpu_steps > 0:
        cpu = cpu_baseline(case, cfg, depth=max(K // 2, 1), nsteps=args.cpu_steps)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "node-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": dt * 1e3 / max(K, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": cfg["name"], "form": "plain FP64 GPU path (drop-in)",
                   "grid_interior": f"{sN - 1}x{M - 1}", "alpha": cfg["alpha"],
                   "tau": cfg["tau"], "history": "full, FP64", "D": cfg["D"],
                   "vmax": cfg["vmax"], "nodes": nodes, "seed": cfg["seed"],
                   "steps_total": K, "history_depth": [0, K - 1],
                   "parallelism": "1 GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
        "final_field_checksum": float(np.sum(final)),
    }
    if chosen:
        line["chosen_forms_same_workload"] = chosen
    print(json.dumps(line))
    s.close()
    ctx.close()


if __name__ == "__main__":
    main()
