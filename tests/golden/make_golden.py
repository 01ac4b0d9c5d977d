"""Write the golden vectors under tests/golden/ from the reference itself.

Runs the harnesses that oracle/ref/Makefile compiles from the reference's own
sources (fr_d_t_g_2d.cpp, mt_minimal.cpp, mt_psi_caputo.cpp,
MF_int_series.cpp) into oracle/_ref/, feeds them seeded inputs and stores
inputs and outputs as .npz fixtures.  Only the fixtures are committed; this
script needs /root/reference and runs in the build container only:

    make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref")


def run(binary: str, inp: list | np.ndarray) -> np.ndarray:
    with tempfile.TemporaryDirectory() as d:
        fi, fo = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        np.asarray(inp, dtype=np.float64).tofile(fi)
        r = subprocess.run([os.path.join(REF, binary), fi, fo], capture_output=True, text=True,
                           env={**os.environ, "OMP_NUM_THREADS": "1"})
        if r.returncode != 0:
            raise RuntimeError(f"{binary} failed ({r.returncode}): {r.stderr[-500:]}")
        return np.fromfile(fo)


def save(name: str, meta: dict, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, meta=np.array(json.dumps(meta)), **arrays)
    print(f"wrote {name}.npz", {k: v.shape for k, v in arrays.items()})


# ---------------------------------------------------------------- 2D LOD
def tg_case(name, seed, func, sN, M, alpha, tau, sigma, dv, C1, S, nsteps, every, fpow=1.0,
            dt_alg=0, dt_eps=1e-8):
    rng = np.random.default_rng(seed)
    nn = (sN + 2) * (M + 2)
    V2 = rng.uniform(0.5, 1.5, nn)
    C0 = rng.uniform(0.0, 1.0, nn)
    int_alg = 6 if dt_alg == 3 else 1
    hdr = [func, sN, M, alpha, tau, sigma, dv, C1, S, nsteps, every, int_alg, fpow, dt_alg, dt_eps]
    out = run("ref_tg", np.concatenate([np.array(hdr, float), V2, C0]))
    nsnap = len([k for k in range(1, nsteps + 1) if k % every == 0 or k == nsteps])
    snaps = out[: nsnap * nn].reshape(nsnap, M + 2, sN + 2)
    kb = out[nsnap * nn:].reshape(8, nsteps + 1)
    meta = dict(source=f"fr_d_t_g_2d.cpp TG_solver_<func,1,0>::calc_step, dt_alg={dt_alg}, "
                       f"int_alg={int_alg}",
                func=func, sN=sN, M=M, alpha=alpha, tau=tau, sigma=sigma, dv=dv, C1=C1, S=S,
                nsteps=nsteps, every=every, fpow=fpow, int_err=1e-6, seed=seed)
    if dt_alg:
        meta.update(dt_alg=dt_alg, dt_eps=dt_eps)
    save(name, meta, V2=V2.reshape(M + 2, sN + 2), C0=C0.reshape(M + 2, sN + 2), snaps=snaps,
         kb=kb)


# ---------------------------------------------------------------- fitting input reader
def fitio_case(name, files, args):
    """Runs the reference's main (mt_minimal.cpp) in fitting mode on the given
    input files (their text is stored with the fixture) and stores what its
    reader (1955-2043) handed to fit_and_solve."""
    with tempfile.TemporaryDirectory() as d:
        for fn, text in files.items():
            with open(os.path.join(d, fn), "w") as f:
                f.write(text)
        fo = os.path.join(d, "out.bin")
        argv = [os.path.join(REF, "ref_fitio"), fo, "Fit", "1", "BS", "2", "NB", "2", "Tm", "1"]
        for a in args:
            argv.append(os.path.join(d, a[1:]) if a.startswith("@") else a)
        r = subprocess.run(argv, capture_output=True, text=True, cwd=d)
        if r.returncode != 0:
            raise RuntimeError(f"ref_fitio failed ({r.returncode}): {r.stderr[-500:]}")
        out = np.fromfile(fo)
    meta = dict(source="mt_minimal.cpp main, fitting input reader (1955-2043)", args=args,
                files=sorted(files))
    save(name, meta, parsed=out, **{"file_" + k.replace(".", "_"): np.array(v)
                                    for k, v in files.items()})


def fitio_all():
    rng = np.random.default_rng(31)

    def rows(n, c, lo=-5, hi=5, trail=""):
        a = rng.uniform(lo, hi, (n, c)).round(7)
        return "".join(" ".join("%.17g" % x for x in r) + trail + "\n" for r in a)

    head = "1 3 2 2 2 1 4 2 0.25\n30 0.7 1.4 1.6 1e-6 200\n"
    body = rows(3, 3) + rows(2, 2) + rows(2, 2) + rows(1, 2) + rows(2, 3)
    ev = rows(4, 4, 0, 1e-6, " ")
    # one file, 25 parameter values (the reader takes 20), default H_mult 10
    fitio_case("fitio_one_file", {"fit_input.txt": head + body + ev +
                                  " ".join("%.10g" % x for x in rng.uniform(0, 1, 25)) + "\n"},
               ["fi_file", "@fit_input.txt"])
    # evapotranspiration and parameters from their own files, H_mult 1
    fitio_case("fitio_files", {"fi.txt": head + body, "et.txt": ev,
                               "pv.txt": " ".join("%.10g" % x for x in rng.uniform(0, 1, 7))},
               ["fi_file", "@fi.txt", "et_file", "@et.txt", "pv_file", "@pv.txt", "H_mult", "1"])
    # empty sections and no parameters (one evapotranspiration column: the
    # reader prints EV_F[0][0], so the reference needs nEV_F >= 1)
    fitio_case("fitio_empty", {"fit_input.txt": "0 2 1 0 0 0 2 1 1\n5 0.5 1 1 0.001 3\n" +
                               rows(2, 3) + rows(1, 2) + rows(2, 3, 0, 1)},
               ["fi_file", "@fit_input.txt"])
    # a file that ends inside the values to fit: the reader gives up
    fitio_case("fitio_short", {"fit_input.txt": head + rows(2, 3)}, ["fi_file", "@fit_input.txt"])
    # exactly 19 parameter values: the reference's counter reports 20
    fitio_case("fitio_19", {"fit_input.txt": head + body + ev +
                            " ".join("%.10g" % x for x in rng.uniform(0, 1, 19)) + "\n"},
               ["fi_file", "@fit_input.txt"])


# ---------------------------------------------------------------- 1D H_solver
def mt_traj(name, binary, seed, BS, NB, alpha, tau, L, k, Ev, Hb, bc, nsteps, every, U0fn,
            beta=None):
    rng = np.random.default_rng(seed)
    N = BS * NB
    x = np.arange(N + 2) / N
    U0 = U0fn(x, rng)
    U0[N + 1] = U0[N]
    hdr = [0, BS, NB, alpha, tau, L, k, Ev, Hb, bc, nsteps, every]
    if beta is not None:
        hdr.append(beta)
    out = run(binary, np.concatenate([np.array(hdr, float), U0]))
    nsnap = len([t for t in range(1, nsteps + 1) if t % every == 0 or t == nsteps])
    snaps = out[: nsnap * (N + 2)].reshape(nsnap, N + 2)
    arrays = dict(U0=U0, snaps=snaps)
    if beta is not None:
        arrays["kb_time"] = out[nsnap * (N + 2):]
    meta = dict(source=f"{binary}: H_solver::calc_step, cv_form=2, k_form=0, gamma=1",
                BS=BS, NB=NB, N=N, alpha=alpha, tau=tau, L=L, k=k, Ev=Ev, Hbottom=Hb,
                bottom_cond=bc, nsteps=nsteps, every=every, beta=beta, seed=seed)
    save(name, meta, **arrays)


def mt_hist(name, binary, seed, BS, NB, alpha, tau, os_, tstep, beta=None):
    rng = np.random.default_rng(seed)
    N = BS * NB
    snaps = np.cumsum(rng.normal(0, 1, (os_, N + 2)), axis=0)
    hdr = [1, BS, NB, alpha, tau, os_, tstep]
    if beta is not None:
        hdr.append(beta)
    out = run(binary, np.concatenate([np.array(hdr, float), snaps.reshape(-1)]))
    meta = dict(source=f"{binary}: H_solver::Om1 time-history sum with U=0", BS=BS, NB=NB,
                N=N, alpha=alpha, tau=tau, os=os_, tstep=tstep, beta=beta, seed=seed)
    save(name, meta, snaps=snaps, Da=out[:1], Om=out[1:])


def psi_weights(name, cases, As):
    tabs = {}
    for q, (alpha, beta, tau) in enumerate(cases):
        out = run("ref_psi", [2, alpha, beta, tau, len(As)] + list(As))
        tabs[f"w{q}"] = out
    meta = dict(source="mt_psi_caputo.cpp calc_int(alpha,a,a+1,A,tau,1), g=t^beta",
                cases=cases, As=list(As))
    save(name, meta, **tabs)


def psi_kb_quad(name):
    """calc_int's quadrature fallback _kb(1, t0, t1, gtj, sing, alpha) alone
    (harness mode 3): intervals away from the target and ending at it
    (sing), g = t^beta for beta 1 (closed form known), 0.8, 1.4 and 80."""
    cases = []
    for alpha in (0.05, 0.3, 0.6, 0.9):
        for beta in (1.0, 0.8, 1.4):
            for t0, t1, tj, sing in ((0.0, 0.01, 0.02, 0), (0.3, 0.4, 0.4, 1), (0.0, 1.0, 1.0, 1),
                                     (0.2, 0.25, 0.9, 0), (1.5, 2.0, 2.0, 1)):
                cases.append((alpha, beta, t0, t1, tj ** beta, sing))
        for t0, t1, tj, sing in ((0.0, 0.01, 0.5, 0), (0.49, 0.5, 0.5, 1), (0.0, 0.01, 0.03, 0)):
            cases.append((alpha, 80.0, t0, t1, tj ** 80.0, sing))
    cases = np.array(cases)
    out = run("ref_psi", [3, len(cases), *cases.reshape(-1)])
    save(name, dict(source="mt_psi_caputo.cpp _kb(1, t0, t1, gtj, sing, alpha)",
                    columns=["alpha", "beta", "t0", "t1", "gtj", "sing"]), cases=cases, kb=out)


# ---------------------------------------------------------------- AB
def mf_all(name):
    arrays, meta = {}, dict(source="MF_int_series.cpp")
    wcases = [(0.5, 0.01, 50, 1000, 1e-10, 120), (0.8, 0.01, 50, 1000, 1e-10, 60),
              (0.3, 0.05, 50, 1000, 1e-10, 40)]
    for q, c in enumerate(wcases):
        arrays[f"w{q}"] = run("ref_mf", [0, *c])
    meta["weights"] = wcases
    rng = np.random.default_rng(77)
    hcases = [(0.6, 0.01, 50, 1000, 1e-10, 50)]
    for q, c in enumerate(hcases):
        fvs = rng.normal(0, 1, c[-1])
        arrays[f"fvs{q}"] = fvs
        arrays[f"h{q}"] = run("ref_mf", [1, *c, *fvs])
    meta["hist"] = hcases
    scases = [(0.5, 1.0, 60, 20, 1, 50, 1000, 1e-10, 0), (0.7, 1.0, 40, 16, 1, 50, 1000, 1e-10, 1),
              (0.3, 0.5, 40, 12, 1, 50, 1000, 1e-10, 2)]
    arrays["solve_err"] = np.array([run("ref_mf", [2, *c])[0] for c in scases])
    meta["solve"] = scases
    ml = []
    for a in (0.3, 0.5, 0.8, 1.0, 2.0):
        for b in (1.0, 0.5, 3.0):
            for z in (-5.0, -1.0, -0.3, 0.0, 0.2, 0.9):
                ml.append((a, b, z, 1e-12))
    arrays["ml_in"] = np.array(ml)
    arrays["ml_out"] = run("ref_mf", [3, len(ml)] + [v for e in ml for v in e])
    save(name, meta, **arrays)


def mf_series(name):
    """solve_and_check with simple = 0: the series history of MF_intF_choice
    per node (harness mode 4) and whole solves (mode 2), on the reference's
    own test configurations (convection_diffusion_test: maxl = 1000;
    convection_diffusion_test_time: 6000 steps, 100 intervals, maxl = 50)."""
    arrays, meta = {}, dict(source="MF_int_series.cpp solve_and_check, simple = 0")
    rng = np.random.default_rng(91)
    # (a, tau, maxl, maxkiter, eps, nmax): series only, series then integral
    # form (switch at i = 51 and i = 9), maxl = 1000
    hcases = [(0.3, 0.02, 50, 1000, 1e-10, 60), (0.6, 0.01, 50, 1000, 1e-10, 80),
              (0.9, 0.01, 50, 1000, 1e-10, 40), (0.5, 0.025, 1000, 1000, 1e-10, 41)]
    for q, c in enumerate(hcases):
        fvs = rng.normal(0, 1, c[-1] - 1)
        arrays[f"fvs{q}"] = fvs
        out = run("ref_mf", [4, *c, *fvs])
        arrays[f"h{q}"] = out[: c[-1] - 1]
        arrays[f"choice{q}"] = out[c[-1] - 1:]
    meta["hist"] = hcases
    # (a, T, n, m, simple, maxl, maxkiter, eps, F_func)
    scases = [(0.9, 1.0, 6000, 100, 0, 50, 1000, 1e-10, 0), (0.5, 1.0, 6000, 100, 0, 50, 1000, 1e-10, 0),
              (0.1, 1.0, 6000, 100, 0, 50, 1000, 1e-10, 0), (0.5, 1.0, 40, 40, 0, 1000, 1000, 1e-10, 1),
              (0.7, 1.0, 80, 24, 0, 50, 1000, 1e-10, 2)]
    arrays["solve_err"] = np.array([run("ref_mf", [2, *c])[0] for c in scases])
    meta["solve"] = scases
    save(name, meta, **arrays)


def mf_recount(name):
    """solve_and_check with simple = 0 where MF_intF's power-series terms
    overflow (T > 1, long series): the sum restarts from the sorted pyramid
    sum of the finite terms (MF_int_series.cpp:600-609)."""
    arrays, meta = {}, dict(source="MF_int_series.cpp solve_and_check, simple = 0, MF_intF recount")
    rng = np.random.default_rng(92)
    hcases = [(0.3, 0.05, 600, 1000, 1e-10, 80)]
    for q, c in enumerate(hcases):
        fvs = rng.normal(0, 1, c[-1] - 1)
        arrays[f"fvs{q}"] = fvs
        out = run("ref_mf", [4, *c, *fvs])
        arrays[f"h{q}"] = out[: c[-1] - 1]
        arrays[f"choice{q}"] = out[c[-1] - 1:]
    meta["hist"] = hcases
    scases = [(0.3, 4.0, 80, 30, 0, 600, 1000, 1e-10, 0), (0.3, 4.0, 80, 30, 0, 600, 1000, 1e-10, 2),
              (0.25, 3.0, 60, 20, 0, 700, 1000, 1e-10, 1)]
    arrays["solve_err"] = np.array([run("ref_mf", [2, *c])[0] for c in scases])
    meta["solve"] = scases
    save(name, meta, **arrays)


def main():
    if not os.path.isdir("/root/reference"):
        sys.exit("the reference sources are needed to write golden vectors")
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    tg_case("tg_a", 11, 0, 16, 12, 0.7, 1e-2, 0.3, 0.05, 1.0, 0.1, 40, 10)
    tg_case("tg_b", 12, 0, 9, 20, 0.4, 1e-3, 1.0, 0.2, 0.0, 0.0, 30, 15)
    tg_case("tg_psi", 13, 3, 10, 8, 0.6, 1e-2, 0.5, 0.1, 0.5, 0.0, 12, 6, fpow=0.7)
    # restricted summation (dt_alg 2): the newest-first history sum stops at
    # the first weight below dt_eps
    tg_case("tg_dt2_a", 16, 0, 12, 10, 0.7, 1e-2, 0.3, 0.05, 1.0, 0.1, 40, 10, dt_alg=2, dt_eps=1e-2)
    tg_case("tg_dt2_b", 17, 0, 9, 14, 0.5, 1e-3, 1.0, 0.2, 0.0, 0.0, 50, 25, dt_alg=2, dt_eps=4e-3)
    # binomial-series history (dt_alg 3), 25 and 10 terms
    tg_case("tg_dt3_a", 14, 0, 12, 10, 0.7, 1e-2, 0.3, 0.05, 1.0, 0.1, 30, 10, dt_alg=3, dt_eps=25)
    tg_case("tg_dt3_b", 15, 0, 9, 14, 0.5, 1e-3, 1.0, 0.2, 0.0, 0.0, 40, 20, dt_alg=3, dt_eps=10)
    # the ia/ibk series approximations of the history (dt_alg 4), across the
    # reference's every-20-steps compaction of its increment lists
    tg_case("tg_dt4_a", 18, 0, 12, 10, 0.7, 1e-2, 0.3, 0.05, 1.0, 0.1, 60, 20, dt_alg=4)
    tg_case("tg_dt4_b", 19, 0, 9, 14, 0.4, 1e-3, 1.0, 0.2, 0.0, 0.0, 90, 30, dt_alg=4)
    mt_traj("mt_flux", "ref_mt", 21, 10, 4, 0.8, 0.01, 1.0, 1e-3, 0.0, -20.0, 1, 60, 20,
            lambda x, r: -20.0 + 2.0 * np.sin(np.pi * x) + 0.1 * r.normal(size=x.size))
    mt_traj("mt_dir_cv2", "ref_mt", 22, 10, 3, 0.6, 0.02, 1.0, 2e-3, -0.5, 0.0, 0, 50, 25,
            lambda x, r: -12.0 + 10.0 * x + 0.2 * r.normal(size=x.size))
    mt_hist("mt_hist", "ref_mt", 23, 10, 3, 0.6, 0.01, 25, 26)
    mt_traj("psi_flux", "ref_psi", 24, 10, 2, 0.75, 0.01, 1.0, 1e-3, 0.0, -20.0, 1, 30, 15,
            lambda x, r: -20.0 + 2.0 * np.sin(np.pi * x) + 0.1 * r.normal(size=x.size),
            beta=0.8)
    mt_hist("psi_hist", "ref_psi", 25, 10, 2, 0.65, 0.01, 18, 19, beta=1.3)
    psi_weights("psi_w", [(0.6, 0.8, 1e-3), (0.9, 1.4, 1e-3), (0.55, 1.0, 1e-2)],
                [1, 2, 5, 17, 64, 200])
    # g = t^beta with large beta and small tau: g(A tau) is so small that
    # some weight series are not finite and calc_int falls back to _kb
    psi_kb_quad("psi_kb_quad")
    psi_weights("psi_w_quad", [(0.05, 80.0, 0.01), (0.3, 40.0, 1e-4), (0.6, 80.0, 1e-4)],
                [3, 10, 50])
    mt_traj("psi_quad_flux", "ref_psi", 26, 10, 2, 0.05, 0.01, 1.0, 1e-3, 0.0, -20.0, 1, 40, 20,
            lambda x, r: -20.0 + 2.0 * np.sin(np.pi * x) + 0.1 * r.normal(size=x.size),
            beta=80.0)
    mf_all("mf")
    mf_series("mf_series")
    mf_recount("mf_recount")
    fitio_all()


if __name__ == "__main__":
    main()
