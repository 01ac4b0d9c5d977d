"""frm::read_fit_input, frm::write_res_rows and frm::write_log_row (include/frm_fitio.hpp): the
fitting input format of mt_minimal.cpp:1955-2043 and the result rows of
1631-1635.  The reference ships no sample input, so the files here are
written from the format; the expected values come from splitting the text
into whitespace-separated tokens in the order the reference reads them.
Host code only (g++), no GPU."""
import os
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("fitio") / "fitio_check")
    # with the address and undefined-behaviour sanitizers: the reader parses
    # files of unknown length, malformed ones included
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-I" + os.path.join(REPO, "include"),
                    os.path.join(REPO, "tests", "cpp", "fitio_check.cpp"), "-o", out], check=True)
    return out


def run(checker, path, hmult, et="-", pv="-"):
    r = subprocess.run([checker, str(path), repr(hmult), str(et), str(pv)], capture_output=True,
                       text=True, check=True)
    lines = r.stdout.splitlines()
    assert lines[0].startswith("ok ")
    if lines[0] == "ok 0":
        return None, lines
    fields = {}
    evf = []
    for ln in lines[1:]:
        if ln.count(" ") >= 3 and ln[0].isdigit():
            break  # result rows
        name, *vals = ln.split(" ")
        if name == "EV_F":
            evf.append([float(v) for v in vals])
        else:
            fields[name] = [float(v) for v in vals]
    fields["EV_F"] = evf
    return fields, lines


def sample(rng, nvalues=4, ninit=3, ncheck=2, nperc=2, nirr=1, nEV_T=3, nEV_F=2, nparam=5,
           ev_inline=True, params_inline=True):
    vals = {k: rng.uniform(-5, 5, n * c).round(6).reshape(n, c) for k, n, c in
            [("values", nvalues, 3), ("init", ninit, 2), ("perc", nperc, 2), ("irr", nirr, 2),
             ("check", ncheck, 3), ("ev", nEV_T, 2 + nEV_F)]}
    params = rng.uniform(0, 1, nparam).round(8)
    head = f"1 {nvalues} {ninit} {ncheck} {nperc} {nirr} {nEV_T} {nEV_F} 0.25\n"
    head += "30 0.7 1.4 1.6 1e-6 200\n"
    body = ""
    for k in ("values", "init", "perc", "irr", "check"):
        for r in vals[k]:
            body += " ".join(repr(float(x)) for x in r) + "\n"
    ev = "".join(" ".join(repr(float(x)) for x in r) + " \n" for r in vals["ev"])
    pv = " ".join(repr(float(x)) for x in params) + "\n"
    text = head + body + (ev if ev_inline else "") + (pv if params_inline else "")
    return text, vals, params, ev, pv


def check_fields(f, vals, params, hmult):
    assert f["values_T"] == list(vals["values"][:, 0])
    assert f["values_Z"] == list(vals["values"][:, 1])
    assert f["values_F"] == list(vals["values"][:, 2] * hmult)
    assert f["init_Z"] == list(vals["init"][:, 0])
    assert f["init_F"] == list(vals["init"][:, 1] * hmult)
    assert f["perc_T"] == list(vals["perc"][:, 0]) and f["perc_A"] == list(vals["perc"][:, 1])
    assert f["irr_T"] == list(vals["irr"][:, 0]) and f["irr_A"] == list(vals["irr"][:, 1])
    assert f["check_T"] == list(vals["check"][:, 0])
    assert f["check_F"] == list(vals["check"][:, 2] * hmult)
    assert f["EV_T"] == list(vals["ev"][:, 0]) and f["LAI"] == list(vals["ev"][:, 1])
    assert f["EV_F"] == [list(vals["ev"][:, 2 + j]) for j in range(vals["ev"].shape[1] - 2)]
    assert f["params"] == list(params[:20])


def test_reads_the_format(checker, tmp_path):
    rng = np.random.default_rng(1)
    text, vals, params, _, _ = sample(rng)
    p = tmp_path / "fit_input.txt"
    p.write_text(text)
    f, lines = run(checker, p, 10.0)
    assert lines[1] == "counts 1 4 3 2 2 1 3 2"
    assert f["ev_mu"] == [0.25] and f["pso"] == [30, 0.7, 1.4, 1.6, 1e-6, 200]
    check_fields(f, vals, params, 10.0)


def test_layout_is_whitespace_separated(checker, tmp_path):
    # the reference reads with fscanf: line breaks and blank lines do not matter
    rng = np.random.default_rng(2)
    text, vals, params, _, _ = sample(rng)
    toks = text.split()
    odd = ""
    for i, t in enumerate(toks):
        odd += t + ("\n\n" if i % 5 == 4 else "  " if i % 3 else "\t")
    p = tmp_path / "fit_input.txt"
    p.write_text(odd)
    f, _ = run(checker, p, 1.0)
    check_fields(f, vals, params, 1.0)


def test_separate_et_and_parameter_files(checker, tmp_path):
    rng = np.random.default_rng(3)
    text, vals, _, ev, _ = sample(rng, nparam=25, ev_inline=False, params_inline=False)
    params = rng.uniform(0, 1, 25).round(8)
    (tmp_path / "fit_input.txt").write_text(text)
    (tmp_path / "et.txt").write_text(ev)
    (tmp_path / "pv.txt").write_text(" ".join(repr(float(x)) for x in params))
    f, _ = run(checker, tmp_path / "fit_input.txt", 2.0, tmp_path / "et.txt", tmp_path / "pv.txt")
    check_fields(f, vals, params, 2.0)  # the first 20 parameters
    assert len(f["params"]) == 20


@pytest.mark.parametrize("nparam", [0, 1, 19, 20])
def test_parameter_count(checker, tmp_path, nparam):
    rng = np.random.default_rng(10 + nparam)
    text, vals, params, _, _ = sample(rng, nparam=nparam)
    p = tmp_path / "fit_input.txt"
    p.write_text(text)
    f, _ = run(checker, p, 1.0)
    assert f["params"] == list(params)


def test_malformed_files_are_rejected(checker, tmp_path):
    rng = np.random.default_rng(4)
    text, *_ = sample(rng, nparam=0)
    toks = text.split()
    # header, values, init, perc, irr, check, ev tokens: cut inside the check rows
    cut = 15 + 4 * 3 + 3 * 2 + 2 * 2 + 1 * 2 + 3
    p = tmp_path / "short.txt"
    p.write_text(" ".join(toks[:cut]))
    assert run(checker, p, 1.0)[0] is None
    bad = list(toks)
    bad[17] = "abc"  # a value to fit
    p.write_text(" ".join(bad))
    assert run(checker, p, 1.0)[0] is None
    p.write_text("1 2 3\n")
    assert run(checker, p, 1.0)[0] is None
    assert run(checker, tmp_path / "missing.txt", 1.0)[0] is None


def test_result_rows_format(checker, tmp_path):
    rng = np.random.default_rng(5)
    text, *_ = sample(rng)
    p = tmp_path / "fit_input.txt"
    p.write_text(text)
    _, lines = run(checker, p, 1.0)
    rows = lines[-4:-1]  # the last line is the log row
    U, w = [-1.5, -0.25, 0.0], [0.3, 0.35, 0.4]
    assert rows == ["%g %g %g %g " % (1000.0, i * 0.05, U[i], w[i]) for i in range(3)]


def test_log_row_format_and_error_measure(checker, tmp_path):
    # check values with F = 0 (model value nonzero: relative term 1), a
    # positive model value (clamped to 0), a relative error above 1 (capped)
    # and an ordinary one; the checker's model values are 0.9 F + 0.5 for
    # the 1st and 4th, 0.9 F - 0.25 otherwise (mt_minimal.cpp:1655-1704)
    F = [0.0, -2.0, -0.1, -1.0]
    text = f"1 1 1 {len(F)} 0 0 1 1 0.25\n30 0.7 1.4 1.6 1e-6 200\n1.0 0.5 -1.0\n0.5 -1.0\n"
    text += "".join(f"{100.0 + i} {0.1 * (i + 1)} {f}\n" for i, f in enumerate(F))
    text += "0.0 1.0 0.5 \n0.3\n"
    p = tmp_path / "fit_input.txt"
    p.write_text(text)
    _, lines = run(checker, p, 1.0)
    v = [0.9 * f + (0.5 if i % 3 == 0 else -0.25) for i, f in enumerate(F)]
    v = [min(x, 0.0) for x in v]  # the model value counted is clamped to <= 0
    assert v == [0.0, -2.05, -0.34, -0.4]
    err = sum((f - x) ** 2 for f, x in zip(F, v))
    rel = [0.0,                   # F = 0 and v = 0: no term
           abs((-2.0 + 2.05) / -2.0), min(abs((-0.1 + 0.34) / -0.1), 1.0),
           abs((-1.0 + 0.4) / -1.0)]
    assert rel[2] == 1.0
    want = "%g " % 100.0 + "".join("%g %g %g " % (0.1 * (i + 1), f, x)
                                   for i, (f, x) in enumerate(zip(F, v)))
    want += "rlw %g twc %g sum_err %g avg_err %g avg_rel_err %g" % (0.31, 12.5, err, err / 4,
                                                                     sum(rel) / 4)
    assert lines[-1] == want


# ---- against the reference's own reader (tests/golden/fitio_*.npz: the
# reference's main in fitting mode, its fit_and_solve interposed to capture
# the parsed fields; see oracle/ref/harness_fitio.cpp)
def _unpack_reference(parsed):
    it = iter(parsed.tolist())
    called = next(it)
    if not called:
        return None
    head = [next(it) for _ in range(15)]
    to_fit, nv, ni, nc, npc, nir, net, nef = (int(x) for x in head[:8])
    f = dict(counts=[to_fit, nv, ni, nc, npc, nir, net, nef], ev_mu=[head[8]],
             pso=[head[9], head[10], head[11], head[12], head[13], head[14]])
    take = lambda n: [next(it) for _ in range(n)]  # noqa: E731
    for k, n in (("values_T", nv), ("values_Z", nv), ("values_F", nv), ("init_Z", ni),
                 ("init_F", ni), ("perc_T", npc), ("perc_A", npc), ("irr_T", nir), ("irr_A", nir),
                 ("check_T", nc), ("check_Z", nc), ("check_F", nc), ("EV_T", net), ("LAI", net)):
        f[k] = take(n)
    f["EV_F"] = [take(net) for _ in range(nef)]
    npar = int(next(it))
    f["nparam_values"] = npar
    f["params"] = take(npar)
    assert next(it, None) is None
    return f


@pytest.mark.parametrize("name", ["fitio_one_file", "fitio_files", "fitio_empty", "fitio_short",
                                  "fitio_19"])
def test_reader_matches_reference_reader(checker, golden, tmp_path, name):
    m, z = golden(name)
    for fn in m["files"]:
        (tmp_path / fn).write_text(str(z["file_" + fn.replace(".", "_")]))
    opts = dict(zip(m["args"][::2], m["args"][1::2]))
    path = lambda k: str(tmp_path / opts[k][1:]) if k in opts else "-"  # noqa: E731
    hmult = float(opts.get("H_mult", 10.0))  # the reference's default Hmult (mt_minimal.cpp:209)
    got, lines = run(checker, path("fi_file"), hmult, path("et_file"), path("pv_file"))
    ref = _unpack_reference(z["parsed"])
    if ref is None:
        assert got is None  # the reference's reader gave up on this file too
        return
    assert got is not None
    c = [int(x) for x in lines[1].split()[1:]]
    assert c == ref["counts"]
    assert got["ev_mu"] == ref["ev_mu"]
    pso = lines[3].split()[1:]
    assert [float(x) for x in pso] == ref["pso"]
    for k in ("values_T", "values_Z", "values_F", "init_Z", "init_F", "perc_T", "perc_A", "irr_T",
              "irr_A", "check_T", "check_Z", "check_F", "EV_T", "LAI", "EV_F"):
        assert got.get(k, []) == ref[k], k
    if name == "fitio_19":
        # 19 values: the reference's counter says 20 with the last one unset;
        # this reader returns the 19 it read
        assert ref["nparam_values"] == 20 and got["params"] == ref["params"][:19]
    else:
        assert got["params"] == ref["params"] and len(got["params"]) == ref["nparam_values"]
