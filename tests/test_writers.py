"""Text outputs: A.txt row format, params.txt layout (chimes_lsq.py printer)."""
import ctypes
import os
import re

import numpy as np

from conftest import golden_dir


def _summary(path):
    """Fit summary numbers of a golden params.txt header."""
    vals = {}
    for line in open(path):
        if not line.startswith("!"):
            break
        if "=" in line:
            k = line.rindex("=")
            vals[line[2:k].strip()] = line[k + 1:].split()[0]
    return vals


def test_params_txt_matches_reference_layout(ch, tmp_path):
    """Given the reference's coefficients and fit summary, the params.txt
    printer reproduces the reference file byte for byte (Date line aside)."""
    d = golden_dir("nonorth2")
    s = ch.Setup(d + "/fm_setup.in")
    x = np.array([float(l.split()[1]) for l in open(d + "/test_suite_params.txt")])
    v = _summary(d + "/params.txt")
    res = ch.FitResult()
    res.eps = float(v["SVD regularization factor"])
    res.eps_abs = float(v["eps (= args.eps*dmax)"])
    res.rms = float(v["RMS force error"])
    res.max_abs_x = float(v["max abs variable"])
    res.nvars = int(v["number of fitting vars"])
    res.bic = float(v["Bayesian Information Criterion"])
    L = ch.lib()
    ch.check(L.cg_write_params(s.h, str(tmp_path).encode(), ch.dptr(x), len(x), 8640, ctypes.byref(res), None, None))
    got = open(tmp_path / "params.txt").read().split("\n")
    ref = open(d + "/params.txt").read().split("\n")
    assert got[1:] == ref[1:]
    assert got[0].startswith("! Date ")


def test_A_text_rows(ch, oracle, tmp_path):
    """A.txt rows: %.16e entries, 3-space separators on force rows, values equal
    to the reference's text within the assembly tolerance."""
    d = golden_dir("nonorth2")
    s = ch.Setup(d + "/fm_setup.in")
    fr = s.frames(limit=1)
    A = oracle.assemble(s, fr)
    L = ch.lib()
    ch.check(L.cg_write_open(s.h, str(tmp_path).encode()))
    cf = fr[0].c()
    ch.check(L.cg_write_frame(s.h, ctypes.byref(cf), ch.dptr(A), A.shape[0]))
    ch.check(L.cg_write_close(s.h))
    lines = open(tmp_path / "A.txt").read().split("\n")
    assert re.fullmatch(r"(-?\d\.\d{16}e[+-]\d{2}   ){12}", lines[0])
    got = np.array([[float(t) for t in l.split()] for l in lines if l.strip()])
    g = np.load(d + "/golden.npz")["A"][: got.shape[0]]
    assert np.allclose(got, g, rtol=1e-10, atol=1e-13)
    assert open(tmp_path / "natoms.txt").read().split("\n")[0] == "288"


import hashlib

from conftest import headers


def _res_from_summary(ch, lines):
    vals = {}
    for line in lines:
        if "=" in line:
            k = line.rindex("=")
            vals[line[2:k].strip()] = line[k + 1:].split()[0]
    res = ch.FitResult()
    res.eps = float(vals["SVD regularization factor"])
    res.eps_abs = float(vals["eps (= args.eps*dmax)"])
    res.rms = float(vals["RMS force error"])
    res.max_abs_x = float(vals["max abs variable"])
    res.nvars = int(vals["number of fitting vars"])
    res.bic = float(vals["Bayesian Information Criterion"])
    return res, int(vals["Number of equations"])


import pytest


@pytest.mark.parametrize("name", sorted(k for k, v in headers().items() if "params_txt" in v))
def test_params_txt_byte_exact_all_tests(ch, tmp_path, name):
    """The params.txt printer (2B/3B/4B tables, exclusions, energy offsets)
    reproduces every reference params.txt given its coefficients."""
    ent = headers()[name]
    open(tmp_path / "fm_setup.in", "w").write(ent["fm_setup"])
    if "traj_list" in ent:
        open(tmp_path / "traj_list.dat", "w").write(ent["traj_list"])
    s = ch.Setup(str(tmp_path / "fm_setup.in"))
    x = np.array(ent["x"])
    res, neq = _res_from_summary(ch, ent["params_txt"]["summary"])
    out = tmp_path / "out"
    out.mkdir()
    ch.check(ch.lib().cg_write_params(s.h, str(out).encode(), ch.dptr(x), len(x), neq, ctypes.byref(res), None, None))
    body = "\n".join(open(out / "params.txt").read().split("\n")[1:]).encode()
    assert len(body) == ent["params_txt"]["bytes"]
    assert hashlib.sha256(body).hexdigest() == ent["params_txt"]["sha256"]
