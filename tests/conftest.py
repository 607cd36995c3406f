"""Shared fixtures: the ctypes binding, the CPU oracle and the golden data."""
import ctypes
import importlib.util
import json
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "chimes_lsq-myllfork_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def ch():
    return _load("chimes", os.path.join(PKG, "chimes.py"))


@pytest.fixture(scope="session")
def synth():
    return _load("synth", os.path.join(PKG, "synth.py"))


class Oracle:
    """ctypes wrapper of oracle/liboracle.so (test infrastructure only)."""

    def __init__(self, ch):
        self.ch = ch
        L = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_assemble_frame.argtypes = [ctypes.POINTER(ch.Model), ctypes.POINTER(ch.Frame),
                                         ctypes.POINTER(ctypes.c_int), dp]
        L.orc_count_clusters.argtypes = [ctypes.POINTER(ch.Model), ctypes.POINTER(ch.Frame),
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_longlong)]
        L.orc_polys.argtypes = [ctypes.POINTER(ch.Model), ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, ctypes.c_int, ctypes.c_double, dp, dp]
        L.orc_fcut.argtypes = [ctypes.POINTER(ch.Model), ctypes.c_double, ctypes.c_double, dp, dp]
        L.orc_volume.argtypes = [dp]
        L.orc_volume.restype = ctypes.c_double
        L.orc_force.argtypes = [ctypes.POINTER(ch.Model), ctypes.POINTER(ch.Frame), dp, ctypes.c_double,
                                ctypes.c_double, dp, dp, dp]
        self.L = L

    def assemble(self, setup, frames):
        small = ctypes.c_int(0)
        out = []
        for f in frames:
            o = np.zeros((setup.frame_rows(f.natoms, f.index), setup.model.ncols))
            cf = f.c()
            n = self.L.orc_assemble_frame(setup.model_ptr, ctypes.byref(cf), ctypes.byref(small),
                                          o.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            assert n == o.shape[0]
            out.append(o)
        return np.vstack(out)

    def force(self, setup, frame, x, penalty_scale=1e4, penalty_dist=0.01):
        F = np.zeros(3 * frame.natoms)
        E = ctypes.c_double()
        V = np.zeros(9)
        x = np.ascontiguousarray(x, dtype=np.float64)
        cf = frame.c()
        dp = ctypes.POINTER(ctypes.c_double)
        rc = self.L.orc_force(setup.model_ptr, ctypes.byref(cf), x.ctypes.data_as(dp), penalty_scale, penalty_dist,
                              F.ctypes.data_as(dp), ctypes.byref(E), V.ctypes.data_as(dp))
        assert rc == 0
        return F, E.value, V


@pytest.fixture(scope="session")
def oracle(ch):
    return Oracle(ch)


def colscaled_error(A, R, floor=1e-3):
    """max |A-R| / max(|R|, floor * max_i |R[:, j]|) (column-scaled relative error)."""
    A = np.asarray(A)
    R = np.asarray(R)
    colmax = np.abs(R).max(axis=0)
    den = np.maximum(np.abs(R), floor * np.maximum(colmax, 1e-300))
    err = np.abs(A - R) / den
    err[(A == 0) & (R == 0)] = 0.0
    return float(err.max()) if err.size else 0.0


def golden_dir(name):
    return os.path.join(GOLDEN, "lsq", name)


def headers():
    return json.load(open(os.path.join(GOLDEN, "headers.json")))


def synth_setup(ch, synth, tmp_path, gen):
    """Write a generated fm_setup.in and return (Setup, [PyFrame])."""
    text, frames, names = gen
    p = os.path.join(str(tmp_path), "fm_setup.in")
    open(p, "w").write(text)
    s = ch.Setup(p)
    fr = [ch.PyFrame(f["natoms"], f["xyz"], f["types"], f["cell"], f["forces"], f["stress"], f["energy"], f["index"])
          for f in frames]
    return s, fr


def md_dir(name):
    return os.path.join(GOLDEN, "lsq", "..", "md", name)


MD_CASES = ["special3b", "special4b", "stress-and-ener-2b1", "stress-and-ener-4b", "stress-and-ener-4b2", "h2o-3bcheby",
            "h2o-4bcheby", "run-ti-nonortho", "run-ti-ortho"]


def md_case(ch, name, tmp_path):
    """(Setup from params.txt, PyFrame of the first coordinate frame, golden forces)."""
    import gzip
    d = md_dir(name)
    p = os.path.join(str(tmp_path), "params.txt")
    with gzip.open(os.path.join(d, "params.txt.gz"), "rb") as g:
        open(p, "wb").write(g.read())
    s = ch.Setup.from_params(p)
    names = s.type_names()
    lines = open(os.path.join(d, "frame.xyzf")).read().split("\n")
    n = int(lines[0].split()[0])
    t = lines[1].split()
    # box line: three lengths, or NON_ORTHO and the cell vectors a, b, c
    cell = np.array([float(v) for v in t[1:10]]) if t[0] == "NON_ORTHO" else np.diag([float(v) for v in t[:3]])
    xyz = np.array([[float(v) for v in lines[2 + i].split()[1:4]] for i in range(n)])
    typ = np.array([names.index(lines[2 + i].split()[0]) for i in range(n)], dtype=np.int32)
    f = ch.PyFrame(n, xyz, typ, cell)
    ref = np.array([float(v) for v in open(os.path.join(d, "forceout.txt")).read().split()])
    assert ref.size == 3 * n
    return s, f, ref


WATER = os.path.join(GOLDEN, "lsq", "water")


def water_setup(ch, tmp_path, name, nframes=None):
    """A FITCOUL water case of the reference test suite in tmp_path: its
    fm_setup.in (NFRAMES optionally lowered) and the shared trajectory."""
    import gzip
    import re
    text = open(os.path.join(WATER, name, "fm_setup.in")).read()
    if nframes is not None:
        text = re.sub(r"(# NFRAMES #[^\n]*\n)[^\n]*\n", lambda m: m.group(1) + "\t%d\n" % nframes, text, count=1)
    open(os.path.join(str(tmp_path), "fm_setup.in"), "w").write(text)
    with gzip.open(os.path.join(WATER, "input.xyzf.gz"), "rb") as g:
        open(os.path.join(str(tmp_path), "input.xyzf"), "wb").write(g.read())
    return ch.Setup(os.path.join(str(tmp_path), "fm_setup.in"))


def fitcoul_text(text, signs, constraints):
    """A generated fm_setup.in turned into a FITCOUL one: charge signs in the
    ATMCHRG column, no stress / energy rows, CHARGE CONSTRAINTS lines."""
    import re
    text = text.replace("# FITCOUL #\n\tfalse", "# FITCOUL #\n\ttrue")
    text = text.replace("# FITSTRS #\n\ttrue", "# FITSTRS #\n\tfalse").replace("# FITENER #\n\ttrue", "# FITENER #\n\tfalse")
    lines = text.split("\n")
    k = lines.index("# TYPEIDX #\t# ATM_TYP #\t# ATMCHRG #\t# ATMMASS #")
    for i, sg in enumerate(signs):
        parts = lines[k + 1 + i].split("\t\t")
        parts[2] = sg
        lines[k + 1 + i] = "\t\t".join(parts)
    e = lines.index("# FCUTTYP #")
    lines[e:e] = ["CHARGE CONSTRAINTS:"] + constraints + [""]
    return "\n".join(lines)


def check_md_printout(meta, natoms, cell, E, V):
    """Energy, pressure and stress a one-step chimes_md run printed for its
    initial configuration at zero temperature: Vtot/N to its 7 decimals, P
    (GPa, with the run's kinetic term of Ktot/N at most 3e-7 kcal/mol) within
    2e-7 GPa, the diagonal stresses to their 4 decimals."""
    if "vtot_per_atom" not in meta:
        return
    assert round(E / natoms, 7) == pytest.approx(meta["vtot_per_atom"], abs=1e-12)
    gpa = 6.9476955  # kcal/mol/A^3 per GPa, as the stress rows of A
    vol = abs(np.linalg.det(np.asarray(cell).reshape(3, 3)))
    W = np.asarray(V).reshape(3, 3)
    assert np.trace(W) / (3 * vol) * gpa == pytest.approx(meta["pressure_gpa"], abs=2e-7)
    for k, a in (("xx", 0), ("yy", 1), ("zz", 2)):
        assert round(W[a, a] / vol * gpa, 4) == pytest.approx(meta["sigma_gpa"][k], abs=1e-9)
