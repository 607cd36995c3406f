"""Hierarchical fit: EXCLUDE 1B/2B INTERACTION with HIERARC true (src/input.C:
760-840) drops the excluded energy-offset and 2-body columns from A.txt
(A_MAT::skip_2b / add_col_of_ones, src/A_Matrix.C:155-208), lists them on the
EXCLD1B / EXCLD2B header lines (src/chimes_lsq.C:592-610), reduces the dim.txt
column count by N_EXCL_LT_3B, and chimes_lsq.py prints zeros for them
(src/chimes_lsq.py:335-349,470-476,652-663).  No reference test case covers
it; these checks pin the layout against the same setup without exclusions."""
import ctypes
import os
import re

import numpy as np
import pytest

EXCL = "# HIERARC #\n\ttrue\nEXCLUDE 1B INTERACTION: 1\nH\nEXCLUDE 2B INTERACTION: 1\nO H\n\n# ENDFILE #"


def _setups(ch, synth, tmp_path):
    text, frames, names = synth.water(nframes=2, nmol=27, o2=8, o3=4)
    out = {}
    for tag, t in (("full", text), ("excl", text.replace("# ENDFILE #", EXCL))):
        d = tmp_path / tag
        d.mkdir()
        open(d / "fm_setup.in", "w").write(t)
        out[tag] = ch.Setup(str(d / "fm_setup.in"))
    fr = [ch.PyFrame(f["natoms"], f["xyz"], f["types"], f["cell"], f["forces"], f["stress"], f["energy"], f["index"])
          for f in frames]
    return out["full"], out["excl"], fr


def test_header_lines_and_kept_columns(ch, synth, tmp_path):
    full, ex, fr = _setups(ch, synth, tmp_path)
    m = ex.model
    ex.write_header(str(tmp_path / "excl"))
    full.write_header(str(tmp_path / "full"))
    h = open(tmp_path / "excl" / "params.header").read()
    assert "EXCLD1B: 1 \nEXCLD2B: 2 \n" in h  # type H = 1; pairs OO, HH, OH
    assert h.replace("EXCLD1B: 1 \nEXCLD2B: 2 \n", "EXCLD1B: false\nEXCLD2B: false\n") == \
        open(tmp_path / "full" / "params.header").read()
    keep = ex.fit_columns()
    dropped = [c for c in range(m.ncols) if c not in keep]
    assert dropped == list(range(2 * 8, 3 * 8)) + [m.ncols_short + 1]
    assert full.fit_columns() == list(range(m.ncols))


def test_A_text_without_excluded_columns(ch, synth, oracle, tmp_path):
    full, ex, fr = _setups(ch, synth, tmp_path)
    A = oracle.assemble(full, fr)
    rows, r = [], 0
    for f in fr:
        n = full.frame_rows(f.natoms, f.index)
        rows.append(A[r:r + n])
        r += n
    for s, tag in ((full, "full"), (ex, "excl")):
        s.write_text(str(tmp_path / tag), fr, rows=rows)
    keep = ex.fit_columns()
    lf = open(tmp_path / "full" / "A.txt").read().split("\n")
    le = open(tmp_path / "excl" / "A.txt").read().split("\n")
    assert len(lf) == len(le)
    for a, b in zip(lf, le):
        assert b.split() == [a.split()[c] for c in keep] if a else b == a
    nf, rows_f = open(tmp_path / "full" / "dim.txt").read().split()
    ne, rows_e = open(tmp_path / "excl" / "dim.txt").read().split()
    assert int(ne) == int(nf) - 8 - 1 == len(keep) and rows_e == rows_f
    for fn in ["b.txt", "b-labeled.txt", "natoms.txt"]:
        assert open(tmp_path / "full" / fn).read() == open(tmp_path / "excl" / fn).read()


def test_params_txt_prints_zeros_for_excluded(ch, synth, tmp_path):
    full, ex, fr = _setups(ch, synth, tmp_path)
    keep = ex.fit_columns()
    x = np.arange(1, len(keep) + 1, dtype=np.float64) * 0.5
    res = ch.FitResult()
    res.nvars = len(keep)
    ch.check(ch.lib().cg_write_params(ex.h, str(tmp_path).encode(), ch.dptr(x), len(x), 100, ctypes.byref(res),
                                      None, None))
    txt = open(tmp_path / "params.txt").read()
    blocks = re.findall(r"PAIRTYPE PARAMS: (\d+) (\S+) (\S+)\n\n((?:\s*\d+ +\S+\n)+)", txt)
    vals = {int(i): [float(l.split()[1]) for l in b.strip().split("\n")] for i, _, _, b in blocks}
    assert vals[2] == [0.0] * 8
    assert vals[0] == list(x[0:8]) and vals[1] == list(x[8:16])
    offs = re.findall(r"ENERGY OFFSET (\d) +(\S+)", txt)
    assert [(int(i), float(v)) for i, v in offs] == [(1, x[-1]), (2, 0.0)]
