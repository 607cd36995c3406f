"""params.txt reader (the file chimes_md reads): model rebuilt from its header,
coefficients placed in A column order, checked against the reference's own
fitted values and, through the oracle's force evaluation, its forceout.txt."""
import os

import numpy as np
import pytest

from conftest import MD_CASES, golden_dir, md_case


def test_nonorth2_coefficients_match_test_suite_params(ch):
    d = golden_dir("nonorth2")
    s = ch.Setup.from_params(os.path.join(d, "params.txt"))
    x, pd, ps = s.coeffs()
    xr = np.array([float(l.split()[1]) for l in open(os.path.join(d, "test_suite_params.txt")) if l.strip()])
    assert x.shape == xr.shape
    assert np.array_equal(x, xr)
    assert (pd, ps) == (0.01, 1.0e4)
    # the rebuilt model is the one the fm_setup.in of the fit describes
    sf = ch.Setup(os.path.join(d, "fm_setup.in"))
    assert sf.model.ncols == s.model.ncols and sf.model.order2b == s.model.order2b


@pytest.mark.parametrize("case", MD_CASES)
def test_params_reader_oracle_forces(ch, oracle, tmp_path, case):
    """Oracle forces with the coefficients read from params.txt reproduce the
    reference chimes_md forceout.txt (14 digits)."""
    s, f, ref = md_case(ch, case, tmp_path)
    x, pd, ps = s.coeffs()
    F, E, V = oracle.force(s, f, x, penalty_scale=ps, penalty_dist=pd)
    # relative to the largest force, floor 1 kcal/mol/A (the TiO2 crystals'
    # net forces cancel to ~1e-7)
    assert np.max(np.abs(F - ref)) <= 1e-10 * max(np.abs(ref).max(), 1.0)
    # energy (with the per-atom offsets): the reference run printed
    # |E - E_cmp| / N to 6 decimals, E_cmp being the energy row of its force.txt
    import json
    from conftest import md_dir
    meta = json.load(open(os.path.join(md_dir(case), "expected.json")))
    if meta["abs_energy_error_per_atom"] is not None:
        err = abs(E - meta["forcetxt_energy"]) / f.natoms
        assert round(err, 6) == pytest.approx(meta["abs_energy_error_per_atom"], abs=1e-12)
    from conftest import check_md_printout
    check_md_printout(meta, f.natoms, f.cell, E, V)


def test_params_reader_refuses_coulomb(ch, tmp_path):
    import gzip
    from conftest import md_dir
    text = gzip.open(os.path.join(md_dir("special4b"), "params.txt.gz"), "rt").read()
    p = tmp_path / "params.txt"
    p.write_text(text.replace("USECOUL: false", "USECOUL: true"))
    with pytest.raises(RuntimeError, match="UNSUPPORTED"):
        ch.Setup.from_params(str(p))
    p.write_text(text.replace("   PAIRS: OO OO OO UNIQUE: 7 TOTAL: 20", "   PAIRS: OO OO OO UNIQUE: 8 TOTAL: 20"))
    with pytest.raises(RuntimeError, match="UNIQUE/TOTAL"):
        ch.Setup.from_params(str(p))


def test_trimmed_params_file_chebyfix_case(ch, tmp_path):
    """The reference's published HN3 force field lists only the power sets that
    carry parameters (UNIQUE: -1); each listed set lands on its column, and a
    column is listed with all its permutations."""
    import gzip
    import os
    import re
    from chebyfix_cases import CHEBYFIX_DIR
    p = os.path.join(str(tmp_path), "params.txt")
    text = gzip.open(os.path.join(CHEBYFIX_DIR, "params.txt.gz"), "rb").read().decode()
    open(p, "w").write(text)
    s = ch.Setup.from_params(p)
    x, pd, ps = s.coeffs()
    assert (pd, ps) == (0.1, 1.0e4)
    m = s.model
    assert (m.order2b, m.c3.order, m.c4.order) == (12, 8, 4)
    assert m.fcut_type == 1 and m.fcut_offset == 0.5 and m.n_offset_cols == 2
    # every parameter of the file is in x (the listed sets of one column share it)
    vals = set()
    for blk in re.split(r"(?:TRIPLETTYPE|QUADRUPLETYPE) PARAMS:", text)[1:]:
        rows = [l.split() for l in blk.split("----------------------------------------------------------------------------")[1].split("\n")]
        for r in rows:
            if len(r) in (7, 10):
                vals.add(float(r[-1]))
            elif r and not r[0].isdigit():
                break
    xs = set(x[m.n2b_cols:m.ncols_short][x[m.n2b_cols:m.ncols_short] != 0].tolist())
    assert vals - {0.0} == xs
    # a block that lists only part of a column's permutations is refused
    lines = text.split("\n")
    head = next(i for i, l in enumerate(lines) if "UNIQUE: -1 TOTAL: 390" in l)
    drop = head + 3 + 388  # row 388, one of three permutations of its column
    assert lines[drop].split()[0] == "388"
    pcol = lines[drop].split()[-2]
    assert sum(1 for l in lines[head + 3:drop] if l.split()[-2:-1] == [pcol]) == 2
    lines[head] = lines[head].replace("TOTAL: 390", "TOTAL: 389")
    lines[drop + 1] = "388" + lines[drop + 1][3:]  # renumber the row after it
    del lines[drop]
    open(p, "w").write("\n".join(lines))
    with pytest.raises(RuntimeError, match="permutations"):
        ch.Setup.from_params(p)
