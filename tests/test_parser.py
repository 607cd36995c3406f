"""Host parser and table builder vs the reference's golden params.header /
ff_groups.map (byte-exact, via sha256) and its error behaviour."""
import hashlib
import os

import pytest

from conftest import headers


@pytest.mark.parametrize("name", sorted(headers().keys()))
def test_header_and_map_byte_exact(ch, tmp_path, name):
    ent = headers()[name]
    open(tmp_path / "fm_setup.in", "w").write(ent["fm_setup"])
    if "traj_list" in ent:
        open(tmp_path / "traj_list.dat", "w").write(ent["traj_list"])
    s = ch.Setup(str(tmp_path / "fm_setup.in"))
    out = tmp_path / "out"
    out.mkdir()
    s.write_header(str(out))
    for fn in ["params.header", "ff_groups.map"]:
        data = open(out / fn, "rb").read()
        assert len(data) == ent[fn]["bytes"], fn
        assert hashlib.sha256(data).hexdigest() == ent[fn]["sha256"], fn
    if "dim" in ent:
        assert int(ent["dim"][0]) == s.model.ncols


def test_column_counts_of_benchmark_orders(ch, synth, tmp_path):
    # Column counts of the benchmark configurations:
    # O/H 12 4 0 -> 136, C 16 9 5 -> 1038, C/H/N/O 12 6 3 -> 10504.
    cases = [(synth.water(nframes=1), 136), (synth.carbon(nframes=1, natoms=64), 1038)]
    for (text, _, _), ncols in cases:
        p = tmp_path / ("s%d.in" % ncols)
        open(p, "w").write(text)
        setup = ch.Setup(str(p))
        assert setup.model.ncols == ncols
    text = synth.fm_setup(["C", "H", "N", "O"], [12.011, 1.008, 14.007, 15.999], 12, 6, 3, 0.8, 5.0, 1.4,
                          smax3=4.0, smax4=3.5)
    open(tmp_path / "chno.in", "w").write(text)
    setup = ch.Setup(str(tmp_path / "chno.in"))
    m = setup.model
    assert (m.n2b_cols, m.c3.ncols, m.c4.ncols, m.ncols) == (120, 2380, 8000, 10504)
    assert (m.c3.total_allowed, m.c4.total_allowed) == (4000, 22260)


def test_allowed_sets_are_permutation_closed(ch, synth, tmp_path):
    """Each column's allowed power sets form one orbit of the atom permutations
    that preserve atom types, and every atom carries a non-zero power."""
    import itertools
    text = synth.fm_setup(["C", "H"], [12.0, 1.0], 6, 4, 3, 0.8, 4.0, 1.4)
    open(tmp_path / "a.in", "w").write(text)
    setup = ch.Setup(str(tmp_path / "a.in"))  # owns the model tables
    m = setup.model
    for cs in (m.c3, m.c4):
        nb, np_ = cs.nbody, cs.npairs
        pairs = [(i, j) for i in range(nb) for j in range(i + 1, nb)]
        for t in range(cs.ntypes):
            types = [cs.atom_type[t * nb + i] for i in range(nb)]
            b, e = cs.perm_offset[t], cs.perm_offset[t + 1]
            sets = {}
            for q in range(b, e):
                pw = tuple(cs.perm_powers[q * np_ + k] for k in range(np_))
                sets.setdefault(cs.perm_col[q], set()).add(pw)
                for a in range(nb):
                    assert any(pw[k] > 0 for k, (i, j) in enumerate(pairs) if a in (i, j))
            assert sorted(sets) == list(range(cs.n_unique[t]))
            for col, orbit in sets.items():
                p0 = next(iter(orbit))
                mat = {}
                for k, (i, j) in enumerate(pairs):
                    mat[(i, j)] = mat[(j, i)] = p0[k]
                closure = set()
                for perm in itertools.permutations(range(nb)):
                    if [types[perm[i]] for i in range(nb)] != types:
                        continue
                    closure.add(tuple(mat[(perm[i], perm[j])] for (i, j) in pairs))
                assert closure == orbit


def _write(tmp_path, text):
    p = tmp_path / "fm_setup.in"
    open(p, "w").write(text)
    return str(p)


def test_error_missing_endfile(ch, tmp_path):
    with pytest.raises(RuntimeError, match="ENDFILE"):
        ch.Setup(_write(tmp_path, "# PAIRTYP #\n\tCHEBYSHEV 8\n"))


def test_error_1000_line_cap(ch, tmp_path):
    text = "".join("# COMMENTKEY #\n" for _ in range(1001)) + "# ENDFILE #\n"
    with pytest.raises(RuntimeError, match="1000 lines"):
        ch.Setup(_write(tmp_path, text))


def test_error_fitcoul_with_stress_and_fixed_charges(ch, synth, tmp_path):
    """FITCOUL with stress rows stops as RUN_SANITY_LSQ does (src/input.C:1336-1337);
    fixed non-zero charges (Ewald forces subtracted from b) are not supported."""
    text = synth.fm_setup(["C"], [12.0], 8, 0, 0, 1.0, 4.0, 1.4).replace("# FITCOUL #\n\tfalse", "# FITCOUL #\n\ttrue")
    with pytest.raises(RuntimeError, match="not compatible with use of ZCalc_Ewald_Deriv"):
        ch.Setup(_write(tmp_path, text))
    text = synth.fm_setup(["C"], [12.0], 8, 0, 0, 1.0, 4.0, 1.4).replace("\t\t0.0\t\t12", "\t\t0.41\t\t12")
    with pytest.raises(RuntimeError, match="UNSUPPORTED"):
        ch.Setup(_write(tmp_path, text))


def test_error_unknown_atom_type_in_trajectory(ch, synth, tmp_path):
    text = synth.fm_setup(["C"], [12.0], 8, 0, 0, 1.0, 4.0, 1.4, nframes=1)
    p = _write(tmp_path, text)
    open(tmp_path / "input.xyzf", "w").write("1\n10.0 10.0 10.0 0 0 0 0\nXx 0 0 0 0 0 0\n")
    s = ch.Setup(p)
    with pytest.raises(RuntimeError, match="Did not recognize atom type Xx"):
        s.frames()


def test_error_missing_box_tokens(ch, synth, tmp_path):
    text = synth.fm_setup(["C"], [12.0], 8, 0, 0, 1.0, 4.0, 1.4, nframes=1)
    p = _write(tmp_path, text)
    open(tmp_path / "input.xyzf", "w").write("1\n10.0 10.0\nC 0 0 0 0 0 0\n")
    with pytest.raises(RuntimeError, match="Missing box length"):
        ch.Setup(p).frames()


def test_error_missing_stress_tokens(ch, synth, tmp_path):
    text = synth.fm_setup(["C"], [12.0], 8, 0, 0, 1.0, 4.0, 1.4, nframes=1)
    p = _write(tmp_path, text)
    open(tmp_path / "input.xyzf", "w").write("1\n10.0 10.0 10.0 1.0\nC 0 0 0 0 0 0\n")
    with pytest.raises(RuntimeError, match="Missing diagonal stress"):
        ch.Setup(p).frames()


def test_trajectory_units(ch, synth, tmp_path):
    """Forces are read in Hartree/Bohr and converted (src/ClassDefs.C:2632-2634)."""
    text = synth.fm_setup(["C"], [12.0], 8, 0, 0, 1.0, 4.0, 1.4, nframes=1)
    p = _write(tmp_path, text)
    open(tmp_path / "input.xyzf", "w").write("1\n10.0 11.0 12.0 1.0 2.0 3.0 -5.5\nC 0.5 0.25 0.125 1.0 -2.0 0.5\n")
    f = ch.Setup(p).frames()[0]
    conv = 627.50960803 * 1.889725989
    assert list(f.forces) == [1.0 * conv, -2.0 * conv, 0.5 * conv]
    assert list(f.cell) == [10.0, 0, 0, 0, 11.0, 0, 0, 0, 12.0]
    assert list(f.stress[:3]) == [1.0, 2.0, 3.0] and f.energy == -5.5
