"""GPU force evaluation (serving kernel) vs the oracle's Force_all/3B/4B."""
import numpy as np
import pytest

from conftest import synth_setup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["chno", "carbon"])
def test_force_matches_oracle(ch, synth, oracle, tmp_path, case):
    if case == "chno":
        gen = synth.chno(nframes=1, nmin=150, nmax=150, seed=41)
    else:
        gen = synth.carbon(nframes=1, natoms=128, seed=42)
    s, fr = synth_setup(ch, synth, tmp_path, gen)
    rng = np.random.default_rng(3)
    x = rng.normal(size=s.model.ncols)
    dev = ch.Device(s)
    F, E, V = dev.force(x, fr[0])
    Fo, Eo, Vo = oracle.force(s, fr[0], x)
    assert np.allclose(F, Fo, rtol=1e-10, atol=1e-10 * np.abs(Fo).max())
    assert E == pytest.approx(Eo, rel=1e-10, abs=1e-10 * abs(Eo))
    assert np.allclose(V, Vo, rtol=1e-10, atol=1e-10 * np.abs(Vo).max())
    # F = A x on the force rows (no pair below s_min + 0.01 in these inputs)
    A, b = dev.assemble(fr)
    assert np.allclose(F, A[: 3 * fr[0].natoms] @ x, rtol=1e-9, atol=1e-9 * np.abs(F).max())
    dev.close()


def test_force_penalty_below_smin(ch, synth, oracle, tmp_path):
    """A pair closer than s_min + 0.01 adds the penalty A_p (s_min + d_p - r)^3."""
    text = synth.fm_setup(["C"], [12.0], 8, 4, 3, 1.2, 4.0, 1.4, smax3=3.5, smax4=3.0)
    p = tmp_path / "fm_setup.in"
    open(p, "w").write(text)
    s = ch.Setup(str(p))
    rng = np.random.default_rng(4)
    pos = synth.random_sequential(rng, 30, np.diag([9.0, 9.0, 9.0]), 1.3)
    pos[1] = pos[0] + np.array([1.15, 0.0, 0.0])  # below s_min
    f = ch.PyFrame(30, pos, np.zeros(30, dtype=np.int32), np.diag([9.0, 9.0, 9.0]))
    x = rng.normal(size=s.model.ncols)
    dev = ch.Device(s)
    F, E, V = dev.force(x, f)
    Fo, Eo, Vo = oracle.force(s, f, x)
    assert np.allclose(F, Fo, rtol=1e-10, atol=1e-10 * np.abs(Fo).max())
    assert E == pytest.approx(Eo, rel=1e-10)
    dev.close()


def test_force_penalty_change_after_graph_capture(ch, synth, oracle, tmp_path):
    """Same-shape calls are captured into a graph; a penalty change afterwards
    is honoured by the next call (equal to a fresh context with that penalty,
    and to the oracle)."""
    text = synth.fm_setup(["C"], [12.0], 8, 4, 3, 1.2, 4.0, 1.4, smax3=3.5, smax4=3.0)
    p = tmp_path / "fm_setup.in"
    open(p, "w").write(text)
    s = ch.Setup(str(p))
    rng = np.random.default_rng(11)
    cell = np.diag([9.0, 9.0, 9.0])
    pos = synth.random_sequential(rng, 30, cell, 1.3)
    pos[1] = pos[0] + np.array([1.15, 0.0, 0.0])  # below s_min
    f = ch.PyFrame(30, pos, np.zeros(30, dtype=np.int32), cell)
    x = rng.normal(size=s.model.ncols)
    dev = ch.Device(s)
    for _ in range(4):  # exact, capacity-sized, captured, replayed
        dev.force(x, f)
    assert dev.force_stats()[2] >= 1  # the fourth call replayed a graph
    dev.set_penalty(0.05, 2.0e4)
    F, E, V = dev.force(x, f)
    fresh = ch.Device(s)
    fresh.set_penalty(0.05, 2.0e4)
    F0, E0, V0 = fresh.force(x, f)
    fresh.close()
    scale = np.abs(F0).max()
    assert np.max(np.abs(F - F0)) <= 1e-12 * scale
    assert E == pytest.approx(E0, rel=1e-12)
    Fo, Eo, Vo = oracle.force(s, f, x, penalty_scale=2.0e4, penalty_dist=0.05)
    assert np.max(np.abs(F - Fo)) <= 1e-10 * scale
    assert E == pytest.approx(Eo, rel=1e-10)
    # and the old penalty differs (the change is visible at all)
    Fd, Ed, Vd = oracle.force(s, f, x)
    assert abs(Ed - Eo) > 1e-6 * abs(Eo)
    dev.close()


def test_force_small_cell_needs_two_layers(ch, synth, oracle, tmp_path):
    """A configuration whose cell is thinner than the cutoff gets the images
    of a second layer (the force path searches as many layers as the cutoff
    needs), like the oracle's force evaluation."""
    text = synth.fm_setup(["C"], [12.0], 8, 4, 0, 1.0, 4.0, 1.4, smax3=3.0)
    p = tmp_path / "fm_setup.in"
    open(p, "w").write(text)
    s = ch.Setup(str(p))
    rng = np.random.default_rng(12)
    cell = np.array([[3.4, 0.0, 0.0], [0.9, 3.6, 0.0], [0.4, -0.6, 3.5]])
    pos = synth.random_sequential(rng, 3, cell, 1.2)
    f = ch.PyFrame(3, pos, np.zeros(3, dtype=np.int32), cell)
    x = rng.normal(size=s.model.ncols)
    dev = ch.Device(s)
    F, E, V = dev.force(x, f)
    dev.close()
    Fo, Eo, Vo = oracle.force(s, f, x)
    assert np.allclose(F, Fo, rtol=1e-10, atol=1e-10 * np.abs(Fo).max())
    assert E == pytest.approx(Eo, rel=1e-10)
    assert np.allclose(V, Vo, rtol=1e-10, atol=1e-10 * np.abs(Vo).max())


def test_force_batch_matches_single_calls(ch, synth, oracle, tmp_path):
    """A batch of configurations of different sizes (one in a small cell)
    gives each configuration the result it gets alone, and the oracle's."""
    text, frames, _ = synth.chno(nframes=4, nmin=40, nmax=400, seed=43)
    s, fr = synth_setup(ch, synth, tmp_path, (text, frames, _))
    rng = np.random.default_rng(5)
    # a 4.9 A cubic cell (below the 5 A cutoff: small-cell lists) in the middle
    cell = np.diag([4.9, 4.9, 4.9])
    pos = synth.random_sequential(rng, 10, cell, 1.0)
    small = ch.PyFrame(10, pos, rng.integers(0, 4, 10).astype(np.int32), cell)
    fr = fr[:2] + [small] + fr[2:]
    x = rng.normal(size=s.model.ncols)
    dev = ch.Device(s)
    Fb, Eb, Vb = dev.force_batch(x, fr)
    for i, f in enumerate(fr):
        F1, E1, V1 = dev.force(x, f)
        Fo, Eo, Vo = oracle.force(s, f, x)
        scale = max(np.abs(Fo).max(), 1e-300)
        assert np.max(np.abs(Fb[i] - F1)) <= 1e-12 * scale
        assert np.max(np.abs(Fb[i] - Fo)) <= 1e-10 * scale
        assert Eb[i] == pytest.approx(E1, rel=1e-12, abs=1e-12 * abs(Eo))
        assert Eb[i] == pytest.approx(Eo, rel=1e-10, abs=1e-10 * abs(Eo))
        assert np.allclose(Vb[i], Vo, rtol=1e-10, atol=1e-10 * np.abs(Vo).max())
    dev.close()


@pytest.mark.parametrize("case", ["special3b", "special4b", "stress-and-ener-2b1", "stress-and-ener-4b",
                                  "stress-and-ener-4b2", "h2o-3bcheby", "h2o-4bcheby", "run-ti-nonortho",
                                  "run-ti-ortho"])
def test_force_matches_reference_forceout(ch, tmp_path, case):
    """GPU forces with a reference params.txt reproduce chimes_md's forceout.txt
    (the h2o cases with fitted charges: Ewald terms included)."""
    print("case", case, flush=True)
    from conftest import md_case
    s, f, ref = md_case(ch, case, tmp_path)
    x, pd, ps = s.coeffs()
    dev = ch.Device(s)
    dev.set_penalty(pd, ps)
    F, E, V = dev.force(x, f)
    dev.close()
    assert np.max(np.abs(F - ref)) <= 1e-10 * max(np.abs(ref).max(), 1.0)
    # energy incl. per-atom offsets, as the reference run checked it: |E - E_cmp| / N
    # to 6 decimals, E_cmp = energy row of its force.txt
    import json
    import os
    from conftest import md_dir
    meta = json.load(open(os.path.join(md_dir(case), "expected.json")))
    if meta["abs_energy_error_per_atom"] is not None:
        assert round(abs(E - meta["forcetxt_energy"]) / f.natoms, 6) == pytest.approx(
            meta["abs_energy_error_per_atom"], abs=1e-12)
    from conftest import check_md_printout
    check_md_printout(meta, f.natoms, f.cell, E, V)


def test_force_list_capacity_overflow_repeats_exactly(ch, synth, oracle, tmp_path):
    """After a dilute configuration the context sizes the lists of the next
    call from its per-atom counts; a dense configuration overflows them, is
    repeated with exact sizes and gives the result of a fresh context (and
    the oracle's); the next call of the same size then fits."""
    text = synth.fm_setup(["C", "H", "N", "O"], [12.011, 1.008, 14.007, 15.999], 12, 6, 3, 0.8, 5.0, 1.4,
                          smax3=4.0, smax4=3.5)
    p = tmp_path / "fm_setup.in"
    open(p, "w").write(text)
    s = ch.Setup(str(p))
    rng = np.random.default_rng(6)
    cell = np.diag([20.0, 20.0, 20.0])
    dil = ch.PyFrame(60, synth.random_sequential(rng, 60, cell, 0.9), rng.integers(0, 4, 60).astype(np.int32), cell)
    den = ch.PyFrame(900, synth.random_sequential(rng, 900, cell, 0.9), rng.integers(0, 4, 900).astype(np.int32),
                     cell)
    x = rng.normal(size=s.model.ncols)
    dev = ch.Device(s)
    dev.force(x, dil)
    F1, E1, V1 = dev.force(x, den)
    assert dev.force_stats() == (2, 1, 0)
    F2, E2, V2 = dev.force(x, den)
    assert dev.force_stats() == (3, 1, 0)
    fresh = ch.Device(s)
    F0, E0, V0 = fresh.force(x, den)
    fresh.close()
    scale = np.abs(F0).max()
    for F, E, V in ((F1, E1, V1), (F2, E2, V2)):
        assert np.max(np.abs(F - F0)) <= 1e-12 * scale
        assert E == pytest.approx(E0, rel=1e-12, abs=1e-12 * abs(E0))
        assert np.allclose(V, V0, rtol=1e-12, atol=1e-12 * np.abs(V0).max())
    Fo, Eo, Vo = oracle.force(s, den, x)
    assert np.max(np.abs(F0 - Fo)) <= 1e-10 * scale
    assert E0 == pytest.approx(Eo, rel=1e-10, abs=1e-10 * abs(Eo))
    dev.close()


def test_force_graph_replay_matches_eager(ch, synth, oracle, tmp_path):
    """MD-like steps of one configuration (same cell and atom count, moved
    atoms, coefficients changed half way) are captured into a HIP graph and
    replayed; every replay equals the result of a context that stays eager
    (it alternates shapes), and the first one the oracle's."""
    text, frames, _ = synth.chno(nframes=1, nmin=300, nmax=300, seed=44)
    s, fr = synth_setup(ch, synth, tmp_path, (text, frames, _))
    f0 = fr[0]
    rng = np.random.default_rng(7)
    other = ch.PyFrame(50, synth.random_sequential(rng, 50, np.diag([12.0] * 3), 0.9),
                       rng.integers(0, 4, 50).astype(np.int32), np.diag([12.0] * 3))
    x = rng.normal(size=s.model.ncols)
    dev, ref = ch.Device(s), ch.Device(s)
    xyz = np.array(f0.xyz, dtype=np.float64).reshape(-1, 3)
    for step in range(8):
        if step == 5:
            x = x * 0.5 + rng.normal(size=s.model.ncols) * 0.1
        pos = xyz + rng.normal(0.0, 0.01, xyz.shape)
        f = ch.PyFrame(f0.natoms, pos, np.array(f0.types, dtype=np.int32), np.array(f0.cell, dtype=np.float64))
        F, E, V = dev.force(x, f)
        Fr, Er, Vr = ref.force(x, f)
        ref.force(x, other)  # the reference context never sees one shape twice in a row
        scale = np.abs(Fr).max()
        assert np.max(np.abs(F - Fr)) <= 1e-12 * scale, step
        assert E == pytest.approx(Er, rel=1e-12, abs=1e-12 * abs(Er)), step
        assert np.allclose(V, Vr, rtol=1e-12, atol=1e-12 * np.abs(Vr).max()), step
        if step == 0:
            Fo, Eo, Vo = oracle.force(s, f, x)
            assert np.max(np.abs(F - Fo)) <= 1e-10 * scale
    calls, repeats, replays = dev.force_stats()
    assert calls == 8 and replays >= 5, (calls, repeats, replays)
    assert ref.force_stats()[2] == 0
    dev.close()
    ref.close()
