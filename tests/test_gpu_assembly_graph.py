"""Repeated assemblies on one context: list capacities from per-atom counts,
graph capture and replay, exact repeat on overflow, and no stale entries of A
(every entry is written, buckets without clusters as zeros).

Bar: a replayed or capacity-sized assembly equals the assembly of a context
that stays eager to 1e-12 column-scaled (the LDS sums run in a different
order from run to run), and the oracle's to 1e-10.
"""
import numpy as np
import pytest

from conftest import colscaled_error, synth_setup

pytestmark = pytest.mark.gpu


def _moved(ch, f, rng, sigma=0.01):
    xyz = np.array(f.xyz, dtype=np.float64).reshape(-1, 3)
    return ch.PyFrame(f.natoms, xyz + rng.normal(0.0, sigma, xyz.shape), np.array(f.types, dtype=np.int32),
                      np.array(f.cell, dtype=np.float64), np.array(f.forces), np.array(f.stress), f.energy, f.index)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_assembly_graph_replay_matches_eager(ch, synth, oracle, tmp_path, prec):
    """A mixed C/H/N/O batch assembled six times with moved atoms: the third
    call is captured and the later ones replay it; each result equals an
    eager context's (graphs off), with other work on the device between the
    replays."""
    print("prec", prec, flush=True)
    P = ch.CG_FP64 if prec == "fp64" else ch.CG_FP32
    gen = synth.chno(nframes=6, nmin=40, nmax=700, seed=61)
    s, fr = synth_setup(ch, synth, tmp_path, gen)
    rng = np.random.default_rng(61)
    dev, ref = ch.Device(s), ch.Device(s)
    ref.assembly_graphs(False)
    x = rng.normal(0, 1e-3, s.model.ncols)
    for step in range(6):
        batch = [_moved(ch, f, rng) for f in fr]
        A, b = dev.assemble(batch, prec=P)
        Ar, br = ref.assemble(batch, prec=P)
        assert np.array_equal(b, br)
        err = colscaled_error(A, Ar)
        print("step", step, "vs eager %.2e" % err, dev.assembly_stats(), flush=True)
        assert err < (1e-12 if prec == "fp64" else 1e-6), step
        if step == 0:
            assert colscaled_error(A, oracle.assemble(s, batch)) < (1e-10 if prec == "fp64" else 5e-3)
        dev.force(x, batch[0])  # other device work between the replays
        dev.synchronize()
    calls, repeats, replays, captures = dev.assembly_stats()
    assert calls == 6 and captures >= 1 and replays >= 2, (calls, repeats, replays, captures)
    assert ref.assembly_stats()[2] == 0
    dev.close()
    ref.close()


def test_assembly_capacity_overflow_repeats_exactly(ch, synth, oracle, tmp_path):
    """After a dilute batch the next assembly sizes its lists from the dilute
    per-atom counts; a dense batch of the same atom count overflows them, is
    repeated with exact sizes and equals the oracle."""
    text = synth.fm_setup(["C", "H", "N", "O"], [12.011, 1.008, 14.007, 15.999], 12, 6, 3, 0.8, 5.0, 1.4,
                          smax3=4.0, smax4=3.5)
    p = tmp_path / "fm_setup.in"
    open(p, "w").write(text)
    s = ch.Setup(str(p))
    rng = np.random.default_rng(62)
    big, small = np.diag([30.0] * 3), np.diag([16.0] * 3)
    dil = ch.PyFrame(300, synth.random_sequential(rng, 300, big, 0.9), rng.integers(0, 4, 300).astype(np.int32),
                     big, rng.normal(size=900), np.zeros(9), 0.0, 0)
    den = ch.PyFrame(300, synth.random_sequential(rng, 300, small, 0.9), rng.integers(0, 4, 300).astype(np.int32),
                     small, rng.normal(size=900), np.zeros(9), 0.0, 0)
    dev = ch.Device(s)
    dev.assemble([dil])
    A1, _ = dev.assemble([den])
    calls, repeats, _, _ = dev.assembly_stats()
    assert calls == 2 and repeats == 1, dev.assembly_stats()
    A2, _ = dev.assemble([den])
    assert dev.assembly_stats()[1] == 1
    Ao = oracle.assemble(s, [den])
    assert colscaled_error(A1, Ao) < 1e-10
    assert colscaled_error(A2, Ao) < 1e-10
    dev.close()


def test_assembly_leaves_no_stale_entries(ch, synth, oracle, tmp_path):
    """A large mixed batch, then a smaller one on the same context (same
    buffers, not cleared): the second A equals the oracle's, including the
    zeros of cluster types its frames do not contain."""
    gen = synth.chno(nframes=4, nmin=200, nmax=600, seed=63)
    s, fr = synth_setup(ch, synth, tmp_path, gen)
    rng = np.random.default_rng(63)
    cell = np.diag([14.0] * 3)
    # carbon and hydrogen only: every cluster type with N or O is empty
    ch_only = ch.PyFrame(150, synth.random_sequential(rng, 150, cell, 0.9), rng.integers(0, 2, 150).astype(np.int32),
                         cell, rng.normal(size=450), np.zeros(9), 0.0, 0)
    dev = ch.Device(s)
    dev.assemble(fr)
    A, _ = dev.assemble([ch_only])
    Ao = oracle.assemble(s, [ch_only])
    assert colscaled_error(A, Ao) < 1e-10
    assert np.all(A[Ao == 0] == 0)
    dev.close()


@pytest.mark.parametrize("case", ["carbon", "chno"])
def test_deterministic_assembly_is_bitwise_stable(ch, synth, oracle, tmp_path, case):
    """With cg_assembly_deterministic every run of a batch gives the same bits
    (repeated calls on one context, eager and replayed, and a fresh context),
    and the oracle's A to 1e-10."""
    print("case", case, flush=True)
    gen = (synth.carbon(nframes=2, natoms=200, seed=64) if case == "carbon"
           else synth.chno(nframes=4, nmin=60, nmax=900, seed=65))
    s, fr = synth_setup(ch, synth, tmp_path, gen)
    runs = []
    dev = ch.Device(s)
    dev.assembly_deterministic(True)
    for _ in range(4):
        runs.append(dev.assemble(fr)[0])
    assert dev.assembly_stats()[2] >= 1  # the last call replays the captured graph
    other = ch.Device(s)
    other.assembly_deterministic(True)
    runs.append(other.assemble(fr)[0])
    for A in runs[1:]:
        assert np.array_equal(A, runs[0])
    assert colscaled_error(runs[0], oracle.assemble(s, fr)) < 1e-10
    dev.close()
    other.close()


def test_same_sizes_new_species(ch, synth, oracle, tmp_path):
    """Batches of the same frame sizes whose atoms change species (other
    clusters, other list sizes) assembled on one context (graphs on, each
    shape twice) equal a fresh eager context's result (1e-12 column-scaled)
    and the oracle's (1e-11 of each column's largest magnitude: these
    150-300-atom frames hold entries near 1e-3 of their column's largest whose
    FP64 sums differ from the oracle's at ~1.3e-10 column-scaled)."""
    gen = synth.chno(nframes=3, nmin=150, nmax=300, seed=72)
    s, fr = synth_setup(ch, synth, tmp_path, gen)
    rng = np.random.default_rng(72)
    dev = ch.Device(s)
    for trial in range(3):
        batch = []
        for f in fr:
            t = np.array(f.types, dtype=np.int32)
            if trial == 1:  # most atoms carbon
                t = np.where(rng.random(t.size) < 0.8, 0, t).astype(np.int32)
            elif trial == 2:  # species permuted
                t = rng.permutation(t).astype(np.int32)
            batch.append(ch.PyFrame(f.natoms, np.array(f.xyz), t, np.array(f.cell), np.array(f.forces),
                                    np.array(f.stress), f.energy, f.index))
        for rep in range(2):
            A, b = dev.assemble(batch)
        fresh = ch.Device(s)
        fresh.assembly_graphs(False)
        Af, _ = fresh.assemble(batch)
        fresh.close()
        e_fresh = colscaled_error(A, Af)
        Ao = oracle.assemble(s, batch)
        e_orc = float((np.abs(A - Ao) / np.maximum(np.abs(Ao).max(axis=0), 1e-300)).max())
        print("trial", trial, "vs fresh %.2e vs oracle %.2e" % (e_fresh, e_orc), dev.assembly_stats(), flush=True)
        assert e_fresh < 1e-12, trial
        assert e_orc < 1e-11, trial
    assert dev.assembly_stats()[3] >= 1  # graphs were captured along the way
    dev.close()
