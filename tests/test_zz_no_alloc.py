"""Allocation-free force entry point (cg_force_noalloc) against cg_force.

Two calls in a row on one context, the second with a smaller configuration
(the scratch array keeps the first call's contents), a model with 4-body
terms and a charged model; and a configuration too large for the scratch,
refused with a message instead of running."""
import numpy as np
import pytest

from conftest import synth_setup

pytestmark = pytest.mark.gpu


def _same(a, b):
    F, E, V = a
    F0, E0, V0 = b
    scale = max(np.abs(F0).max(), 1e-300)
    assert np.max(np.abs(F - F0)) <= 1e-12 * scale
    assert E == pytest.approx(E0, rel=1e-12, abs=1e-12 * abs(E0))
    assert np.max(np.abs(V - V0)) <= 1e-12 * max(np.abs(V0).max(), 1e-300)


def test_noalloc_matches_force_two_calls(ch, synth, oracle, tmp_path):
    text, frames, names = synth.chno(nframes=2, nmin=900, nmax=900, seed=71)
    s, fr = synth_setup(ch, synth, tmp_path, (text, frames, names))
    small = synth.chno(nframes=1, nmin=150, nmax=150, seed=72)[1][0]
    fs = ch.PyFrame(small["natoms"], small["xyz"], small["types"], small["cell"])
    x = np.random.default_rng(13).normal(size=s.model.ncols)
    dev = ch.Device(s)
    ref_big = dev.force(x, fr[0])
    ref_small = dev.force(x, fs)
    na = ch.Device(s)
    print("noalloc: 900-atom C/H/N/O call", flush=True)
    _same(na.force_noalloc(x, fr[0]), ref_big)
    print("noalloc: 150-atom C/H/N/O call", flush=True)
    _same(na.force_noalloc(x, fs), ref_small)
    # the context's ordinary path still works and agrees afterwards
    _same(na.force(x, fs), ref_small)
    _same(na.force_noalloc(x, fr[0]), ref_big)
    Fo, Eo, Vo = oracle.force(s, fs, x)
    F, E, V = na.force_noalloc(x, fs)
    assert np.max(np.abs(F - Fo)) <= 1e-10 * np.abs(Fo).max()
    na.close()
    dev.close()


def test_noalloc_carbon_order5(ch, synth, tmp_path):
    s, fr = synth_setup(ch, synth, tmp_path, synth.carbon(nframes=2, natoms=256, seed=73))
    x = np.random.default_rng(14).normal(size=s.model.ncols)
    dev = ch.Device(s)
    refs = [dev.force(x, f) for f in fr]
    dev.close()
    na = ch.Device(s)
    for f, r in zip(fr, refs):
        print("noalloc: 256-atom carbon call", flush=True)
        _same(na.force_noalloc(x, f), r)
    na.close()


def test_noalloc_refuses_too_large(ch, synth, tmp_path):
    text, frames, names = synth.chno(nframes=1, nmin=100, nmax=100, seed=74)
    s, fr = synth_setup(ch, synth, tmp_path, (text, frames, names))
    # 28^3 atoms on a 1.3 A simple-cubic lattice: its neighbour lists alone
    # exceed the scratch
    n = 28
    g = np.arange(n) * 1.3
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    cell = np.diag([n * 1.3] * 3)
    big = ch.PyFrame(len(pos), pos, np.zeros(len(pos), dtype=np.int32), cell)
    x = np.random.default_rng(15).normal(size=s.model.ncols)
    na = ch.Device(s)
    print("noalloc: %d-atom configuration, expected refusal" % big.natoms, flush=True)
    with pytest.raises(RuntimeError, match="scratch"):
        na.force_noalloc(x, big)
    # the context and the scratch are usable afterwards
    F, E, V = na.force_noalloc(x, fr[0])
    dev = ch.Device(s)
    _same((F, E, V), dev.force(x, fr[0]))
    dev.close()
    na.close()
