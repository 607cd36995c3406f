"""Continuous-batching force server: every result equals the configuration's
result alone, whatever batch it was served in."""
import threading

import numpy as np
import pytest

from conftest import synth_setup

pytestmark = pytest.mark.gpu


def _requests(ch, synth, tmp_path, n, seed):
    text, frames, names = synth.chno(nframes=n, nmin=64, nmax=1500, seed=seed)
    return synth_setup(ch, synth, tmp_path, (text, frames, names))


def test_serve_results_equal_alone(ch, synth, tmp_path):
    s, fr = _requests(ch, synth, tmp_path, 12, 61)
    x = np.random.default_rng(7).normal(size=s.model.ncols)
    dev = ch.Device(s)
    alone = [dev.force(x, f) for f in fr]
    dev.close()
    srv = ch.Server(s, x, workers=2)
    tickets = [None] * len(fr)

    def submit(idx):
        for i in idx:
            tickets[i] = srv.submit(fr[i])

    th = [threading.Thread(target=submit, args=(range(k, len(fr), 3),)) for k in range(3)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for i, tk in enumerate(tickets):
        F, E, V, lat = srv.poll(tk)
        Fa, Ea, Va = alone[i]
        scale = max(np.abs(Fa).max(), 1e-300)
        assert np.max(np.abs(F - Fa)) <= 1e-12 * scale, i
        assert E == pytest.approx(Ea, rel=1e-12, abs=1e-12 * abs(Ea))
        assert np.max(np.abs(V - Va)) <= 1e-12 * max(np.abs(Va).max(), 1e-300)
        assert lat > 0.0
    batches, requests = srv.stats()
    assert requests == len(fr) and 1 <= batches <= len(fr)
    srv.close()


def test_serve_budget_admits_one_at_a_time(ch, synth, tmp_path):
    """A budget below every request's estimate serves each request alone."""
    s, fr = _requests(ch, synth, tmp_path, 5, 62)
    x = np.random.default_rng(8).normal(size=s.model.ncols)
    srv = ch.Server(s, x, workers=1, budget_bytes=1)
    tk = [srv.submit(f) for f in fr]
    res = [srv.poll(t) for t in tk]
    assert all(r is not None for r in res)
    assert srv.stats() == (len(fr), len(fr))
    srv.close()


def test_serve_rejects_bad_type(ch, synth, tmp_path):
    s, fr = _requests(ch, synth, tmp_path, 1, 63)
    srv = ch.Server(s, np.zeros(s.model.ncols), workers=1)
    bad = ch.PyFrame(fr[0].natoms, fr[0].xyz, np.full(fr[0].natoms, 7, dtype=np.int32), fr[0].cell)
    with pytest.raises(RuntimeError):
        srv.submit(bad)
    t = srv.submit(fr[0])
    assert srv.poll(t) is not None
    srv.close()
