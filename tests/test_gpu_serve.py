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


def test_serve_mixed_batch_equals_alone(ch, synth, tmp_path):
    """Requests held in the queue and then released are served in shared
    batches (fewer batches than requests), and every result equals the
    configuration's result alone."""
    s, fr = _requests(ch, synth, tmp_path, 10, 64)
    x = np.random.default_rng(9).normal(size=s.model.ncols)
    dev = ch.Device(s)
    alone = [dev.force(x, f) for f in fr]
    dev.close()
    srv = ch.Server(s, x, workers=1, budget_bytes=1 << 34, max_batch_atoms=1 << 20)
    srv.hold(True)
    tickets = [srv.submit(f) for f in fr]
    srv.hold(False)
    for i, tk in enumerate(tickets):
        F, E, V, lat = srv.poll(tk)
        Fa, Ea, Va = alone[i]
        scale = max(np.abs(Fa).max(), 1e-300)
        assert np.max(np.abs(F - Fa)) <= 1e-12 * scale, i
        assert E == pytest.approx(Ea, rel=1e-12, abs=1e-12 * abs(Ea))
        assert np.max(np.abs(V - Va)) <= 1e-12 * max(np.abs(Va).max(), 1e-300)
    batches, requests = srv.stats()
    assert requests == len(fr)
    assert batches == 1, batches  # one worker, everything admitted at once
    srv.close()


def test_serve_failure_stays_with_its_request(ch, synth, tmp_path):
    """A request that fails on its own (a cell thinner than cutoff / 7, which
    the force path refuses) in a shared batch reports its error; the other
    requests of that batch get their results alone."""
    s, fr = _requests(ch, synth, tmp_path, 4, 65)
    x = np.random.default_rng(10).normal(size=s.model.ncols)
    dev = ch.Device(s)
    alone = [dev.force(x, f) for f in fr]
    dev.close()
    thin = np.diag([0.5, 20.0, 20.0])
    bad = ch.PyFrame(2, np.array([[0.1, 1.0, 1.0], [0.2, 5.0, 5.0]]), np.zeros(2, dtype=np.int32), thin)
    srv = ch.Server(s, x, workers=1, budget_bytes=1 << 34, max_batch_atoms=1 << 20)
    srv.hold(True)
    tickets = [srv.submit(f) for f in fr[:2]] + [srv.submit(bad)] + [srv.submit(f) for f in fr[2:]]
    srv.hold(False)
    res = [srv.poll_status(t) for t in tickets]
    st, msg = res[2][0], res[2][1]
    assert st != 0 and "cell width" in msg, (st, msg)
    good = res[:2] + res[3:]
    for (st, msg, F, E, V), (Fa, Ea, Va) in zip(good, alone):
        assert st == 0, msg
        assert np.max(np.abs(F - Fa)) <= 1e-12 * max(np.abs(Fa).max(), 1e-300)
        assert E == pytest.approx(Ea, rel=1e-12, abs=1e-12 * abs(Ea))
    assert srv.stats() == (1, 5)
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
