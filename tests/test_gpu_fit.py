"""GPU fit vs the reference's fitted parameters and vs a thin-SVD oracle."""
import numpy as np
import pytest

from conftest import golden_dir, synth_setup

pytestmark = pytest.mark.gpu


def thin_svd_fit(A, b, eps=1e-5):
    """chimes_lsq.py's estimator (src/chimes_lsq.py:187-214) with the thin SVD:
    x = V diag(1/s for s > eps*s_max) U^T b."""
    U, S, VT = np.linalg.svd(A, full_matrices=False)
    keep = S > eps * S.max()
    x = VT.T[:, keep] @ ((U[:, keep].T @ b) / S[keep])
    r = A @ x - b
    return x, np.sqrt((r @ r) / len(b)), int(keep.sum())


def test_nonorth2_fit_matches_reference(ch):
    d = golden_dir("nonorth2")
    s = ch.Setup(d + "/fm_setup.in")
    fr = s.frames()
    dev = ch.Device(s)
    A, b = dev.assemble(fr, copy=False)
    x, res = dev.fit(b, eps=1e-5, refine=3)
    xr = np.array([float(l.split()[1]) for l in open(d + "/test_suite_params.txt")])
    # run_test_suite.sh: parameters within 1e-9 absolute, forces within 1e-7
    assert np.max(np.abs(x - xr)) < 1e-9 * max(1.0, np.abs(xr).max())
    y = dev.matvec(x, len(b))
    yr = np.load(d + "/golden.npz")["force"]
    assert np.max(np.abs(np.round(y, 7) - yr) / np.maximum(np.abs(yr), 1e-6)) < 1e-5
    assert res.nvars == 12
    assert res.rms == pytest.approx(4.0724e-03, rel=1e-4)
    dev.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_carbon_fit_vs_thin_svd(ch, synth, oracle, tmp_path, prec):
    s, fr = synth_setup(ch, synth, tmp_path, synth.carbon(nframes=3, natoms=128, seed=31))
    dev = ch.Device(s)
    A, b = dev.assemble(fr, prec=prec)
    x, res = dev.fit(b, eps=1e-5, refine=4)
    Ao = oracle.assemble(s, fr)
    xo, rmso, nk = thin_svd_fit(Ao, b)
    if prec == 0:
        assert res.rms == pytest.approx(rmso, rel=1e-8)
    else:
        assert res.rms == pytest.approx(rmso, rel=1e-6)
    dev.close()
