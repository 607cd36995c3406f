"""LassoLars on the GPU (chimes_lsq.py --algorithm=lassolars: scikit-learn's
LassoLars, fit_intercept=False, fit_path=False) against scikit-learn 1.7.2 run
on the same design matrix (the GPU's A copied back), and the drop-in CLI
against the reference's lsq2 output.

Bars: the same number of path steps and set of non-zero coefficients, and
coefficients within 1e-5 of max|x| -- of scikit-learn's on the same A, and of
the reference's lsq2 output (a different scikit-learn build on the
reference's A).  The Gram matrices here have condition numbers near 1e13, and
a path of hundreds of FP64 steps carries the different summation orders of the
Gram, A^T b and the triangular solves into 1e-7..1e-6 of max|x|: measured
1.6e-7 (alpha 1e-3, 325 steps) and 1.0e-6 (alpha 1e-4, 591 steps) against
scikit-learn 1.7.2, and 2.9e-8 against the reference's lsq2 run, where two
scikit-learn builds differ by 2.3e-7."""
import os
import warnings

import numpy as np
import pytest

from conftest import water_setup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("alpha", [1e-3, 1e-4])
def test_lassolars_matches_sklearn(ch, tmp_path, alpha):
    from sklearn import linear_model
    print("alpha", alpha, flush=True)
    s = water_setup(ch, tmp_path, "h2o-4bcheby", nframes=4)
    fr = s.frames(str(tmp_path))
    dev = ch.Device(s)
    A, b = dev.assemble(fr)
    C, cb = s.constraint_rows()
    dev.append_rows(C)
    A = np.vstack([A, C])
    b = np.concatenate([b, cb])
    x, res = dev.fit_lassolars(b, alpha)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        reg = linear_model.LassoLars(alpha=alpha, fit_intercept=False, fit_path=False, max_iter=100000)
        reg.fit(A, b)
    xr = np.ravel(reg.coef_)
    d = np.abs(x - xr).max() / np.abs(xr).max()
    print("GPU %d steps (%.1f ms), sklearn %s; nnz %d / %d; max|x - x_sk|/max|x_sk| %.2e" % (
        res.iters, res.ms_solve, reg.n_iter_, np.count_nonzero(x), np.count_nonzero(xr), d))
    assert res.iters == reg.n_iter_
    assert np.array_equal(x != 0, xr != 0)
    assert d <= 1e-5
    assert res.nvars == int((np.abs(xr) > 1e-5).sum())
    y = A @ x
    assert res.rms == pytest.approx(np.sqrt(np.mean((b - y) ** 2)), rel=1e-10)
    dev.close()
