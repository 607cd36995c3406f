"""FITCOUL on the GPU: the Ewald charge columns of A against the oracle and
the reference's golden rows, the constraint rows in the fit, and the drop-in
executable on the reference's h2o-4bcheby case.

Bars: FP64 A within 1e-10 column-scaled relative error (with -0.0 on the
columns of pairs without the atom's type, as the reference prints them);
FP32 A within 5e-3 column-scaled and 3e-5 of the column's largest magnitude.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import PKG, WATER, colscaled_error, fitcoul_text, headers, synth_setup, water_setup

pytestmark = pytest.mark.gpu


def _fp32_ok(A32, Ao):
    colmax = np.maximum(np.abs(Ao).max(axis=0), 1e-300)
    return colscaled_error(A32, Ao) < 5e-3 and float((np.abs(A32 - Ao) / colmax).max()) < 3e-5


@pytest.mark.parametrize("name", ["h2o-4bcheby", "h2o-4bcheby-small"])
def test_water_cases_match_oracle_and_golden(ch, oracle, tmp_path, name):
    print("case", name, flush=True)
    s = water_setup(ch, tmp_path, name)
    fr = s.frames(str(tmp_path))
    batch = [fr[0], fr[7], fr[-1]]
    dev = ch.Device(s)
    A, b = dev.assemble(batch)
    Ao = oracle.assemble(s, batch)
    assert colscaled_error(A, Ao) < 1e-10
    P = s.model.ncols_short
    assert np.array_equal(np.signbit(A[:, P:]), np.signbit(Ao[:, P:]))
    G = np.load(os.path.join(WATER, name, "golden.npz"))["A"]
    assert colscaled_error(A[-152:], G[:152]) < 1e-10
    A32, _ = dev.assemble(batch, prec=ch.CG_FP32)
    assert _fp32_ok(A32, Ao)
    dev.close()


def test_three_types_ragged_frames(ch, synth, oracle, tmp_path):
    """C/H/O-like 3-type frames of different sizes and cells in one batch,
    one large enough that the real-space cutoff is below half the cell
    (uncapped Fincham parameters), one small (capped)."""
    rng = np.random.default_rng(11)
    frames = []
    for k, (n, L) in enumerate([(40, [9.0, 9.5, 10.0]), (1200, [20.0, 20.0, 20.0]), (90, [11.0, 11.0, 12.5])]):
        pos = rng.uniform(0, 1, (n, 3)) * np.array(L)
        frames.append(dict(natoms=n, xyz=pos.reshape(-1), types=rng.integers(0, 3, n).astype(np.int32),
                           cell=np.diag(L).reshape(-1), forces=rng.normal(0, 1, 3 * n), stress=np.zeros(6),
                           energy=0.0, index=k))
    text = synth.fm_setup(["C", "H", "O"], [12.0, 1.0, 16.0], 6, 2, 0, 0.5, 3.2, 1.4, nframes=3)
    text = fitcoul_text(text, ["+", "+", "-"], ["CC HH OO CH CO HO 1 1 1 1 1 1 0.0"] * 5)
    s, fr = synth_setup(ch, synth, tmp_path, (text, frames, ["C", "H", "O"]))
    a, rc, k = ch.ewald_params(1200, np.diag([20.0, 20.0, 20.0]))
    assert rc < 10.0
    dev = ch.Device(s)
    A, b = dev.assemble(fr)
    Ao = oracle.assemble(s, fr)
    assert A.shape == Ao.shape and s.model.n_charge_cols == 6
    assert colscaled_error(A, Ao) < 1e-10
    P = s.model.ncols_short
    assert np.array_equal(np.signbit(A[:, P:]), np.signbit(Ao[:, P:]))
    dev.close()


def test_fit_with_constraint_rows(ch, oracle, tmp_path):
    """Constraint rows appended on the device enter the fit.  Against a numpy
    thin-SVD fit of the oracle's [A; C] (eps 1e-5; 85 of 419 singular values
    kept, the spectrum dense around the cut):
    * CG_FIT_QR (SVD of A through its LQ factor): A x within 1e-8 of max|A x|,
      RMS within 1e-10;
    * the FP64 Gram path: same kept count, A x within 1e-6 of max|A x|, RMS
      within 1e-7 -- eigenvectors of A^T A mix directions whose singular values
      differ by ~1e-6 s_max near the cut at ~1e-16 s_max^2 / (s_i^2 - s_j^2)."""
    s = water_setup(ch, tmp_path, "h2o-4bcheby", nframes=4)
    fr = s.frames(str(tmp_path))
    dev = ch.Device(s)
    A, b = dev.assemble(fr)
    C, cb = s.constraint_rows()
    dev.append_rows(C)
    Ao = np.vstack([oracle.assemble(s, fr), C])
    bo = np.concatenate([b, cb])
    U, S, Vt = np.linalg.svd(Ao, full_matrices=False)
    keep = S > 1e-5 * S.max()
    xr = Vt[keep].T @ ((U[:, keep].T @ bo) / S[keep])
    yr = Ao @ xr
    rms_r = np.sqrt(np.mean((bo - yr) ** 2))
    for mode, ytol, rtol in [(ch.CG_FIT_QR, 1e-8, 1e-10), (ch.CG_GRAM_FP64, 1e-6, 1e-7)]:
        x, res = dev.fit(bo, gram=mode)
        y = Ao @ x
        print("mode %d: nvars %d/%d, max |y - yr| / max|yr| %.2e, rms rel diff %.2e, %.1f + %.1f ms" % (
            mode, res.nvars, int(keep.sum()), np.max(np.abs(y - yr)) / np.abs(yr).max(), abs(res.rms / rms_r - 1),
            res.ms_gram, res.ms_solve))
        assert res.nvars == int(keep.sum())
        assert np.max(np.abs(y - yr)) <= ytol * np.abs(yr).max()
        assert res.rms == pytest.approx(rms_r, rel=rtol)
    dev.close()


def _numbers(line):
    return [float(t) for t in re.findall(r"[-+]?\d+\.\d+e[-+]\d+", line)]


def test_cli_h2o_4bcheby_end_to_end(ch, tmp_path):
    """`chimes_lsq_gpu fm_setup.in --fit` on the reference's h2o-4bcheby case
    (33 frames, 416 Chebyshev + 3 Ewald columns, 2 constraint rows)."""
    d = os.path.join(WATER, "h2o-4bcheby")
    shutil.copy(os.path.join(d, "fm_setup.in"), tmp_path / "fm_setup.in")
    import gzip
    with gzip.open(os.path.join(WATER, "input.xyzf.gz"), "rb") as g:
        open(tmp_path / "input.xyzf", "wb").write(g.read())
    exe = os.path.join(PKG, "chimes_lsq_gpu")
    r = subprocess.run([exe, "fm_setup.in", "--fit", "--ranks", "36"], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    # standard output: the run's fm_setup.out (rank 0 of 36) under compare.pl semantics
    from stdout_compare import differences
    bad = differences(gzip.open(os.path.join(d, "fm_setup.out.gz"), "rt").read(), r.stdout)
    assert not bad, bad
    g = np.load(os.path.join(d, "golden.npz"))
    A = np.loadtxt(tmp_path / "A.txt")
    assert A.shape == (9506, 419)
    assert colscaled_error(A[-154:], g["A"]) < 1e-10
    assert np.array_equal(np.loadtxt(tmp_path / "b.txt"), g["b"])
    assert open(tmp_path / "dim.txt").read().split() == ["419", "9506"]
    # params.txt: the reference's summary; every coefficient against the
    # reference's fitted x printed by the same printer (pinned byte for byte by
    # the params.txt tests), within 1e-8 of max|x|
    ref = headers()["h2o-4bcheby"]
    ours = open(tmp_path / "params.txt").read().split("\n")
    summary = [l for l in ours[1:13] if l.startswith("!")]
    for a, b in zip(summary, ref["params_txt"]["summary"]):
        assert re.sub(r"[-+]?\d+\.\d+e[-+]\d+", "#", a) == re.sub(r"[-+]?\d+\.\d+e[-+]\d+", "#", b), (a, b)
        for u, v in zip(_numbers(a), _numbers(b)):
            assert u == pytest.approx(v, rel=1e-4), (a, b)
        if "fitting vars" in a:
            assert a == b
    import ctypes
    chm = ch
    s = chm.Setup(str(tmp_path / "fm_setup.in"))
    xr = np.array(ref["x"])
    res = chm.FitResult()
    gold = tmp_path / "gold"
    gold.mkdir()
    chm.check(chm.lib().cg_write_params(s.h, str(gold).encode(), chm.dptr(xr), len(xr), 9506, ctypes.byref(res),
                                        None, None))
    theirs = open(gold / "params.txt").read().split("\n")
    assert len(theirs) == len(ours)
    tol = 1e-8 * np.abs(xr).max()
    worst = 0.0
    for a, b in list(zip(ours, theirs))[13:]:
        assert re.sub(r"[-+]?\d+\.\d+e[-+]\d+", "#", a) == re.sub(r"[-+]?\d+\.\d+e[-+]\d+", "#", b), (a, b)
        for u, v in zip(_numbers(a), _numbers(b)):
            worst = max(worst, abs(u - v))
            assert abs(u - v) <= tol, (a, b)
    print("largest coefficient difference %.2e (max |x| %.4e)" % (worst, np.abs(xr).max()))
    # force.txt (A x at 7 digits): within the print resolution
    y = np.loadtxt(tmp_path / "force.txt")
    rel = np.max(np.abs(y - g["force"]) / np.maximum(np.abs(g["force"]), 1e-3 * np.abs(g["force"]).max()))
    print("force.txt max rel diff %.2e" % rel)
    assert rel < 1e-6


@pytest.mark.parametrize("fn,args,xtol", [
    ("params.svd.weight.txt", ["--weights", "weights.txt"], 1e-8),
    ("params.svd.1.0e-04.txt", ["--eps", "1.0e-04"], 1e-8),
    # sklearn's Ridge solves (A^T A + alpha I) x = A^T b by Cholesky: relative
    # error up to ~1e-16 s_max^2 / alpha = 4e-4 here (s_max 6.5e4, alpha 1e-3)
    ("params.ridge.1.0e-03.txt", ["--algorithm", "ridge", "--alpha", "1.0e-03"], 1e-3),
    # a LARS path of ~560 steps; scikit-learn 1.7.2 on the oracle's A is within 2.3e-7
    ("params.lassolars.1.0e-04.txt", ["--algorithm", "lassolars", "--alpha", "1.0e-04"], 1e-5)])
def test_cli_lsq2_weights_eps_ridge(ch, tmp_path, fn, args, xtol):
    """chimes_lsq.py's weighted SVD, eps 1e-4 SVD and ridge runs of the
    reference's lsq2 case (h2o-4bcheby setup) through the drop-in CLI: the
    summary lines to their printed digits and every coefficient against the
    reference's (within xtol of max|x|)."""
    import ctypes
    import gzip
    from test_ewald import _lsq2_text, _summary, x_from_params
    print("case", fn, flush=True)
    s = water_setup(ch, tmp_path, "h2o-4bcheby")
    with gzip.open(os.path.join(WATER, "lsq2", "weights.txt.gz"), "rb") as g:
        open(tmp_path / "weights.txt", "wb").write(g.read())
    exe = os.path.join(PKG, "chimes_lsq_gpu")
    r = subprocess.run([exe, "fm_setup.in", "--fit", "--no-A"] + args, cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    gold = _lsq2_text(fn)
    ours = open(tmp_path / "params.txt").read()
    vg, vo = _summary(gold), _summary(ours)
    assert vg.keys() == vo.keys()
    for k in vg:
        if k == "number of fitting vars" or k.startswith("Number"):
            assert vo[k] == vg[k], k
        else:
            assert float(vo[k]) == pytest.approx(float(vg[k]), rel=1e-4), k
    assert [l for l in ours.split("\n") if "weighting" in l] == [l for l in gold.split("\n") if "weighting" in l]
    xg = x_from_params(ch, s, gold, tmp_path)
    xo = x_from_params(ch, s, ours, tmp_path)
    d = np.abs(xo - xg).max() / np.abs(xg).max()
    print("max |x - x_ref| / max|x_ref| = %.2e" % d)
    assert d <= xtol


@pytest.mark.parametrize("case", ["h2o-3bcheby", "h2o-4bcheby"])
def test_md_ewald_force_energy_virial_vs_oracle(ch, oracle, tmp_path, case):
    """Force calls on a model fitted with charges (chimes_md's Ewald terms):
    forces, energy and virial against the oracle (1e-10), alone and batched
    with the same configuration shifted and a smaller 48-atom cut (net charge
    0), repeated calls (no graph capture for these models)."""
    from conftest import md_case
    s, f, ref = md_case(ch, case, tmp_path)
    x, pd, ps = s.coeffs()
    dev = ch.Device(s)
    dev.set_penalty(pd, ps)
    F, E, V = dev.force(x, f)
    Fo, Eo, Vo = oracle.force(s, f, x, penalty_scale=ps, penalty_dist=pd)
    assert np.max(np.abs(F - Fo)) <= 1e-10 * np.abs(Fo).max()
    assert E == pytest.approx(Eo, rel=1e-10)
    assert np.allclose(V, Vo, rtol=1e-10, atol=1e-10 * np.abs(Vo).max())
    f2 = ch.PyFrame(f.natoms, f.xyz + 0.31, f.types, f.cell)
    sel = np.concatenate([np.where(f.types == 0)[0][:16], np.where(f.types == 1)[0][:32]])
    f3 = ch.PyFrame(48, f.xyz.reshape(-1, 3)[sel].reshape(-1), f.types[sel], np.diag([9.0, 9.5, 10.0]))
    Fb, Eb, Vb = dev.force_batch(x, [f, f2, f3])
    for i, fr in enumerate([f, f2, f3]):
        Fo, Eo, Vo = oracle.force(s, fr, x, penalty_scale=ps, penalty_dist=pd)
        assert np.max(np.abs(Fb[i] - Fo)) <= 1e-10 * np.abs(Fo).max(), i
        assert Eb[i] == pytest.approx(Eo, rel=1e-10), i
        assert np.allclose(Vb[i], Vo, rtol=1e-10, atol=1e-10 * np.abs(Vo).max()), i
    for _ in range(3):
        F2, E2, V2 = dev.force(x, f)
        assert np.max(np.abs(F2 - F)) <= 1e-12 * np.abs(F).max() and E2 == pytest.approx(E, rel=1e-12)
    dev.close()
