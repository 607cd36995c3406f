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
    for k, (n, L) in enumerate([(40, [9.0, 9.5, 10.0]), (420, [24.0, 25.0, 26.0]), (90, [11.0, 11.0, 12.5])]):
        pos = rng.uniform(0, 1, (n, 3)) * np.array(L)
        frames.append(dict(natoms=n, xyz=pos.reshape(-1), types=rng.integers(0, 3, n).astype(np.int32),
                           cell=np.diag(L).reshape(-1), forces=rng.normal(0, 1, 3 * n), stress=np.zeros(6),
                           energy=0.0, index=k))
    text = synth.fm_setup(["C", "H", "O"], [12.0, 1.0, 16.0], 6, 2, 0, 0.5, 3.2, 1.4, nframes=3)
    text = fitcoul_text(text, ["+", "+", "-"], ["CC HH OO CH CO HO 1 1 1 1 1 1 0.0"] * 5)
    s, fr = synth_setup(ch, synth, tmp_path, (text, frames, ["C", "H", "O"]))
    a, rc, k = ch.ewald_params(420, np.diag([24.0, 25.0, 26.0]))
    assert rc < 12.0
    dev = ch.Device(s)
    A, b = dev.assemble(fr)
    Ao = oracle.assemble(s, fr)
    assert A.shape == Ao.shape and s.model.n_charge_cols == 6
    assert colscaled_error(A, Ao) < 1e-10
    P = s.model.ncols_short
    assert np.array_equal(np.signbit(A[:, P:]), np.signbit(Ao[:, P:]))
    dev.close()


def test_fit_with_constraint_rows(ch, oracle, tmp_path):
    """Constraint rows appended on the device enter the fit: the GPU fit of
    [A; C] equals a numpy thin-SVD fit of the oracle's [A; C] (eps 1e-5)."""
    s = water_setup(ch, tmp_path, "h2o-4bcheby", nframes=4)
    fr = s.frames(str(tmp_path))
    dev = ch.Device(s)
    A, b = dev.assemble(fr)
    C, cb = s.constraint_rows()
    dev.append_rows(C)
    Ao = np.vstack([oracle.assemble(s, fr), C])
    bo = np.concatenate([b, cb])
    x, res = dev.fit(bo)
    U, S, Vt = np.linalg.svd(Ao, full_matrices=False)
    keep = S > 1e-5 * S.max()
    xr = Vt[keep].T @ ((U[:, keep].T @ bo) / S[keep])
    y, yr = Ao @ x, Ao @ xr
    assert res.nvars == int(keep.sum())
    assert np.max(np.abs(y - yr)) <= 1e-6 * np.abs(yr).max()
    assert res.rms == pytest.approx(np.sqrt(np.mean((bo - yr) ** 2)), rel=1e-8)
    dev.close()


def _numbers(line):
    return [float(t) for t in re.findall(r"[-+]?\d+\.\d+e[-+]\d+", line)]


def test_cli_h2o_4bcheby_end_to_end(tmp_path):
    """`chimes_lsq_gpu fm_setup.in --fit` on the reference's h2o-4bcheby case
    (33 frames, 416 Chebyshev + 3 Ewald columns, 2 constraint rows)."""
    d = os.path.join(WATER, "h2o-4bcheby")
    shutil.copy(os.path.join(d, "fm_setup.in"), tmp_path / "fm_setup.in")
    import gzip
    with gzip.open(os.path.join(WATER, "input.xyzf.gz"), "rb") as g:
        open(tmp_path / "input.xyzf", "wb").write(g.read())
    exe = os.path.join(PKG, "chimes_lsq_gpu")
    r = subprocess.run([exe, "fm_setup.in", "--fit"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    g = np.load(os.path.join(d, "golden.npz"))
    A = np.loadtxt(tmp_path / "A.txt")
    assert A.shape == (9506, 419)
    assert colscaled_error(A[-154:], g["A"]) < 1e-10
    assert np.array_equal(np.loadtxt(tmp_path / "b.txt"), g["b"])
    assert open(tmp_path / "dim.txt").read().split() == ["419", "9506"]
    # params.txt: the reference's lines; numbers within the fit's tolerance
    ref = headers()["h2o-4bcheby"]
    ours = open(tmp_path / "params.txt").read().split("\n")
    summary = [l for l in ours[1:13] if l.startswith("!")]
    for a, b in zip(summary, ref["params_txt"]["summary"]):
        assert re.sub(r"[-+]?\d+\.\d+e[-+]\d+", "#", a) == re.sub(r"[-+]?\d+\.\d+e[-+]\d+", "#", b), (a, b)
        for u, v in zip(_numbers(a), _numbers(b)):
            assert u == pytest.approx(v, rel=1e-4), (a, b)
        if "fitting vars" in a:
            assert a == b
    xs = [float(l.split()[-1]) for l in ours if l.startswith("q_")]
    xr = np.array(ref["x"])
    assert len(xs) == 3
    assert np.allclose(xs, xr[416:419], rtol=1e-4, atol=1e-6 * np.abs(xr).max())
    y = np.loadtxt(tmp_path / "force.txt")
    rel = np.max(np.abs(y - g["force"]) / np.maximum(np.abs(g["force"]), 1e-3 * np.abs(g["force"]).max()))
    print("force.txt max rel diff %.2e" % rel)
    assert rel < 1e-5
