"""The drop-in executable: `chimes_lsq_gpu fm_setup.in --fit` in a copy of the
reference's nonorth2 case writes the files chimes_lsq + chimes_lsq.py write,
with the reference's contents."""
import hashlib
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import PKG, colscaled_error, golden_dir, headers

pytestmark = pytest.mark.gpu


def _numbers(line):
    return [float(t) for t in re.findall(r"[-+]?\d+\.\d+e[-+]\d+", line)]


def test_cli_nonorth2_end_to_end(tmp_path):
    d = golden_dir("nonorth2")
    for fn in ["fm_setup.in", "input.xyzf"]:
        shutil.copy(os.path.join(d, fn), tmp_path / fn)
    exe = os.path.join(PKG, "chimes_lsq_gpu")
    r = subprocess.run([exe, "fm_setup.in", "--fit", "--ranks", "36"], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    import gzip
    from stdout_compare import differences
    bad = differences(gzip.open(os.path.join(d, "fm_setup.out.gz"), "rt").read(), r.stdout)
    assert not bad, bad
    g = np.load(os.path.join(d, "golden.npz"))
    A = np.loadtxt(tmp_path / "A.txt")
    assert A.shape == g["A"].shape
    assert colscaled_error(A, g["A"]) < 1e-10
    # A.txt text layout: 16-digit scientific, three-space separators on force rows
    first = open(tmp_path / "A.txt").readline()
    assert re.match(r"^ ?-?\d\.\d{16}e[-+]\d\d   ", first)
    assert np.array_equal(np.loadtxt(tmp_path / "b.txt"), g["b"])
    assert open(tmp_path / "dim.txt").read().split() == open(os.path.join(d, "dim.txt")).read().split()
    h = headers()["nonorth2"]
    for fn in ["params.header", "ff_groups.map"]:
        data = open(tmp_path / fn, "rb").read()
        assert hashlib.sha256(data).hexdigest() == h[fn]["sha256"], fn
    # params.txt: same lines as the reference's fit; numbers within the fit's tolerance
    ours = open(tmp_path / "params.txt").read().split("\n")[1:]
    ref = open(os.path.join(d, "params.txt")).read().split("\n")[1:]
    assert len(ours) == len(ref)
    for a, b in zip(ours, ref):
        na, nb = _numbers(a), _numbers(b)
        assert re.sub(r"[-+]?\d+\.\d+e[-+]\d+", "#", a) == re.sub(r"[-+]?\d+\.\d+e[-+]\d+", "#", b), (a, b)
        for u, v in zip(na, nb):
            assert u == pytest.approx(v, rel=1e-4, abs=1e-9 * 358.7), (a, b)
    y = np.loadtxt(tmp_path / "force.txt")
    assert np.max(np.abs(y - g["force"]) / np.maximum(np.abs(g["force"]), 1e-6)) < 1e-5


def test_cli_split_files_over_ranks(tmp_path):
    """SPLITFI true with SKPFRMS 1 over 3 ranks: the per-rank A.<rank>.txt
    concatenate to the A.txt of the same run without SPLITFI, the dim files
    tile the rows, and b.txt is the same; frames are in round-robin rank order
    (rank 0: frames 0, 3, 6, 9)."""
    d = golden_dir("nonorth2")
    text = open(os.path.join(d, "fm_setup.in")).read()
    text = text.rstrip("\n").replace("# ENDFILE #", "") + "\n# SPLITFI #\n\ttrue\n# SKPFRMS #\n\t1\n\n# ENDFILE #\n"
    runs = {}
    for split in (True, False):
        w = tmp_path / ("split" if split else "whole")
        w.mkdir()
        shutil.copy(os.path.join(d, "input.xyzf"), w / "input.xyzf")
        open(w / "fm_setup.in", "w").write(text if split else text.replace("\ttrue\n# SKPFRMS #", "\tfalse\n# SKPFRMS #"))
        exe = os.path.join(PKG, "chimes_lsq_gpu")
        r = subprocess.run([exe, "fm_setup.in", "--ranks", "3"], cwd=str(w), capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        runs[split] = w
    s, whole = runs[True], runs[False]
    assert not os.path.exists(s / "A.txt") and not os.path.exists(s / "dim.txt")
    parts = [open(s / ("A.%04d.txt" % r)).read() for r in range(3)]
    joined, ref = "".join(parts), open(whole / "A.txt").read()
    # same layout; values equal to rounding (LDS accumulation order differs from run to run)
    num = r"[-+]?\d+\.\d+e[-+]\d+"
    assert re.sub(num, "#", joined) == re.sub(num, "#", ref)
    a = np.array([float(v) for v in re.findall(num, joined)])
    b = np.array([float(v) for v in re.findall(num, ref)])
    assert a.shape == b.shape and np.max(np.abs(a - b)) <= 1e-13 * np.abs(b).max()
    total = int(open(whole / "dim.txt").read().split()[1])
    ncols = int(open(whole / "dim.txt").read().split()[0])
    start = 0
    for r in range(3):
        n = parts[r].count("\n")
        assert open(s / ("dim.%04d.txt" % r)).read() == "%d %d %d %d\n" % (ncols, start, start + n - 1, total)
        start += n
    assert start == total
    assert open(s / "b.txt").read() == open(whole / "b.txt").read()
    # rank order: frames 0 3 6 9 | 2 5 8 | 1 4 7 (frame i to rank r when (i + r) % 3 == 0)
    g = np.load(os.path.join(d, "golden.npz"))
    rows = len(g["b"]) // 10
    order = [0, 3, 6, 9, 2, 5, 8, 1, 4, 7]
    want = np.concatenate([g["b"][rows * i: rows * (i + 1)] for i in order])
    assert np.array_equal(np.loadtxt(s / "b.txt"), want)


def test_cli_hierarchical_fit(ch, synth, oracle, tmp_path):
    """HIERARC with EXCLUDE 1B (H) and 2B (O-H): A.txt is the oracle's A without
    those columns, the fit is the thin-SVD fit of that reduced A, and params.txt
    prints zeros for the excluded pair and energy offset."""
    text, frames, names = synth.water(nframes=6, nmol=27, o2=8, o3=4)
    text = text.replace("# ENDFILE #", "# HIERARC #\n\ttrue\nEXCLUDE 1B INTERACTION: 1\nH\n"
                                       "EXCLUDE 2B INTERACTION: 1\nO H\n\n# ENDFILE #")
    open(tmp_path / "fm_setup.in", "w").write(text)
    synth.write_xyzf(str(tmp_path / "input.xyzf"), frames, names)
    exe = os.path.join(PKG, "chimes_lsq_gpu")
    r = subprocess.run([exe, "fm_setup.in", "--fit"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    s = ch.Setup(str(tmp_path / "fm_setup.in"))
    fr = s.frames(str(tmp_path))
    keep = s.fit_columns()
    Ar = oracle.assemble(s, fr)[:, keep]
    A = np.loadtxt(tmp_path / "A.txt")
    assert A.shape == Ar.shape and colscaled_error(A, Ar) < 1e-10
    b = np.loadtxt(tmp_path / "b.txt")
    U, S, Vt = np.linalg.svd(Ar, full_matrices=False)
    k = S > 1e-5 * S.max()
    x = Vt[k].T @ ((U[:, k].T @ b) / S[k])
    y = Ar @ x
    yg = np.loadtxt(tmp_path / "force.txt")
    assert np.max(np.abs(yg - y)) <= 2e-6 * np.abs(y).max()
    txt = open(tmp_path / "params.txt").read()
    assert "EXCLD1B: 1 \nEXCLD2B: 2 \n" in txt
    blk = re.search(r"PAIRTYPE PARAMS: 2 O H\n\n((?:\s*\d+ +\S+\n)+)", txt).group(1)
    assert all(float(l.split()[1]) == 0.0 for l in blk.strip().split("\n"))
    assert re.search(r"ENERGY OFFSET 2 +(\S+)", txt).group(1) == "0.0000000000000e+00"
    rms = float(re.search(r"! RMS force error\s*=\s*(\S+)", txt).group(1))
    # params.txt prints the RMS with 5 significant digits (%11.4e)
    assert rms == pytest.approx(np.sqrt(np.mean((b - y) ** 2)), rel=1e-4)


def test_cli_skip_frames_stdout(tmp_path):
    """The reference's skip_frames case (FITCOUL water, SKPFRMS 1: round robin
    over 36 ranks): standard output against its fm_setup.out as rank 0 of 36
    (compare.pl semantics; that kept file prints the minimum distances with
    3 decimals)."""
    import gzip
    from conftest import WATER
    from stdout_compare import STALE_MIN_DISTANCE_DECIMALS, differences, round_min_distances
    d = os.path.join(WATER, "skip_frames")
    shutil.copy(os.path.join(d, "fm_setup.in"), tmp_path / "fm_setup.in")
    with gzip.open(os.path.join(WATER, "input.xyzf.gz"), "rb") as g:
        open(tmp_path / "input.xyzf", "wb").write(g.read())
    exe = os.path.join(PKG, "chimes_lsq_gpu")
    r = subprocess.run([exe, "fm_setup.in", "--no-A", "--ranks", "36"], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = round_min_distances(r.stdout, STALE_MIN_DISTANCE_DECIMALS["skip_frames"])
    bad = differences(gzip.open(os.path.join(d, "fm_setup.out.gz"), "rt").read(), out)
    assert not bad, bad
