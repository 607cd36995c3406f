"""The drop-in executable on 19 of the reference's LSQ tests end to end
(`chimes_lsq_gpu fm_setup.in --fit`, the SVD of chimes_lsq.py with its
default eps 1e-5): its standard output against the run's fm_setup.out under
compare.pl semantics (as rank 0 of 36, the runs' layout), b.txt byte for byte, the fit summary (number of fitted
variables equal, printed numbers within 1e-4), every coefficient against the
reference's test_suite_params.txt within 1e-8 of max|x|, and force.txt
(A x at 7 digits) within 1e-6 relative (floor 1e-3 of its largest value).

Cases: 2-, 3- and 4-body models, diagonal / full / first-N stress rows and
energy rows (stress-and-ener-*), special 3B/4B cutoffs, inverse-r transforms,
several trajectory files (MULTI) or one file of three trajectories, Ewald
charge columns with constraint rows (h2o-*)."""
import gzip
import hashlib
import json
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, PKG, WATER

pytestmark = pytest.mark.gpu

CASES_DIR = os.path.join(GOLDEN, "lsq", "cases")
CASES = sorted(os.listdir(CASES_DIR)) if os.path.isdir(CASES_DIR) else []


def _numbers(line):
    return [float(t) for t in re.findall(r"[-+]?\d+\.\d+e[-+]\d+", line)]


@pytest.mark.parametrize("case", CASES)
def test_cli_reference_case(ch, tmp_path, case):
    from test_ewald import x_from_params
    print("case", case, flush=True)
    d = os.path.join(CASES_DIR, case)
    meta = json.load(open(os.path.join(d, "meta.json")))
    for fn in os.listdir(d):
        if fn == "fm_setup.in" or fn.endswith(".dat"):
            shutil.copy(os.path.join(d, fn), tmp_path / fn)
    for t in meta["trajectories"]:
        src = os.path.join(WATER, "input.xyzf.gz") if t["water"] else os.path.join(d, t["name"] + ".gz")
        with gzip.open(src, "rb") as g:
            open(tmp_path / t["name"], "wb").write(g.read())
    exe = os.path.join(PKG, "chimes_lsq_gpu")
    # the reference runs behind the golden files used 36 MPI ranks
    r = subprocess.run([exe, "fm_setup.in", "--fit", "--no-A", "--ranks", "36"], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    # standard output = the run's fm_setup.out under compare.pl semantics
    from stdout_compare import STALE_MIN_DISTANCE_DECIMALS, differences, round_min_distances
    out = r.stdout
    if case in STALE_MIN_DISTANCE_DECIMALS:
        out = round_min_distances(out, STALE_MIN_DISTANCE_DECIMALS[case])
    bad = differences(gzip.open(os.path.join(d, "fm_setup.out.gz"), "rt").read(), out)
    assert not bad, "\n".join("ref %s: %r\nour %s: %r" % (a[0], a[1], b[0], b[1]) for a, b in bad)
    assert os.path.exists(tmp_path / "frames.0000.log") and os.path.exists(tmp_path / "frames.0035.log")
    data = open(tmp_path / "b.txt", "rb").read()
    assert (len(data), hashlib.sha256(data).hexdigest()) == (meta["b_bytes"], meta["b_sha256"])
    ours = open(tmp_path / "params.txt").read()
    summary = [l for l in ours.split("\n")[1:13] if l.startswith("!")]
    assert len(summary) == len(meta["summary"])
    for a, b in zip(summary, meta["summary"]):
        assert re.sub(r"[-+]?\d+\.\d+e[-+]\d+", "#", a) == re.sub(r"[-+]?\d+\.\d+e[-+]\d+", "#", b), (a, b)
        if "fitting vars" in a or "Number of" in a:
            assert a == b
        for u, v in zip(_numbers(a), _numbers(b)):
            assert u == pytest.approx(v, rel=1e-4), (a, b)
    s = ch.Setup(str(tmp_path / "fm_setup.in"))
    xo = x_from_params(ch, s, ours, tmp_path)
    xr = np.array(meta["x"])
    assert xo.shape == xr.shape
    dx = np.abs(xo - xr).max() / np.abs(xr).max()
    with gzip.open(os.path.join(d, "force.txt.gz"), "rt") as g:
        yr = np.array([float(v) for v in g.read().split()])
    y = np.loadtxt(tmp_path / "force.txt")
    dy = np.max(np.abs(y - yr) / np.maximum(np.abs(yr), 1e-3 * np.abs(yr).max()))
    print("%s: max|x - x_ref|/max|x_ref| %.2e, force.txt %.2e" % (case, dx, dy))
    assert dx <= 1e-8
    assert dy <= 1e-6
