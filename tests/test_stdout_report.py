"""Standard output of the drop-in executable against the reference runs'
fm_setup.out (compare.pl semantics, tests/stdout_compare.py): the part printed
while the input is read and the cluster tables are built (input echo, cluster
power tables, parameter counts), here without a GPU, for every reference
case.  The whole output of each run, with its per-frame lines and closing
statistics, is compared on the GPU (test_gpu_cases.py)."""
import ctypes
import gzip
import os
import shutil

import pytest

from conftest import GOLDEN, WATER
from stdout_compare import differences

CASES_DIR = os.path.join(GOLDEN, "lsq", "cases")
CASES = sorted(os.listdir(CASES_DIR))
WATER_CASES = ["h2o-2bcheby", "h2o-4bcheby", "h2o-4bcheby-small", "skip_frames"]
SETUP_END = "Setting up the matrices for A, Coulomb forces, and overbonding..."


def _report_setup(ch, path, tmp_path):
    libc = ctypes.CDLL(None)
    libc.fopen.restype = ctypes.c_void_p
    libc.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    libc.fclose.argtypes = [ctypes.c_void_p]
    L = ch.lib()
    L.cg_report_setup.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    s = ch.Setup(path)
    out = os.path.join(str(tmp_path), "report.txt")
    fp = libc.fopen(out.encode(), b"w")
    ch.check(L.cg_report_setup(s.h, fp))
    libc.fclose(fp)
    return open(out).read()


def _golden_prefix(path):
    text = gzip.open(path, "rt").read()
    k = text.index(SETUP_END) + len(SETUP_END) + 1
    return text[:k]


@pytest.mark.parametrize("case", CASES + ["water/" + c for c in WATER_CASES])
def test_setup_part_of_stdout(ch, tmp_path, monkeypatch, case):
    if case.startswith("water/"):
        d = os.path.join(WATER, case[6:])
    else:
        d = os.path.join(CASES_DIR, case)
    for fn in os.listdir(d):
        if fn == "fm_setup.in" or fn.endswith(".dat"):
            shutil.copy(os.path.join(d, fn), tmp_path / fn)
    monkeypatch.chdir(tmp_path)  # the run reads fm_setup.in from its working directory
    ours = _report_setup(ch, "fm_setup.in", tmp_path)
    ref = _golden_prefix(os.path.join(d, "fm_setup.out.gz"))
    bad = differences(ref, ours)
    assert not bad, "\n".join("ref %s: %r\nour %s: %r" % (a[0], a[1], b[0], b[1]) for a, b in bad)
