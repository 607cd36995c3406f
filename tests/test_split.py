"""Split output (SPLITFI) and the reference's rank-order frame layout
(SKPFRMS), pinned by the reference run test_suite-lsq/tatb: 36 frames over
36 MPI ranks, round robin, one A/dim file per rank (src/A_Matrix.C:626-720,
src/chimes_lsq.C:263,305-319, src/functions.C:182-222)."""
import gzip
import json
import os
import shutil

import pytest

from conftest import golden_dir

TATB = os.path.join(golden_dir("tatb"))


def _tatb(ch, tmp_path):
    for fn in ["fm_setup.in"]:
        shutil.copy(os.path.join(TATB, fn), tmp_path / fn)
    with gzip.open(os.path.join(TATB, "SCAN_all.xyzf.gz"), "rb") as g:
        open(tmp_path / "SCAN_all.xyzf", "wb").write(g.read())
    return ch.Setup(str(tmp_path / "fm_setup.in"))


def test_round_robin_order_matches_reference_ranks(ch, tmp_path):
    s = _tatb(ch, tmp_path)
    meta = json.load(open(os.path.join(TATB, "split.json")))
    order, rank = s.frame_order(meta["nranks"])
    got = {}
    for f, r in zip(order, rank):
        got.setdefault(str(r), []).append(f)
    assert got == meta["frames"]
    assert rank == sorted(rank)


def test_split_files_match_reference(ch, tmp_path):
    """dim.<rank>.txt, b-labeled.txt and natoms.txt of the 36-rank run byte for
    byte, and b.txt line for line except its stress rows (the A text is left
    out: the reference's A files are not kept).  The kept b.txt has zeros and
    the cell length on the 9 stress rows of every frame while the same run's
    b-labeled.txt has the stress tensor; A_Matrix.C:472-496 writes the same
    STRESS_TENSORS_X/Y/Z values to both files, so those rows come from an
    older build.  They are checked against b-labeled.txt instead."""
    s = _tatb(ch, tmp_path)
    meta = json.load(open(os.path.join(TATB, "split.json")))
    frames = s.frames(str(tmp_path))
    order, rank = s.frame_order(meta["nranks"])
    out = tmp_path / "out"
    out.mkdir()
    s.write_text(str(out), [frames[i] for i in order], rank)
    for r, text in meta["dims"].items():
        assert open(out / ("dim.%04d.txt" % int(r))).read() == text, r
    assert sorted(p for p in os.listdir(out) if p.startswith("dim.")) == ["dim.%04d.txt" % int(r) for r in
                                                                         sorted(meta["dims"], key=int)]
    for fn in ["b-labeled.txt", "natoms.txt"]:
        with gzip.open(os.path.join(TATB, fn + ".gz"), "rb") as g:
            assert open(out / fn, "rb").read() == g.read(), fn
    with gzip.open(os.path.join(TATB, "b.txt.gz"), "rt") as g:
        gold = g.read().split("\n")
    with gzip.open(os.path.join(TATB, "b-labeled.txt.gz"), "rt") as g:
        lab = g.read().split("\n")
    mine = open(out / "b.txt").read().split("\n")
    assert len(mine) == len(gold) == len(lab)
    nstress = 0
    for k, (a, b, l) in enumerate(zip(mine, gold, lab)):
        if l.startswith("s_"):
            nstress += 1
            assert float(a) == pytest.approx(float(l.split()[1]), rel=1e-5, abs=1e-6), k
        else:
            assert a == b, k
    assert nstress == 36 * 9
    assert not os.path.exists(out / "A.txt") and not os.path.exists(out / "dim.txt")


@pytest.mark.parametrize("nframes,nranks", [(10, 4), (3, 5), (7, 7), (36, 1), (1, 3)])
def test_contiguous_blocks(ch, tmp_path, nframes, nranks):
    """Without SKPFRMS every frame goes to exactly one rank, in order, in blocks
    of ceil-divided size (divide_atoms); ranks past the frame count get none."""
    text = open(os.path.join(TATB, "fm_setup.in")).read()
    text = text.replace("# SKPFRMS #\n        1", "# SKPFRMS #\n        0").replace("\t36\n", "\t%d\n" % nframes, 1)
    p = tmp_path / "fm_setup.in"
    open(p, "w").write(text)
    s = ch.Setup(str(p))
    assert s.nframes == nframes
    order, rank = s.frame_order(nranks)
    assert order == list(range(nframes))
    assert rank == sorted(rank)
    used = min(nranks, nframes)
    sizes = [rank.count(r) for r in range(nranks)]
    assert sizes[:used] == [((r + 1) * nframes + used - 1) // used - (r * nframes + used - 1) // used
                            for r in range(used)]
    assert all(n == 0 for n in sizes[used:])


def test_skip_frames_takes_every_kth_frame_of_a_rank(ch, tmp_path):
    """SKPFRMS k >= 1: rank r takes frames i with (i + r) % nranks == 0, then
    every k-th of them (round robin with skipping)."""
    text = open(os.path.join(TATB, "fm_setup.in")).read().replace("# SKPFRMS #\n        1", "# SKPFRMS #\n        3")
    p = tmp_path / "fm_setup.in"
    open(p, "w").write(text)
    s = ch.Setup(str(p))
    order, rank = s.frame_order(4)
    for r in range(4):
        mine = [f for f, q in zip(order, rank) if q == r]
        cand = [i for i in range(36) if (i + r) % 4 == 0]
        assert mine == [i for i in cand if ((i + r) // 4) % 3 == 0]
    assert len(order) == len(set(order))


def test_header_and_map_match_reference(ch, tmp_path):
    """params.header and ff_groups.map of the 4-species 2+3+4-body run.  The kept
    header predates the EXCLD1B/EXCLD2B lines that chimes_lsq.C:594-610 prints
    (like the kept headers of h2o-2bcheby2 and skip_frames); apart from them it
    is byte for byte the same."""
    s = _tatb(ch, tmp_path)
    out = tmp_path / "hdr"
    out.mkdir()
    s.write_header(str(out))
    mine = open(out / "params.header").read()
    assert "EXCLD1B: false\nEXCLD2B: false\n" in mine
    with gzip.open(os.path.join(TATB, "params.header.gz"), "rt") as g:
        assert mine.replace("EXCLD1B: false\nEXCLD2B: false\n", "") == g.read()
    with gzip.open(os.path.join(TATB, "ff_groups.map.gz"), "rb") as g:
        assert open(out / "ff_groups.map", "rb").read() == g.read()


def test_frame_logs_per_rank(ch, tmp_path):
    """frames.NNNN.log of each of the 36 ranks: the frames it processed, as
    the reference run's logs list them, each followed by its atom counts per
    type and its time (src/chimes_lsq.C:718-823)."""
    import ctypes
    import re
    import numpy as np
    s = _tatb(ch, tmp_path)
    meta = json.load(open(os.path.join(TATB, "split.json")))
    frames = s.frames(workdir=str(tmp_path))
    cf = (ch.Frame * len(frames))(*[f.c() for f in frames])
    secs = np.arange(len(frames), dtype=np.float64) * 0.25
    L = ch.lib()
    L.cg_write_frame_logs.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ch.Frame),
                                      ctypes.POINTER(ctypes.c_double)]
    ch.check(L.cg_write_frame_logs(s.h, str(tmp_path).encode(), meta["nranks"], cf, ch.dptr(secs)))
    names = s.type_names()
    for r in range(meta["nranks"]):
        lines = open(tmp_path / ("frames.%04d.log" % r)).read().split("\n")
        got = [int(l.split()[2]) for l in lines if l.startswith("Processing frame")]
        assert got == meta["frames"][str(r)], r
        for i in got:
            k = lines.index("Processing frame %d" % i)
            cnt = np.bincount(frames[i].types, minlength=len(names))
            assert lines[k + 1].split() == [w for t, n in zip(names, cnt) for w in (t, str(n))]
            assert re.match(r"^Finished frame %d in %.1f seconds$" % (i, secs[i]), lines[k + 2]), lines[k + 2]
