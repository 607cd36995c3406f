"""Shared inputs of the out-of-range (CHEBYFIX) tests: generated 2+3+4-body
C/H models with pairs below S_MINIM inside 3- and 4-body clusters, and the
reference's chebyfix case (published HN3 force field, its first frame with a
few atoms moved below S_MINIM)."""
import gzip
import json
import os

import numpy as np

from conftest import GOLDEN

MODES = [("SMOOTH", 0.01), ("SMOOTH", 0.025), ("CONSTANT_DERIV", None), ("ZERO_DERIV", None)]
MODE_IDS = ["smooth-0.01", "smooth-0.025", "constant-deriv", "zero-deriv"]
FIX_CODE = {"ZERO_DERIV": 0, "CONSTANT_DERIV": 1, "SMOOTH": 2}
SMIN = {("C", "C"): 1.0, ("C", "H"): 0.9, ("H", "C"): 0.9, ("H", "H"): 0.8}


def fm_setup_text(synth, mode, dist, orders=(10, 6, 4)):
    text = synth.fm_setup(["C", "H"], [12.011, 1.008], orders[0], orders[1], orders[2], SMIN, 4.5, 1.3,
                          smax3=4.0, smax4=3.5)
    line = "\t%s %g" % (mode, dist) if dist is not None else "\t" + mode
    return text.replace("# ENDFILE #", "# CHEBYFIX #\n" + line + "\n# ENDFILE #")


def close_frames(ch, synth, n=2, natoms=48, seed=81):
    """Frames whose closest pairs sit below their S_MINIM (0.65-0.95 of it)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        cell = np.diag([9.5, 9.0, 10.0])
        pos = synth.random_sequential(rng, natoms, cell, 1.05)
        typ = rng.integers(0, 2, natoms).astype(np.int32)
        # pull a few atoms towards a neighbour, below S_MINIM of the pair
        for a in rng.choice(natoms, 6, replace=False):
            d = pos - pos[a]
            r = np.linalg.norm(d, axis=1)
            r[a] = 1e9
            b = int(np.argmin(r))
            s = SMIN[("CH"[typ[a]], "CH"[typ[b]])]
            target = s * rng.uniform(0.65, 0.95)
            pos[a] = pos[b] - d[b] / r[b] * target
        out.append(ch.PyFrame(natoms, pos, typ, cell, rng.normal(size=3 * natoms), np.zeros(6), -10.0 * natoms, i))
    return out


def pairs_below_smin(frame, names, smin, cutoff):
    """Pairs (minimum image) closer than both their S_MINIM and `cutoff`."""
    x = np.asarray(frame.xyz).reshape(-1, 3)
    cell = np.asarray(frame.cell).reshape(3, 3)
    inv = np.linalg.inv(cell)
    cnt = 0
    for i in range(len(x)):
        d = x[i + 1:] - x[i]
        f = d @ inv
        f -= np.round(f)
        r = np.linalg.norm(f @ cell, axis=1)
        for j, rr in zip(range(i + 1, len(x)), r):
            s = smin[(names[frame.types[i]], names[frame.types[j]])]
            if rr < s and rr < cutoff:
                cnt += 1
    return cnt


CHEBYFIX_DIR = os.path.join(GOLDEN, "md", "chebyfix")


def hn3_case(ch, tmp_path):
    """(Setup of the published HN3 params, its run's CHEBYFIX setting, the first
    frame with one N-N, one N-H and one H-H pair moved below S_MINIM)."""
    p = os.path.join(str(tmp_path), "params.txt")
    with gzip.open(os.path.join(CHEBYFIX_DIR, "params.txt.gz"), "rb") as g:
        open(p, "wb").write(g.read())
    s = ch.Setup.from_params(p)
    meta = json.load(open(os.path.join(CHEBYFIX_DIR, "expected.json")))
    names = s.type_names()
    lines = open(os.path.join(CHEBYFIX_DIR, "frame.xyzf")).read().split("\n")
    n = int(lines[0].split()[0])
    box = [float(v) for v in lines[1].split()[:3]]
    xyz = np.array([[float(v) for v in lines[2 + i].split()[1:4]] for i in range(n)])
    typ = np.array([names.index(lines[2 + i].split()[0]) for i in range(n)], dtype=np.int32)
    cell = np.diag(box)
    smin = {}
    for p_ in range(s.model.npairs):
        a, b = names[s.model.pair_t1[p_]], names[s.model.pair_t2[p_]]
        smin[(a, b)] = smin[(b, a)] = s.model.pair_smin[p_]
    moved = set()
    for (ta, tb), frac in ((("N", "N"), 0.9), (("N", "H"), 0.88), (("H", "H"), 0.85)):
        ia, ib = names.index(ta), names.index(tb)
        best = None
        for a in np.where(typ == ia)[0]:
            if a in moved:
                continue
            d = xyz - xyz[a]
            d -= np.round(d / box) * box
            r = np.linalg.norm(d, axis=1)
            r[a] = 1e9
            r[typ != ib] = 1e9
            r[list(moved)] = 1e9
            b = int(np.argmin(r))
            if best is None or r[b] < best[0]:
                best = (r[b], a, b, d[b])
        r, a, b, d = best
        xyz[b] = xyz[a] + d / r * smin[(ta, tb)] * frac
        moved.update((int(a), int(b)))
    f = ch.PyFrame(n, xyz, typ, cell)
    return s, meta, f, names, smin
