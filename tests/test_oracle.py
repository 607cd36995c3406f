"""Pins of the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each oracle function is pinned here by at least one independent check:
  assembly (Deriv_2B/3B/4B + row layout)  -> the reference's golden A rows:
      nonorth2 (full 8640x12, triclinic 2-body), test_4atoms.2 (last frame,
      2+3+4-body, C/H, 3B/4B exclusions), fitstrs (176 rows incl. stress rows)
  set_polys                                -> T_n(cos t) = cos(n t) and
      dT_n/dr by central finite differences (Morse, inverse-r, identity maps)
  set_polys_out_of_range (SMOOTH/CONSTANT/ZERO) -> continuity at s_min and the
      closed forms of the three extrapolations
  get_fcut (CUBIC, TERSOFF)                -> closed forms + finite differences
  cell volume                              -> |det H| for triclinic cells
  neighbour / cluster enumeration          -> brute-force O(N^k) count on small
      periodic cells with explicit images
  force evaluation                         -> F = A x on the force rows and
      F = -dE/dx by finite differences (incl. the r < s_min penalty)
"""
import ctypes
import itertools
import math

import numpy as np
import pytest

from conftest import colscaled_error, golden_dir


def _setup_frames(ch, name):
    s = ch.Setup(golden_dir(name) + "/fm_setup.in")
    return s, s.frames()


@pytest.mark.parametrize("name", ["nonorth2", "test_4atoms.2", "fitstrs"])
def test_oracle_matches_reference_A(ch, oracle, name):
    s, fr = _setup_frames(ch, name)
    A = oracle.assemble(s, fr)
    g = np.load(golden_dir(name) + "/golden.npz")
    G = g["A"]
    if name == "nonorth2":
        assert A.shape == G.shape
        sub = A
    else:
        sub = A[-G.shape[0]:, :G.shape[1]]
    assert colscaled_error(sub, G) < 1e-10


def test_nonorth2_b_exact(ch, tmp_path):
    """b.txt reproduces the reference bytes (forces converted, frame order)."""
    s, fr = _setup_frames(ch, "nonorth2")
    L = ch.lib()
    ch.check(L.cg_write_open(s.h, str(tmp_path).encode()))
    for f in fr:
        rows = np.zeros((s.frame_rows(f.natoms, f.index), s.model.ncols))
        cf = f.c()
        ch.check(L.cg_write_frame(s.h, ctypes.byref(cf), ch.dptr(rows), rows.shape[0]))
    ch.check(L.cg_write_close(s.h))
    g = np.load(golden_dir("nonorth2") + "/golden.npz")
    b = np.loadtxt(tmp_path / "b.txt")
    assert np.array_equal(b, g["b"])
    assert open(tmp_path / "dim.txt").read() == open(golden_dir("nonorth2") + "/dim.txt").read()


def _model(ch, synth, tmp_path, transform="MORSE", fix="SMOOTH 0.01", fcut="CUBIC"):
    text = synth.fm_setup(["C"], [12.0], 12, 0, 0, 1.0, 5.0, 1.4, fcut=fcut)
    text = text.replace("# CHBTYPE #\n\tMORSE", "# CHBTYPE #\n\t" + transform)
    text = text.replace("# ENDFILE #", "# CHEBYFIX #\n\t%s\n# ENDFILE #" % fix)
    p = tmp_path / ("m_%s_%s_%s.in" % (transform, fix.split()[0], fcut.split()[0]))
    open(p, "w").write(text)
    return ch.Setup(str(p))


def _polys(oracle, s, r, snum=12):
    Tn = np.zeros(snum + 1)
    Tnd = np.zeros(snum + 1)
    m = s.model
    dp = ctypes.POINTER(ctypes.c_double)
    oracle.L.orc_polys(s.model_ptr, 0, r, m.pair_xdiff[0], m.pair_xavg[0], snum, m.pair_smin[0],
                       Tn.ctypes.data_as(dp), Tnd.ctypes.data_as(dp))
    return Tn, Tnd


@pytest.mark.parametrize("transform", ["MORSE", "INVRSE_R", "DEFAULT"])
def test_polys_closed_form_and_derivative(ch, synth, oracle, tmp_path, transform):
    s = _model(ch, synth, tmp_path, transform=transform)
    m = s.model
    lam, smin, smax = 1.4, 1.0, 5.0
    for r in [1.3, 2.2, 3.7, 4.9]:
        Tn, Tnd = _polys(oracle, s, r)
        if transform == "MORSE":
            x = (math.exp(-r / lam) - m.pair_xavg[0]) / m.pair_xdiff[0]
            xmin, xmax = math.exp(-smax / lam), math.exp(-smin / lam)
        elif transform == "INVRSE_R":
            x = (1 / r - m.pair_xavg[0]) / m.pair_xdiff[0]
            xmin, xmax = 1 / smax, 1 / smin
        else:
            x = (r - m.pair_xavg[0]) / m.pair_xdiff[0]
            xmin, xmax = smin, smax
        assert m.pair_xavg[0] == pytest.approx((xmin + xmax) / 2)
        assert m.pair_xdiff[0] == pytest.approx((xmax - xmin) / 2)
        th = math.acos(max(-1.0, min(1.0, x)))
        for n in range(13):
            assert Tn[n] == pytest.approx(math.cos(n * th), abs=1e-11)
        h = 1e-6
        Tp, _ = _polys(oracle, s, r + h)
        Tm, _ = _polys(oracle, s, r - h)
        fd = (Tp - Tm) / (2 * h)
        assert np.allclose(Tnd, fd, rtol=1e-6, atol=1e-6 * np.abs(fd).max())


@pytest.mark.parametrize("fix", ["SMOOTH 0.01", "CONSTANT_DERIV", "ZERO_DERIV"])
def test_out_of_range_extrapolation(ch, synth, oracle, tmp_path, fix):
    s = _model(ch, synth, tmp_path, fix=fix)
    smin = 1.0
    T0, D0 = _polys(oracle, s, smin)
    for dr in [1e-3, 0.02, 0.2]:
        r = smin - dr
        T, D = _polys(oracle, s, r)
        if fix.startswith("SMOOTH"):
            f = math.exp(-dr / 0.01)
            assert np.allclose(T, T0 + 0.01 * (f - 1) * D0, rtol=1e-12, atol=1e-12)
            assert np.allclose(D, D0 * f, rtol=1e-12, atol=1e-300)
        elif fix == "CONSTANT_DERIV":
            assert np.allclose(T, T0 - dr * D0, rtol=1e-12, atol=1e-12)
            assert np.allclose(D, D0)
        else:
            assert np.allclose(T, T0) and np.all(D == 0)
    # continuity at s_min from below
    T, D = _polys(oracle, s, smin - 1e-9)
    assert np.allclose(T, T0, atol=1e-6)


@pytest.mark.parametrize("fcut", ["CUBIC", "TERSOFF 0.5"])
def test_fcut_closed_form(ch, synth, oracle, tmp_path, fcut):
    s = _model(ch, synth, tmp_path, fcut=fcut)
    rmax = 4.0
    dp = ctypes.POINTER(ctypes.c_double)
    for r in [0.5, 1.5, 2.5, 3.0, 3.5, 3.99]:
        fc, fcd = ctypes.c_double(), ctypes.c_double()
        oracle.L.orc_fcut(s.model_ptr, r, rmax, ctypes.byref(fc), ctypes.byref(fcd))
        if fcut == "CUBIC":
            assert fc.value == pytest.approx((1 - r / rmax) ** 3, rel=1e-14)
        else:
            th = rmax - 0.5 * rmax
            ref = 1.0 if r < th else 0.5 + 0.5 * math.sin((r - th) / (rmax - th) * 3.14159265359 + 3.14159265359 / 2)
            assert fc.value == pytest.approx(ref, rel=1e-13)
        h = 1e-6
        a, b, d = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        oracle.L.orc_fcut(s.model_ptr, r + h, rmax, ctypes.byref(a), ctypes.byref(d))
        oracle.L.orc_fcut(s.model_ptr, r - h, rmax, ctypes.byref(b), ctypes.byref(d))
        assert fcd.value == pytest.approx((a.value - b.value) / (2 * h), abs=1e-7)


def test_volume_is_det(oracle):
    rng = np.random.default_rng(3)
    for _ in range(5):
        a = rng.uniform(8, 14)
        cell = np.array([[a, 0, 0], [rng.uniform(-3, 3), rng.uniform(8, 14), 0],
                         [rng.uniform(-3, 3), rng.uniform(-3, 3), rng.uniform(8, 14)]])
        c = np.ascontiguousarray(cell.reshape(9))
        v = oracle.L.orc_volume(c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        assert v == pytest.approx(abs(np.linalg.det(cell)), rel=1e-12)


def _brute_counts(pos, types, cell, L, rc2, rc3, rc4, pair_ok, id_ok):
    """Clusters by explicit images: pairs a1 <= parent(a2), triplets/quads with
    every pair inside its cutoff, following the reference's ordering rules."""
    n = len(pos)
    inv = np.linalg.inv(cell)
    frac = pos @ inv
    frac -= np.floor(frac)
    wrapped = frac @ cell
    imgs = []
    for s in itertools.product(range(-L, L + 1), repeat=3):
        for j in range(n):
            imgs.append((j, wrapped[j] + np.array(s) @ cell))
    counts = [0, 0, 0]
    for i in range(n):
        nb = [(j, p) for (j, p) in imgs if j >= i and np.linalg.norm(p - wrapped[i]) > 1e-12]
        d = [np.linalg.norm(p - wrapped[i]) for (_, p) in nb]
        counts[0] += sum(1 for (k, (j, p)) in enumerate(nb) if pair_ok(types[i], types[j], d[k]))
        n3 = [k for k in range(len(nb)) if d[k] < rc3]
        for a, b in itertools.permutations(n3, 2):
            if nb[a][0] > nb[b][0]:
                continue
            if np.linalg.norm(nb[a][1] - nb[b][1]) < rc3 and id_ok((types[i], types[nb[a][0]], types[nb[b][0]])):
                counts[1] += 1
        n4 = [k for k in range(len(nb)) if d[k] < rc4]
        for a, b, c in itertools.permutations(n4, 3):
            if nb[a][0] > nb[b][0] or nb[b][0] > nb[c][0]:
                continue
            ps = [nb[a][1], nb[b][1], nb[c][1]]
            if all(np.linalg.norm(ps[x] - ps[y]) < rc4 for x in range(3) for y in range(x + 1, 3)):
                counts[2] += 1
    return counts


def test_enumeration_brute_force(ch, synth, oracle, tmp_path):
    text = synth.fm_setup(["C", "H"], [12.0, 1.0], 6, 3, 2, 0.9, 3.2, 1.4, smax3=3.0, smax4=2.6)
    open(tmp_path / "b.in", "w").write(text)
    s = ch.Setup(str(tmp_path / "b.in"))
    rng = np.random.default_rng(7)
    cell = np.array([[7.5, 0, 0], [1.2, 7.0, 0], [-0.8, 0.9, 7.2]])
    pos = synth.random_sequential(rng, 24, cell, 1.0)
    types = rng.integers(0, 2, 24)
    f = ch.PyFrame(24, pos, types, cell)
    cnt = (ctypes.c_longlong * 3)()
    small = ctypes.c_int(0)
    cf = f.c()
    assert oracle.L.orc_count_clusters(s.model_ptr, ctypes.byref(cf), ctypes.byref(small), cnt) == 0
    ref = _brute_counts(pos, types, cell, 1, 3.2, 3.0, 2.6, lambda a, b, r: 0.9 < r < 3.2, lambda t: True)
    assert list(cnt) == ref


def test_force_rows_are_A_x_and_gradient(ch, synth, oracle, tmp_path):
    """Force evaluation equals A x (no pair below s_min) and -dE/dx numerically."""
    gen = synth.chno(nframes=1, nmin=40, nmax=40, seed=11, o2=8, o3=4, o4=3)
    text, frames, _ = gen
    open(tmp_path / "f.in", "w").write(text)
    s = ch.Setup(str(tmp_path / "f.in"))
    fr = frames[0]
    f = ch.PyFrame(fr["natoms"], fr["xyz"], fr["types"], fr["cell"])
    rng = np.random.default_rng(5)
    x = rng.normal(size=s.model.ncols)
    F, E, V = oracle.force(s, f, x)
    A = oracle.assemble(s, [f])
    # rows with d < s_min+d_p would differ (penalty); the generator keeps d >= 0.9 > 0.8 + 0.01
    assert np.allclose(F, A[: 3 * f.natoms] @ x, rtol=1e-10, atol=1e-10 * np.abs(F).max())
    h = 1e-5
    for k in rng.choice(3 * f.natoms, 6, replace=False):
        xp = f.xyz.copy(); xp[k] += h
        xm = f.xyz.copy(); xm[k] -= h
        Ep = oracle.force(s, ch.PyFrame(f.natoms, xp, f.types, f.cell), x)[1]
        Em = oracle.force(s, ch.PyFrame(f.natoms, xm, f.types, f.cell), x)[1]
        assert F[k] == pytest.approx(-(Ep - Em) / (2 * h), rel=1e-5, abs=1e-6 * np.abs(F).max())


@pytest.mark.parametrize("fcut", ["CUBIC", "TERSOFF 0.5"])
@pytest.mark.parametrize("mode,dist", [("SMOOTH", 0.01), ("SMOOTH", 0.025), ("CONSTANT_DERIV", None),
                                       ("ZERO_DERIV", None)])
def test_oracle_force_is_minus_gradient_below_smin(ch, synth, oracle, tmp_path, mode, dist, fcut):
    """The oracle's forces in every out-of-range mode are -dE/dx of its own
    energy (central differences) on 2+3+4-body C/H configurations whose 3- and
    4-body clusters hold pairs below S_MINIM: each mode's extrapolation and its
    derivative belong together, so a dropped term or a wrong sign in either
    shows here."""
    from chebyfix_cases import SMIN, close_frames, fm_setup_text, pairs_below_smin
    text = fm_setup_text(synth, mode, dist, orders=(8, 5, 4)).replace("# FCUTTYP #\n\tCUBIC", "# FCUTTYP #\n\t" + fcut)
    p = tmp_path / "fm_setup.in"
    open(p, "w").write(text)
    s = ch.Setup(str(p))
    f = close_frames(ch, synth, n=1, natoms=40, seed=83)[0]
    assert pairs_below_smin(f, ["C", "H"], SMIN, 3.5) >= 3
    x = np.random.default_rng(31).normal(size=s.model.ncols)
    F, E, V = oracle.force(s, f, x, penalty_scale=0.0)
    xyz = np.array(f.xyz).reshape(-1, 3)
    atoms = np.argsort(-np.abs(F.reshape(-1, 3)).max(axis=1))[:4]
    h = 1e-6
    for a in atoms:
        for c in range(3):
            e = []
            for sgn in (1, -1):
                y = xyz.copy()
                y[a, c] += sgn * h
                g = ch.PyFrame(f.natoms, y, f.types, f.cell)
                e.append(oracle.force(s, g, x, penalty_scale=0.0)[1])
            fd = -(e[0] - e[1]) / (2 * h)
            assert fd == pytest.approx(F[3 * a + c], rel=1e-6, abs=1e-7 * np.abs(F).max())
