"""FITCOUL Ewald charge columns (ZCalc_Ewald_Deriv, src/Ewald.C:477-750) and
charge-constraint rows (src/A_Matrix.C:567-624), CPU side: the oracle's
restatement against the reference's golden A rows of h2o-4bcheby (half
neighbour lists) and h2o-4bcheby-small (unordered lists, PERM_SCALE), against
minus the gradient of the Ewald pair energy by finite differences, and the
text output (b.txt, dim.txt, constraint rows) against the reference's
36-rank skip_frames run."""
import ctypes
import gzip
import math
import os

import numpy as np
import pytest

from conftest import WATER, colscaled_error, fitcoul_text, synth_setup, water_setup


def _orc_params(oracle, natoms, cell):
    a, r, k = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
    assert oracle.L.orc_ewald_params(natoms, cell.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     ctypes.byref(a), ctypes.byref(r), ctypes.byref(k)) == 0
    return a.value, r.value, k.value


@pytest.mark.parametrize("name", ["h2o-4bcheby", "h2o-4bcheby-small"])
def test_oracle_matches_golden_rows(ch, oracle, tmp_path, name):
    """Frame 32 of the reference run: its last 152 force rows (416 Chebyshev
    columns + 3 Ewald columns) and the two constraint rows after it."""
    s = water_setup(ch, tmp_path, name)
    g = np.load(os.path.join(WATER, name, "golden.npz"))
    G = g["A"]
    assert G.shape == (154, s.model.ncols) and s.model.n_charge_cols == 3
    fr = s.frames(str(tmp_path))
    A = oracle.assemble(s, [fr[-1]])
    assert colscaled_error(A[-152:], G[:152]) < 1e-10
    # columns of pairs without the atom's type are -0.0, as printed
    P = s.model.ncols_short
    assert np.array_equal(np.signbit(A[-152:, P:]), np.signbit(G[:152, P:]))
    rows, b = s.constraint_rows()
    assert np.array_equal(rows, G[152:]) and np.array_equal(b, [0.0, 0.0])


def test_ewald_parameters(ch, oracle):
    """optimal_ewald_params: the library and the oracle agree; in the golden
    water box (96 atoms, 8.609 A) the cutoff is capped at half the cell and
    alpha follows from it; in a large cell the uncapped Fincham values hold."""
    L = 8.6091309
    a, rc, k = ch.ewald_params(96, np.diag([L, L, L]))
    assert (a, rc, k) == _orc_params(oracle, 96, np.diag([L, L, L]))
    p = -math.log(1e-6) * 0.8
    assert rc == pytest.approx(L / 2, rel=1e-15) and a == pytest.approx(math.sqrt(p) / rc, rel=1e-14)
    assert k == math.ceil(2 * a * math.sqrt(p) / 1.1 * L / (2 * math.pi))
    cell = np.diag([40.0, 42.0, 44.0])
    a, rc, k = ch.ewald_params(5000, cell)
    assert (a, rc, k) == _orc_params(oracle, 5000, cell)
    V = 40.0 * 42.0 * 44.0
    a0 = math.sqrt(math.pi) * (5.5 * 5000 / V ** 2) ** (1 / 6)
    assert a == pytest.approx(a0, rel=1e-14) and rc == pytest.approx(1.1 * math.sqrt(p) / a0, rel=1e-14)
    assert rc < 20.0


def _ewald_energy(pos, types, L, alpha, rc, kmax, pair_of):
    """Ewald pair energy per unit charge product, summed per pair type:
    erfc(a r)/r over minimum-image pairs within rc, plus
    (4 pi / V) sum_k exp(-k^2/4a^2)/k^2 cos(k.r_ij) over the k set of the
    reference (|k_x|,|k_y|,|k_z| <= kmax, 0 < |n|^2 < 50)."""
    n = len(types)
    E = np.zeros(max(pair_of.values()) + 1)
    V = L[0] * L[1] * L[2]
    ks = []
    for kx in range(-kmax, kmax + 1):
        for ky in range(-kmax, kmax + 1):
            for kz in range(-kmax, kmax + 1):
                q = kx * kx + ky * ky + kz * kz
                if 0 < q < 50:
                    ks.append((kx, ky, kz))
    K = np.array(ks, dtype=float) * (2 * math.pi / np.asarray(L))
    ksq = (K ** 2).sum(1)
    kf = np.exp(-ksq / (4 * alpha ** 2)) / ksq
    for i in range(n):
        for j in range(i):
            d = pos[j] - pos[i]
            dm = d - L * np.rint(d / L)
            r = np.sqrt((dm ** 2).sum())
            p = pair_of[(types[i], types[j])]
            if r < rc:
                E[p] += math.erfc(alpha * r) / r
            E[p] += (4 * math.pi / V) * (kf * np.cos(K @ d)).sum()
    return E


def test_oracle_columns_are_minus_energy_gradient(ch, synth, oracle, tmp_path):
    """Each Ewald column of a force row is -dE_p/dr (central differences) of the
    pair type's Ewald energy: pins the signs, the half-space k factor 2, the
    real-space prefactor and the pair-type mapping (3 types)."""
    rng = np.random.default_rng(7)
    L = np.array([7.3, 8.1, 8.9])
    n = 14
    pos = rng.uniform(0, 1, (n, 3)) * L
    types = np.array([i % 3 for i in range(n)], dtype=np.int32)
    text = synth.fm_setup(["O", "H", "N"], [16.0, 1.0, 14.0], 4, 0, 0, 0.5, 3.0, 1.4)
    text = fitcoul_text(text, ["-", "+", "+"], ["OO HH NN OH ON HN 1 1 1 1 1 1 0.0"] * 5)
    fr = dict(natoms=n, xyz=pos.reshape(-1), types=types, cell=np.diag(L).reshape(-1), forces=np.zeros(3 * n),
              stress=np.zeros(6), energy=0.0, index=0)
    s, frs = synth_setup(ch, synth, tmp_path, (text, [fr], ["O", "H", "N"]))
    m = s.model
    assert m.n_charge_cols == 6 and m.ncols == m.ncols_short + 6
    A = oracle.assemble(s, frs)
    alpha, rc, kmax = ch.ewald_params(n, np.diag(L))
    ipm = [m.int_pair_map[i] for i in range(9)]
    pair_of = {(a, b): ipm[a * 3 + b] for a in range(3) for b in range(3)}
    h = 1e-5
    P = m.ncols_short
    for i in (0, 1, 5):
        for c in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, c] += h
            pm[i, c] -= h
            g = -(_ewald_energy(pp, types, L, alpha, rc, kmax, pair_of) -
                  _ewald_energy(pm, types, L, alpha, rc, kmax, pair_of)) / (2 * h)
            got = A[3 * i + c, P:P + 6]
            assert np.allclose(got, g, rtol=1e-6, atol=1e-7 * np.abs(A[:, P:]).max()), (i, c, got, g)
    # Newton's third law per column
    for c in range(3):
        assert np.all(np.abs(A[c::3, P:].sum(axis=0)) < 1e-11 * np.abs(A[:, P:]).sum(axis=0))


def test_constraint_rows_and_b_of_the_36_rank_run(ch, tmp_path):
    """skip_frames (FITCOUL, SKPFRMS 1, 36 ranks, 33 frames): b.txt byte for byte
    (frames in rank order, then the two constraint values) and dim.txt (43
    columns, 9506 rows with the constraint rows)."""
    s = water_setup(ch, tmp_path, "skip_frames")
    frames = s.frames(str(tmp_path))
    order, rank = s.frame_order(36)
    out = tmp_path / "out"
    out.mkdir()
    s.write_text(str(out), [frames[i] for i in order], rank)
    with gzip.open(os.path.join(WATER, "skip_frames", "b.txt.gz"), "rb") as g:
        assert open(out / "b.txt", "rb").read() == g.read()
    assert open(out / "dim.txt").read() == open(os.path.join(WATER, "skip_frames", "dim.txt")).read()


def test_constraint_rows_in_A_text(ch, oracle, tmp_path):
    """The A text of a FITCOUL run: force rows carry the Ewald columns, the
    constraint rows follow the last frame as the reference prints them
    ("0.0 " per short-range column, then the weights in pair-index order)."""
    s = water_setup(ch, tmp_path, "h2o-4bcheby", nframes=2)
    frames = s.frames(str(tmp_path))
    A = oracle.assemble(s, frames)
    rows = [A[:288], A[288:]]
    out = tmp_path / "out"
    out.mkdir()
    s.write_text(str(out), frames, rows=rows)
    lines = open(out / "A.txt").read().split("\n")
    assert lines[-1] == "" and len(lines) == 576 + 2 + 1
    P = s.model.ncols_short
    assert lines[-3] == "0.0 " * P + "1.0000000000000000e+04 4.0000000000000000e+04 4.0000000000000000e+04 "
    assert lines[-2] == "0.0 " * P + "1.0000000000000000e+04 -4.0000000000000000e+04 0.0000000000000000e+00 "
    first = lines[0].split("   ")
    assert len(first) == s.model.ncols + 1 and first[-1] == ""
    got = np.array([[float(v) for v in l.split()] for l in lines[:576]])
    assert np.array_equal(got, A)
    assert open(out / "dim.txt").read() == "%d %d\n" % (s.model.ncols, 578)
    b = open(out / "b.txt").read().split("\n")
    assert b[-3:] == ["0.0000000000000000e+00", "0.0000000000000000e+00", ""]


def test_fitcoul_refusals(ch, synth, tmp_path):
    """Stress fits with Ewald columns stop as in the reference (RUN_SANITY_LSQ);
    energy rows (which the reference prints without the Ewald columns) and
    hierarchical charge fits are refused."""
    base = synth.fm_setup(["O", "H"], [16.0, 1.0], 4, 0, 0, 0.5, 3.0, 1.4)
    good = fitcoul_text(base, ["-", "+"], ["OO HH OH 1 1 1 0.0", "OO OH HH 1 2 3 0.5"])
    for bad in [good.replace("# FITSTRS #\n\tfalse", "# FITSTRS #\n\ttrue"),
                good.replace("# FITENER #\n\tfalse", "# FITENER #\n\ttrue"),
                good.replace("# ENDFILE #", "# HIERARC #\n\ttrue\nEXCLUDE 2B INTERACTION: 1\nO H\n\n# ENDFILE #")]:
        open(tmp_path / "fm_setup.in", "w").write(bad)
        with pytest.raises(RuntimeError):
            ch.Setup(str(tmp_path / "fm_setup.in"))
    open(tmp_path / "fm_setup.in", "w").write(good)
    s = ch.Setup(str(tmp_path / "fm_setup.in"))
    rows, b = s.constraint_rows()
    # npairs - 1 rows; weights land in pair-index order (OO 0, HH 1, OH 2)
    assert rows.shape == (2, s.model.ncols) and list(rows[0, -3:]) == [1.0, 1.0, 1.0]
    assert list(rows[1, -3:]) == [1.0, 3.0, 2.0] and list(b) == [0.0, 0.5]


# ------------------------------------------------------------ lsq2 printer
LSQ2 = os.path.join(WATER, "lsq2")
_NUM = r"[-+]?\d+\.\d+e[-+]\d+"


def _lsq2_text(fn):
    with gzip.open(os.path.join(LSQ2, fn + ".gz"), "rt") as g:
        return g.read()


def _body(lines):
    k = next(i for i, l in enumerate(lines) if l.startswith("USECOUL:"))
    return lines[k:]


def x_from_params(ch, s, text, tmp_path):
    """The coefficient vector a params.txt prints: the printer's output for
    x = 1..ncols names the column of every printed number."""
    import re
    n = s.model.ncols
    probe = np.arange(1, n + 1, dtype=float)
    res = ch.FitResult()
    d = tmp_path / "probe"
    d.mkdir(exist_ok=True)
    ch.check(ch.lib().cg_write_params(s.h, str(d).encode(), ch.dptr(probe), n, 1, ctypes.byref(res), None, None))
    lp = _body(open(d / "params.txt").read().split("\n"))
    lg = _body(text.split("\n"))
    assert len(lp) == len(lg)
    x = np.full(n, np.nan)
    for a, b in zip(lp, lg):
        for u, v in zip(re.findall(_NUM, a), re.findall(_NUM, b)):
            x[int(round(float(u))) - 1] = float(v)
    assert not np.isnan(x).any()
    return x


def _summary(text):
    import re
    vals = {}
    for l in text.split("\n")[:16]:
        if l.startswith("! ") and "=" in l:
            k, v = l[2:].rsplit("=", 1)
            vals[k.strip()] = v.strip()
    return vals


@pytest.mark.parametrize("fn,alg", [("params.svd.weight.txt", "svd"), ("params.svd.1.0e-04.txt", "svd"),
                                    ("params.ridge.1.0e-03.txt", "ridge"),
                                    ("params.lassolars.1.0e-04.txt", "lassolars")])
def test_params_printer_weights_and_ridge(ch, tmp_path, fn, alg):
    """params.txt of chimes_lsq.py's weighted SVD, eps 1e-4 SVD and ridge runs
    (test_suite-lsq/lsq2) byte for byte from their coefficients: the ridge
    summary lines and the weighting-file line."""
    s = water_setup(ch, tmp_path, "h2o-4bcheby")  # lsq2 runs the same fm_setup.in
    text = _lsq2_text(fn)
    x = x_from_params(ch, s, text, tmp_path)
    v = _summary(text)
    res = ch.FitResult()
    res.rms = float(v["RMS force error"])
    res.max_abs_x = float(v["max abs variable"])
    res.nvars = int(v["number of fitting vars"])
    res.bic = float(v["Bayesian Information Criterion"])
    if alg == "ridge":
        res.algorithm = ch.CG_ALG_RIDGE
        res.alpha = float(v["Ridge alpha"])
    elif alg == "lassolars":
        res.algorithm = ch.CG_ALG_LASSOLARS
        res.alpha = float(v["LASSO alpha"])
    else:
        res.eps_abs = float(v["eps (= args.eps*dmax)"])
        res.eps = float(v["SVD regularization factor"])
    if "weight" in fn:
        res.weights_file = b"weights.txt"
    out = tmp_path / "out"
    out.mkdir()
    neq = int(v["Number of equations"])
    ch.check(ch.lib().cg_write_params(s.h, str(out).encode(), ch.dptr(x), len(x), neq, ctypes.byref(res), None, None))
    ours = open(out / "params.txt").read().split("\n")[1:]
    assert ours == text.split("\n")[1:]


# ------------------------------------------------------- chimes_md's Ewald
KE = 332.0637157615209


def _md_energy(pos, q, H, alpha, rc, kmax):
    """Ewald energy (kcal/mol) of point charges q (e) at pos in the cell with
    column vectors H: erfc(a r)/r over minimum-image pairs within rc, the
    reciprocal sum over n != 0 with |n|^2 < kmax^2 (k = 2 pi H^-T n), minus the
    self term -- the textbook Ewald sum at fixed alpha, rc and n-set."""
    n = len(q)
    Hi = np.linalg.inv(H)
    E = 0.0
    for i in range(n):
        for j in range(i + 1, n):
            d = pos[j] - pos[i]
            s = Hi @ d
            d = H @ (s - np.rint(s))
            r = np.sqrt(d @ d)
            if r < rc:
                E += KE * q[i] * q[j] * math.erfc(alpha * r) / r
    V = abs(np.linalg.det(H))
    ns = [(a, b, c) for a in range(-kmax, kmax + 1) for b in range(-kmax, kmax + 1) for c in range(-kmax, kmax + 1)
          if 0 < a * a + b * b + c * c < kmax * kmax]
    K = 2 * math.pi * np.array(ns, dtype=float) @ Hi  # rows: k = 2 pi H^-T n
    k2 = (K ** 2).sum(1)
    Sk = (q[None, :] * np.exp(1j * (K @ pos.T))).sum(1)
    E += 2 * math.pi * KE / V * (np.exp(-k2 / (4 * alpha ** 2)) / k2 * np.abs(Sk) ** 2).sum()
    E -= KE * alpha / math.sqrt(math.pi) * (q ** 2).sum()
    return E


@pytest.mark.parametrize("nO,nH", [(4, 8), (5, 8)])
def test_md_ewald_energy_forces_virial(ch, synth, oracle, tmp_path, nO, nH):
    """The oracle's chimes_md Ewald terms (only the charge products of x set):
    the energy equals the textbook Ewald sum with the same parameters, forces
    are minus its gradient, and the virial is minus its strain derivative
    (diagonal and shear components, the sheared cell handled by the same sum).
    The (5, 8) case has a net charge of -0.8 e: every atom's charge is shifted
    by the mean, as check_charges does."""
    rng = np.random.default_rng(3 + nO)
    L = np.array([7.1, 7.6, 8.3])
    n = nO + nH
    pos = []
    while len(pos) < n:
        p = rng.uniform(0, 1, 3) * L
        if all(np.linalg.norm((p - o) - L * np.rint((p - o) / L)) > 1.0 for o in pos):
            pos.append(p)
    pos = np.array(pos)
    types = np.array([0] * nO + [1] * nH, dtype=np.int32)
    text = synth.fm_setup(["O", "H"], [16.0, 1.0], 4, 0, 0, 0.5, 3.0, 1.4)
    text = fitcoul_text(text, ["-", "+"], ["OO HH OH 1 1 1 0.0", "OO HH OH 1 -1 0 0.0"])
    fr = dict(natoms=n, xyz=pos.reshape(-1), types=types, cell=np.diag(L).reshape(-1), forces=np.zeros(3 * n),
              stress=np.zeros(6), energy=0.0, index=0)
    s, frs = synth_setup(ch, synth, tmp_path, (text, [fr], ["O", "H"]))
    m = s.model
    qO, qH = -0.8, 0.4
    x = np.zeros(m.ncols)
    P = m.ncols_short
    for p in range(m.npairs):
        qa = qO if m.pair_t1[p] == 0 else qH
        qb = qO if m.pair_t2[p] == 0 else qH
        x[P + p] = KE * qa * qb
    F, E, V = oracle.force(s, frs[0], x)
    q = np.where(types == 0, qO, qH).astype(float)
    q -= q.sum() / n
    alpha, rc, kmax = ch.ewald_params(n, np.diag(L))
    H = np.diag(L)
    E0 = _md_energy(pos, q, H, alpha, rc, kmax)
    assert E == pytest.approx(E0, rel=1e-12)
    h = 1e-5
    for k in range(0, 3 * n, 5):
        pp, pm = pos.copy(), pos.copy()
        pp.flat[k] += h
        pm.flat[k] -= h
        g = -(_md_energy(pp, q, H, alpha, rc, kmax) - _md_energy(pm, q, H, alpha, rc, kmax)) / (2 * h)
        assert F[k] == pytest.approx(g, rel=1e-6, abs=1e-7 * np.abs(F).max())
    W = V.reshape(3, 3)
    for a, b in [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]:
        eps = np.zeros((3, 3))
        eps[a, b] += 0.5
        eps[b, a] += 0.5
        dE = []
        for sg in (1, -1):
            D = np.eye(3) + sg * h * eps
            dE.append(_md_energy(pos @ D.T, q, D @ H, alpha, rc, kmax))
        w = -(dE[0] - dE[1]) / (2 * h)
        assert W[a, b] == pytest.approx(w, rel=1e-6, abs=1e-7 * np.abs(W).max()), (a, b)
        assert W[a, b] == W[b, a]
