"""GPU fit vs the reference's fitted parameters and vs a thin-SVD oracle."""
import numpy as np
import pytest

from conftest import golden_dir, synth_setup

pytestmark = pytest.mark.gpu


def thin_svd_fit(A, b, eps=1e-5):
    """chimes_lsq.py's estimator (src/chimes_lsq.py:187-214) with the thin SVD:
    x = V diag(1/s for s > eps*s_max) U^T b."""
    U, S, VT = np.linalg.svd(A, full_matrices=False)
    keep = S > eps * S.max()
    x = VT.T[:, keep] @ ((U[:, keep].T @ b) / S[keep])
    r = A @ x - b
    return x, np.sqrt((r @ r) / len(b)), int(keep.sum())


def test_nonorth2_fit_matches_reference(ch):
    d = golden_dir("nonorth2")
    s = ch.Setup(d + "/fm_setup.in")
    fr = s.frames()
    dev = ch.Device(s)
    A, b = dev.assemble(fr, copy=False)
    x, res = dev.fit(b, eps=1e-5, refine=3)
    xr = np.array([float(l.split()[1]) for l in open(d + "/test_suite_params.txt")])
    # Coefficients: this least-squares problem has condition 2.2e4 (thin SVD of
    # the golden A); perturbing A by 3e-11 relative (the assembly's measured
    # agreement with the golden A, inside its 1e-10 bar) moves the reference
    # coefficients by up to 3.5e-7 (max |x| = 359).  The bound is 3x that
    # sensitivity; the fit's own bar is the force RMS (1e-8) checked below.
    assert np.max(np.abs(x - xr)) < 1e-6 * np.abs(xr).max()
    y = dev.matvec(x, len(b))
    yr = np.load(d + "/golden.npz")["force"]
    assert np.max(np.abs(np.round(y, 7) - yr) / np.maximum(np.abs(yr), 1e-6)) < 1e-5
    assert res.nvars == 12
    assert res.rms == pytest.approx(4.0724e-03, rel=1e-4)  # params.txt prints 5 digits
    # force RMS against the thin-SVD estimator on the golden A: 1e-8
    g = np.load(d + "/golden.npz")
    xs, rmss, _ = thin_svd_fit(g["A"], g["b"])
    assert res.rms == pytest.approx(rmss, rel=1e-8)
    dev.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_carbon_fit_vs_thin_svd(ch, synth, oracle, tmp_path, prec):
    """Force RMS of the GPU fit vs the thin-SVD estimator on the oracle's FP64 A:
    1e-8 relative for the FP64 path, 1e-6 for FP32 A + FP64 Gram + refinement.
    8 frames x 128 atoms: 3120 rows x 1038 columns (rows well above columns, as
    in a production fit; with rows ~ columns the FP32 rounding of A alone moves
    the RMS by ~1e-6)."""
    s, fr = synth_setup(ch, synth, tmp_path, synth.carbon(nframes=8, natoms=128, seed=31))
    dev = ch.Device(s)
    A, b = dev.assemble(fr, prec=prec)
    x, res = dev.fit(b, eps=1e-5, refine=4)
    Ao = oracle.assemble(s, fr)
    xo, rmso, nk = thin_svd_fit(Ao, b)
    if prec == 0:
        assert res.rms == pytest.approx(rmso, rel=1e-8)
    else:
        assert res.rms == pytest.approx(rmso, rel=1e-6)
    # the reported RMS is that of the returned x
    r = Ao @ x - b
    assert np.sqrt(r @ r / len(b)) == pytest.approx(rmso, rel=1e-6 if prec else 1e-8)
    assert abs(res.nvars - nk) <= (0 if prec == 0 else 2)
    dev.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_gram_matches_numpy(ch, synth, tmp_path, prec):
    """G = A^T A and A^T b of the device A against numpy on the copied-back A
    (FP64 MFMA; FP32 entries are widened, so the products are exact)."""
    s, fr = synth_setup(ch, synth, tmp_path, synth.water(nframes=2))
    dev = ch.Device(s)
    A, b = dev.assemble(fr, prec=prec)
    G, atb = dev.gram(b)
    Gr = A.T @ A
    scale = np.sqrt(np.outer(np.diag(Gr), np.diag(Gr))) + 1e-300
    assert np.max(np.abs(G - Gr) / scale) < 1e-12
    assert np.max(np.abs(G - G.T)) == 0.0
    ar = A.T @ b
    assert np.max(np.abs(atb - ar) / (np.sqrt(np.diag(Gr)) * np.linalg.norm(b) + 1e-300)) < 1e-12
    if prec == 1:
        Gb, _ = dev.gram(None, mode=ch.CG_GRAM_BF16X3)
        err = np.max(np.abs(Gb - Gr) / scale)
        assert err < 1e-4, err        # split-BF16x3: ~2^-16 per product, FP32 slab sums
    dev.close()


def test_bf16x3_fit_guarded(ch, synth, tmp_path):
    """The split-BF16x3 Gram resolves only the large singular values; the
    refinement must never return an iterate worse than its starting point
    (the two calls start from the same bits: A^T b and the Gram are summed
    in a fixed order)."""
    s, fr = synth_setup(ch, synth, tmp_path, synth.carbon(nframes=3, natoms=128, seed=31))
    dev = ch.Device(s)
    dev.assemble(fr, prec=ch.CG_FP32, copy=False)
    _, b = dev.assemble(fr, prec=ch.CG_FP32, copy=False)
    x0, r0 = dev.fit(b, eps=1e-5, refine=0, gram=ch.CG_GRAM_BF16X3)
    x1, r1 = dev.fit(b, eps=1e-5, refine=6, gram=ch.CG_GRAM_BF16X3)
    assert np.all(np.isfinite(x1)) and r1.rms <= r0.rms
    with pytest.raises(RuntimeError):
        dev.assemble(fr, prec=ch.CG_FP64, copy=False)
        dev.fit(b, gram=ch.CG_GRAM_BF16X3)
    dev.close()


def test_fit_is_bitwise_repeatable(ch, synth, tmp_path):
    """Two fits of the same A give the same bits (A^T b, A^T r and the Gram
    add their partial sums in a fixed order), in FP64 and on an FP32 A."""
    s, fr = synth_setup(ch, synth, tmp_path, synth.carbon(nframes=3, natoms=128, seed=33))
    dev = ch.Device(s)
    for prec in (ch.CG_FP64, ch.CG_FP32):
        _, b = dev.assemble(fr, prec=prec, copy=False)
        xa, ra = dev.fit(b, eps=1e-5, refine=3)
        xb, rb = dev.fit(b, eps=1e-5, refine=3)
        assert np.array_equal(xa, xb) and ra.rms == rb.rms, prec
    dev.close()
