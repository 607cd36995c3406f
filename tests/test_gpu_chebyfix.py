"""Out-of-range treatment below S_MINIM (Cheby::set_polys_out_of_range,
src/Cheby.C:294-443; CHEBYFIX, src/input.C:713-747) on the GPU, against the
oracle at 1e-10: design-matrix rows and force evaluation of 2+3+4-body models
whose 3- and 4-body clusters hold pairs below S_MINIM, in every mode (SMOOTH at
the default 0.01 and at 0.025, CONSTANT_DERIV, ZERO_DERIV), and the reference's
chebyfix force field (published HN3 2+3+4-body, Tersoff cutoffs, special 3B/4B
cutoffs, energy offsets, trimmed parameter file) in each mode."""
import numpy as np
import pytest

from chebyfix_cases import FIX_CODE, MODE_IDS, MODES, SMIN, close_frames, fm_setup_text, hn3_case, pairs_below_smin
from conftest import colscaled_error

pytestmark = pytest.mark.gpu


def _setup(ch, synth, tmp_path, mode, dist):
    p = tmp_path / "fm_setup.in"
    open(p, "w").write(fm_setup_text(synth, mode, dist))
    s = ch.Setup(str(p))
    assert s.model.fix_type == FIX_CODE[mode]
    if dist is not None:
        assert s.model.smooth_dist == dist
    return s


@pytest.mark.parametrize("mode,dist", MODES, ids=MODE_IDS)
def test_assembly_below_smin(ch, synth, oracle, tmp_path, mode, dist):
    print("assembly", mode, dist, flush=True)
    s = _setup(ch, synth, tmp_path, mode, dist)
    # 6 or more allowed 4-body sets per column: the quotient-form set loops
    c4 = s.model.c4
    assert c4.total_allowed >= 6 * c4.ncols
    fr = close_frames(ch, synth)
    for f in fr:  # the 3- and 4-body families see pairs below S_MINIM
        assert pairs_below_smin(f, ["C", "H"], SMIN, 3.5) >= 3
    dev = ch.Device(s)
    A, b = dev.assemble(fr)
    dev.close()
    Ao = oracle.assemble(s, fr)
    assert colscaled_error(A, Ao) <= 1e-10
    # the 3- and 4-body columns are not all zero on the rows of the close atoms
    n2 = s.model.n2b_cols
    assert np.abs(Ao[:, n2:s.model.ncols_short]).max() > 0


@pytest.mark.parametrize("mode,dist", MODES, ids=MODE_IDS)
def test_force_below_smin(ch, synth, oracle, tmp_path, mode, dist):
    print("force", mode, dist, flush=True)
    s = _setup(ch, synth, tmp_path, mode, dist)
    fr = close_frames(ch, synth, n=2, seed=82)
    x = np.random.default_rng(21).normal(size=s.model.ncols)
    dev = ch.Device(s)
    for f in fr:
        F, E, V = dev.force(x, f)
        Fo, Eo, Vo = oracle.force(s, f, x)
        scale = np.abs(Fo).max()
        assert np.max(np.abs(F - Fo)) <= 1e-10 * scale
        assert E == pytest.approx(Eo, rel=1e-10, abs=1e-10 * abs(Eo))
        assert np.max(np.abs(V - Vo)) <= 1e-10 * np.abs(Vo).max()
    dev.close()


@pytest.mark.parametrize("mode,dist", [("RUN", None)] + MODES, ids=["as-run"] + MODE_IDS)
def test_reference_chebyfix_model(ch, oracle, tmp_path, mode, dist):
    """The chebyfix run's force field on its first frame (three pairs moved
    below S_MINIM) under its own CHEBYFIX (SMOOTH 0.025) and every other mode."""
    print("chebyfix model", mode, dist, flush=True)
    s, meta, f, names, smin = hn3_case(ch, tmp_path)
    if mode == "RUN":
        mode, dist = meta["chebyfix"], meta["smooth_distance"]
    s.set_chebyfix(FIX_CODE[mode], dist if dist is not None else 0.01)
    assert pairs_below_smin(f, names, smin, 4.0) >= 3
    x, pd, ps = s.coeffs()
    dev = ch.Device(s)
    dev.set_penalty(pd, ps)
    F, E, V = dev.force(x, f)
    dev.close()
    Fo, Eo, Vo = oracle.force(s, f, x, penalty_scale=ps, penalty_dist=pd)
    scale = np.abs(Fo).max()
    assert np.max(np.abs(F - Fo)) <= 1e-10 * scale
    assert E == pytest.approx(Eo, rel=1e-10)
    assert np.max(np.abs(V - Vo)) <= 1e-10 * np.abs(Vo).max()
