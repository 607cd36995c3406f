"""GPU design-matrix assembly vs the oracle and the reference's golden rows.

Bar: FP64 A entries within 1e-10 column-scaled relative error
(|A_gpu - A_ref| / max(|A_ref|, 1e-3 max_i |A_ref[:, j]|)).  FP32 (tables,
products and accumulation in FP32): within 5e-3 column-scaled and 3e-5 of the
column's largest magnitude -- FP32 sums over many clusters of both signs lose
relative accuracy on entries that are small compared with their column.
"""
import numpy as np
import pytest

from conftest import colscaled_error, golden_dir, synth_setup

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL32 = 5e-3
TOL32_COLMAX = 3e-5


def fp32_ok(A32, Ao):
    colmax = np.maximum(np.abs(Ao).max(axis=0), 1e-300)
    rel = float((np.abs(A32 - Ao) / colmax).max())
    cs = colscaled_error(A32, Ao)
    print('fp32 colscaled %.2e colmax-rel %.2e' % (cs, rel))
    return cs < TOL32 and rel < TOL32_COLMAX


@pytest.fixture(scope="module")
def dev_cache():
    return {}


def _dev(ch, s):
    return ch.Device(s)


@pytest.mark.parametrize("name", ["nonorth2", "test_4atoms.2", "fitstrs"])
def test_golden_cases(ch, oracle, name):
    print("case", name, flush=True)
    s = ch.Setup(golden_dir(name) + "/fm_setup.in")
    fr = s.frames()
    dev = _dev(ch, s)
    A, b = dev.assemble(fr)
    Ao = oracle.assemble(s, fr)
    assert colscaled_error(A, Ao) < TOL64
    G = np.load(golden_dir(name) + "/golden.npz")["A"]
    sub = A if name == "nonorth2" else A[-G.shape[0]:, :G.shape[1]]
    assert colscaled_error(sub, G) < TOL64
    A32, _ = dev.assemble(fr, prec=ch.CG_FP32)
    assert fp32_ok(A32, Ao)
    dev.close()


@pytest.mark.parametrize("case", ["carbon", "water", "chno", "small_cell", "high_orders"])
def test_synthetic_cases(ch, synth, oracle, tmp_path, case):
    print("case", case, flush=True)
    if case == "carbon":
        gen = synth.carbon(nframes=1, natoms=128, seed=21)
    elif case == "water":
        gen = synth.water(nframes=2, seed=22)
    elif case == "chno":
        # 2.8 allowed 4-body sets per column: the product-form set loops
        gen = synth.chno(nframes=3, nmin=40, nmax=220, seed=23)
    elif case == "high_orders":
        # 3-body order 10 and 4-body order 6: the generic table kernels (the
        # hand-scheduled set loops cover 3-body orders up to 9, 4-body up to 5)
        gen = synth.carbon(nframes=2, natoms=48, seed=28, o2=12, o3=10, o4=6)
    else:
        # 8 carbon atoms in a 5 A cell: cell extent below the cutoff -> unordered
        # neighbour lists and PERM_SCALE 1/2, 1/6, 1/24 (NEIGHBORS::DO_UPDATE_SMALL)
        text, frames, names = synth.carbon(nframes=1, natoms=8, density=1.6, seed=24, o2=8, o3=4, o4=3)
        gen = (text, frames, names)
    s, fr = synth_setup(ch, synth, tmp_path, gen)
    c4 = s.model.c4
    if case == "chno":
        assert c4.total_allowed < 6 * c4.ncols
    elif case == "carbon":  # orders 16/9/5: 17.5 sets per column, quotient form
        assert c4.total_allowed >= 6 * c4.ncols
    dev = _dev(ch, s)
    A, b = dev.assemble(fr)
    Ao = oracle.assemble(s, fr)
    assert A.shape == Ao.shape
    assert colscaled_error(A, Ao) < TOL64
    A32, _ = dev.assemble(fr, prec=ch.CG_FP32)
    assert fp32_ok(A32, Ao)
    dev.close()


def test_cell_list_frame_matches_oracle(ch, synth, oracle, tmp_path):
    """A frame large enough for the cell-list neighbour search (1400 carbon
    atoms, 23.4 A cube: 5 bins of >= 4.5 A per axis) against the oracle's
    all-pairs search; a 2+3-body model keeps the oracle fast."""
    text, frames, names = synth.carbon(nframes=1, natoms=1400, seed=26, o2=8, o3=4, o4=0)
    s, fr = synth_setup(ch, synth, tmp_path, (text, frames, names))
    assert np.floor(np.asarray(fr[0].cell)[0] / 4.5) >= 5
    dev = _dev(ch, s)
    A, b = dev.assemble(fr)
    Ao = oracle.assemble(s, fr)
    assert colscaled_error(A, Ao) < TOL64
    ms, ncl = dev.detail()
    import ctypes
    cnt = (ctypes.c_longlong * 3)()
    cf = fr[0].c()
    small = ctypes.c_int(0)
    assert oracle.L.orc_count_clusters(s.model_ptr, ctypes.byref(cf), ctypes.byref(small), cnt) == 0
    assert list(ncl[:2]) == list(cnt[:2])
    dev.close()


def test_newton_third_law_full_size(ch, synth, tmp_path):
    """At the benchmark size (512-atom carbon frames, orders 16/9/5) the force
    rows of every column sum to zero over the atoms of a frame."""
    s, fr = synth_setup(ch, synth, tmp_path, synth.carbon(nframes=4, seed=25))
    dev = _dev(ch, s)
    A, b = dev.assemble(fr)
    r = 0
    for f in fr:
        n = f.natoms
        F = A[r:r + 3 * n]
        for c in range(3):
            comp = F[c::3]
            scale = np.abs(comp).sum(axis=0) + 1e-300
            assert np.all(np.abs(comp.sum(axis=0)) <= 1e-11 * scale)
        r += s.frame_rows(n, f.index)
    assert r == A.shape[0]
    # translation invariance: shifting every atom leaves A unchanged
    fr2 = [ch.PyFrame(f.natoms, f.xyz + 0.37, f.types, f.cell, f.forces, f.stress, f.energy, f.index) for f in fr]
    A2, _ = dev.assemble(fr2)
    assert colscaled_error(A2, A) < 1e-10
    dev.close()


def test_empty_and_ragged_frames(ch, synth, tmp_path):
    """Frames without any cluster (isolated atoms) and frames of very different
    sizes in one batch."""
    text, frames, names = synth.chno(nframes=2, nmin=30, nmax=300, seed=26)
    s, fr = synth_setup(ch, synth, tmp_path, (text, frames, names))
    # frame of 3 atoms far apart: no pair inside any cutoff
    iso = ch.PyFrame(3, np.array([[1.0, 1, 1], [8, 8, 8], [15, 15, 15]]), np.array([0, 1, 2]),
                     np.diag([24.0, 24.0, 24.0]), index=2)
    batch = [fr[0], iso, fr[1]]
    dev = _dev(ch, s)
    A, b = dev.assemble(batch)
    from conftest import Oracle
    Ao = Oracle(ch).assemble(s, batch)
    assert colscaled_error(A, Ao) < TOL64
    r0 = s.frame_rows(fr[0].natoms, 0)
    n_iso = s.frame_rows(3, 2)
    assert np.all(A[r0:r0 + 9, : s.model.ncols_short] == 0)
    dev.close()


def test_autotune_configurations_give_the_same_A(ch, synth, oracle, tmp_path):
    """Every tile configuration the autotuner tries, per frame-size class, assembles
    the same A as the default one (and as the oracle); the chosen ones are then
    used, and a context created later for the same model starts from them."""
    gen = synth.chno(nframes=6, nmin=40, nmax=700, seed=45)
    s, fr = synth_setup(ch, synth, tmp_path, gen)
    classes = {0 if f.natoms < 128 else 1 if f.natoms < 512 else 2 if f.natoms < 1024 else 3 for f in fr}
    dev = ch.Device(s)
    A0, b0 = dev.assemble(fr)
    choice, ms = dev.autotune(fr)
    nk = len(ch.Device.TILE_CONFIGS)
    for cls in range(4):
        if cls in classes:
            assert choice[0][cls] == 0
            assert all(0 <= choice[w][cls] < nk for w in (1, 2))
            assert np.all(ms[:, 1:, cls] > 0)
        else:
            assert all(choice[w][cls] == -1 for w in range(3))
            assert np.all(ms[:, :, cls] == 0)
    # the autotuner leaves the assembly of the kept choices
    rows = A0.shape[0]
    A2 = np.zeros_like(A0)
    ch.check(dev.L.cg_copy_A(dev.h, 0, rows, ch.dptr(A2)), dev.h)
    A1, b1 = dev.assemble(fr)
    assert colscaled_error(A2, A0) < 1e-12
    assert colscaled_error(A1, A0) < 1e-12
    assert np.array_equal(b1, b0)
    Ao = oracle.assemble(s, fr)
    assert colscaled_error(A1, Ao) < 1e-10
    used = dev.tile_choice()
    dev.close()
    dev2 = ch.Device(s)
    for cls in classes:
        assert list(dev2.tile_choice()[1:, cls]) == list(used[1:, cls])
    dev2.close()


def test_mixed_tile_configurations_side_by_side(ch, synth, oracle, tmp_path):
    """Frame-size classes with different tile configurations are launched side by
    side on separate streams; A equals the default tiles' A and the oracle's, in
    FP64 and FP32."""
    gen = synth.chno(nframes=6, nmin=40, nmax=700, seed=46)
    s, fr = synth_setup(ch, synth, tmp_path, gen)
    dev = ch.Device(s)
    A0, b0 = dev.assemble(fr)
    A032, _ = dev.assemble(fr, prec=ch.CG_FP32)
    with pytest.raises(RuntimeError):
        dev.tile_set([[1, 0, 0, 0], [0] * 4, [0] * 4])
    for mix in ([[0] * 4, [2, 4, 0, 3], [4, 2, 3, 0]], [[0] * 4, [1, 3, 2, 4], [3, 1, 4, 2]]):
        dev.tile_set(mix)
        assert dev.tile_choice().tolist() == mix
        A1, b1 = dev.assemble(fr)
        assert colscaled_error(A1, A0) < 1e-12
        assert np.array_equal(b1, b0)
        A132, _ = dev.assemble(fr, prec=ch.CG_FP32)
        assert colscaled_error(A132, A032) < 1e-6
    assert colscaled_error(A1, oracle.assemble(s, fr)) < 1e-10
    dev.close()


def test_full_size_carbon_frame_matches_oracle(ch, synth, oracle, tmp_path):
    """One frame of the benchmark workload at full size (512 carbon atoms at
    2.2 g/cc, orders 16/9/5, 1542 rows x 1038 columns) against the oracle, FP64
    and FP32."""
    print("512-atom carbon frame, orders 16/9/5", flush=True)
    s, fr = synth_setup(ch, synth, tmp_path, synth.carbon(nframes=1, natoms=512, seed=102))
    dev = ch.Device(s)
    A, b = dev.assemble(fr)
    A32, _ = dev.assemble(fr, prec=ch.CG_FP32)
    dev.close()
    Ao = oracle.assemble(s, fr)
    assert A.shape == (1542, 1038)
    err = colscaled_error(A, Ao)
    print("FP64 colscaled error %.2e" % err)
    assert err <= TOL64
    assert fp32_ok(A32, Ao)


def test_large_triclinic_chno_frame_matches_oracle(ch, synth, oracle, tmp_path):
    """A 2048-atom C/H/N/O frame in a triclinic cell (orders 12/6/3, 10 504
    columns, cell lists) against the oracle."""
    print("2048-atom C/H/N/O triclinic frame, orders 12/6/3", flush=True)
    gen = synth.chno(nframes=1, nmin=2048, nmax=2048, seed=104)
    s, fr = synth_setup(ch, synth, tmp_path, gen)
    cell = np.asarray(fr[0].cell).reshape(3, 3)
    assert np.count_nonzero(cell - np.diag(np.diag(cell))) > 0  # triclinic
    dev = ch.Device(s)
    A, b = dev.assemble(fr)
    dev.close()
    Ao = oracle.assemble(s, fr)
    assert A.shape == (3 * 2048 + 3 + 3, s.model.ncols)
    err = colscaled_error(A, Ao)
    print("FP64 colscaled error %.2e" % err)
    assert err <= TOL64
