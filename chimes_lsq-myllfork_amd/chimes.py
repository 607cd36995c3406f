"""ctypes binding of libchimes_gpu.so (C ABI declared in include/chimes_gpu.h).

Used by the tests and bench.py; the C++ command-line drivers link the library
directly.  The HIP entry points fail loudly (OSError / RuntimeError) when the
library was not built -- there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libchimes_gpu.so")

CG_FP64, CG_FP32 = 0, 1
CG_FIX_ZERO_DERIV, CG_FIX_CONSTANT_DERIV, CG_FIX_SMOOTH = 0, 1, 2
CG_GRAM_FP64, CG_GRAM_BF16X3 = 0, 1
CG_FIT_QR = 2
CG_ALG_SVD, CG_ALG_RIDGE, CG_ALG_LASSOLARS = 0, 1, 2
STATUS = {0: "CG_OK", 1: "CG_ERR_ARG", 2: "CG_ERR_INPUT", 3: "CG_ERR_UNSUPPORTED",
          4: "CG_ERR_DEVICE", 5: "CG_ERR_MEMORY", 6: "CG_ERR_NUMERIC", 7: "CG_ERR_EOF"}

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


class ClusterSet(C.Structure):
    _fields_ = [("nbody", C.c_int), ("npairs", C.c_int), ("order", C.c_int), ("ntypes", C.c_int),
                ("total_allowed", C.c_int), ("ncols", C.c_int), ("max_cutoff", C.c_double),
                ("excluded", _ip), ("atom_type", _ip), ("col_offset", _ip), ("n_unique", _ip),
                ("perm_offset", _ip), ("perm_powers", _ip), ("perm_col", _ip), ("perm_equiv", _ip),
                ("perm_count", _ip), ("s_min", _dp), ("s_max", _dp), ("x_avg", _dp), ("x_diff", _dp),
                ("pair_type", _ip), ("id_size", C.c_int), ("id_map", _ip), ("id_pair_map", _ip),
                ("id_atom_perm", _ip)]


class Model(C.Structure):
    _fields_ = [("ntypes", C.c_int), ("type_names", C.c_void_p), ("npairs", C.c_int),
                ("pair_t1", _ip), ("pair_t2", _ip), ("int_pair_map", _ip),
                ("pair_smin", _dp), ("pair_smax", _dp), ("pair_lambda", _dp), ("pair_xavg", _dp),
                ("pair_xdiff", _dp), ("transform", C.c_int), ("order2b", C.c_int),
                ("deriv_const", C.c_double), ("cheby_low", C.c_int), ("cheby_high", C.c_int),
                ("fcut_type", C.c_int), ("fcut_offset", C.c_double), ("fix_type", C.c_int),
                ("smooth_dist", C.c_double), ("nlayers", C.c_int), ("small_mode", C.c_int),
                ("fit_stress", C.c_int), ("fit_energy", C.c_int), ("nstress", C.c_int),
                ("nener", C.c_int), ("n2b_cols", C.c_int), ("ncols_short", C.c_int),
                ("n_charge_cols", C.c_int), ("n_offset_cols", C.c_int), ("ncols", C.c_int), ("max_cutoff_2b", C.c_double),
                ("c3", ClusterSet), ("c4", ClusterSet), ("charge_sign", C.POINTER(C.c_int))]


class Frame(C.Structure):
    _fields_ = [("natoms", C.c_int), ("xyz", _dp), ("type", _ip), ("cell", C.c_double * 9),
                ("forces", _dp), ("stress", C.c_double * 6), ("energy", C.c_double),
                ("index", C.c_int)]


class FitResult(C.Structure):
    _fields_ = [("eps", C.c_double), ("rms", C.c_double), ("max_abs_x", C.c_double), ("nvars", C.c_int),
                ("eps_abs", C.c_double), ("bic", C.c_double), ("ms_gram", C.c_double),
                ("ms_solve", C.c_double), ("ms_refine", C.c_double), ("iters", C.c_int),
                ("algorithm", C.c_int), ("alpha", C.c_double), ("weights_file", C.c_char * 256)]


_lib = None


def lib(path=None):
    """Load libchimes_gpu.so (raises OSError if it is missing)."""
    global _lib
    if _lib is None or path:
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise OSError("libchimes_gpu.so is not built (%s); run __graft_entry__.build()" % p)
        L = C.CDLL(p)
        L.cg_last_error.restype = C.c_char_p
        L.cg_last_error.argtypes = [C.c_void_p]
        L.cg_setup_read.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        L.cg_params_read.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        L.cg_params_coeffs.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.cg_setup_free.argtypes = [C.c_void_p]
        L.cg_setup_set_chebyfix.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.cg_setup_model.restype = C.POINTER(Model)
        L.cg_setup_model.argtypes = [C.c_void_p]
        L.cg_setup_nframes.argtypes = [C.c_void_p]
        L.cg_model_frame_rows.restype = C.c_int64
        L.cg_model_frame_rows.argtypes = [C.POINTER(Model), C.c_int, C.c_int]
        L.cg_traj_open.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_void_p)]
        L.cg_traj_next.argtypes = [C.c_void_p, C.POINTER(Frame)]
        L.cg_traj_close.argtypes = [C.c_void_p]
        L.cg_write_open.argtypes = [C.c_void_p, C.c_char_p]
        L.cg_write_frame.argtypes = [C.c_void_p, C.POINTER(Frame), _dp, C.c_int]
        L.cg_write_close.argtypes = [C.c_void_p]
        L.cg_write_open_opts.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        L.cg_write_rank.argtypes = [C.c_void_p, C.c_int]
        L.cg_frame_order.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]
        L.cg_write_header.argtypes = [C.c_void_p, C.c_char_p]
        def opt(name, argtypes):
            if hasattr(L, name):
                getattr(L, name).argtypes = argtypes
        if hasattr(L, "cg_create"):
            L.cg_create.argtypes = [C.POINTER(Model), C.c_int, C.POINTER(C.c_void_p)]
            L.cg_destroy.argtypes = [C.c_void_p]
            L.cg_assemble.argtypes = [C.c_void_p, C.POINTER(Frame), C.c_int, C.c_int,
                                      C.POINTER(C.c_void_p), _dp, C.POINTER(C.c_int64)]
            L.cg_copy_A.argtypes = [C.c_void_p, C.c_int64, C.c_int64, _dp]
            L.cg_last_timing.argtypes = [C.c_void_p, _dp, _dp, _dp]
            L.cg_last_detail.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_int64)]
            L.cg_synchronize.argtypes = [C.c_void_p]
        opt("cg_fit", [C.c_void_p, _dp, C.c_double, C.c_int, C.c_int, _dp, C.POINTER(FitResult)])
        opt("cg_gram", [C.c_void_p, C.c_int, _dp, _dp, _dp])
        opt("cg_matvec", [C.c_void_p, _dp, _dp])
        opt("cg_write_params", [C.c_void_p, C.c_char_p, _dp, C.c_int, C.c_int64,
                                C.POINTER(FitResult), _dp, _dp])
        opt("cg_force", [C.c_void_p, _dp, C.POINTER(Frame), _dp, _dp, _dp])
        opt("cg_force_batch", [C.c_void_p, _dp, C.POINTER(Frame), C.c_int, _dp, _dp, _dp])
        opt("cg_force_noalloc", [C.c_void_p, _dp, C.POINTER(Frame), _dp, _dp, _dp])
        opt("cg_set_penalty", [C.c_void_p, C.c_double, C.c_double])
        opt("cg_set_fit_columns", [C.c_void_p, C.POINTER(C.c_int), C.c_int])
        opt("cg_set_fit_weights", [C.c_void_p, _dp, C.c_int64])
        opt("cg_fit_ridge", [C.c_void_p, _dp, C.c_double, _dp, C.POINTER(FitResult)])
        opt("cg_fit_lassolars", [C.c_void_p, _dp, C.c_double, C.c_int, _dp, C.POINTER(FitResult)])
        opt("cg_setup_fit_columns", [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)])
        opt("cg_autotune", [C.c_void_p, C.POINTER(Frame), C.c_int, C.c_int, C.POINTER(C.c_int), _dp])
        opt("cg_tile_choice", [C.c_void_p, C.POINTER(C.c_int)])
        opt("cg_tile_set", [C.c_void_p, C.POINTER(C.c_int)])
        opt("cg_force_stats", [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)])
        opt("cg_assembly_graphs", [C.c_void_p, C.c_int])
        opt("cg_assembly_global_above", [C.c_void_p, C.c_int])
        opt("cg_assembly_deterministic", [C.c_void_p, C.c_int])
        opt("cg_assembly_pair_tables", [C.c_void_p, C.c_int, C.POINTER(C.c_int)])
        opt("cg_assembly_stats", [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int64)])
        opt("cg_copy_probe", [C.c_void_p, C.c_size_t])
        opt("cg_append_rows", [C.c_void_p, _dp, C.c_int])
        opt("cg_ewald_params", [C.c_int, _dp, _dp, _dp, C.POINTER(C.c_int)])
        opt("cg_setup_constraint_rows", [C.c_void_p, _dp, _dp, C.POINTER(C.c_int)])
        opt("cg_serve_create", [C.POINTER(Model), C.c_int, _dp, C.c_int, C.c_size_t, C.c_int,
                                C.POINTER(C.c_void_p)])
        opt("cg_serve_submit", [C.c_void_p, C.POINTER(Frame), C.POINTER(C.c_int64)])
        opt("cg_serve_poll", [C.c_void_p, C.c_int64, C.c_int, C.POINTER(C.c_int), _dp, _dp, _dp, _dp])
        opt("cg_serve_stats", [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)])
        opt("cg_serve_counters", [C.c_void_p, C.POINTER(C.c_int64), C.c_int])
        opt("cg_serve_hold", [C.c_void_p, C.c_int])
        opt("cg_serve_destroy", [C.c_void_p])
        _lib = L
    return _lib


def ewald_params(natoms, cell):
    """(alpha, rc, kmax) of a frame's Ewald columns."""
    cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
    a, r, k = C.c_double(), C.c_double(), C.c_int()
    check(lib().cg_ewald_params(int(natoms), dptr(cell), C.byref(a), C.byref(r), C.byref(k)))
    return a.value, r.value, k.value


def check(st, ctx=None):
    if st != 0:
        msg = lib().cg_last_error(ctx)
        raise RuntimeError("%s: %s" % (STATUS.get(st, st), msg.decode() if msg else ""))


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


class PyFrame:
    """A frame whose arrays are owned by numpy (safe to keep)."""

    def __init__(self, natoms, xyz, types, cell, forces=None, stress=None, energy=0.0, index=0):
        self.natoms = int(natoms)
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1)
        self.types = np.ascontiguousarray(types, dtype=np.int32)
        self.cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
        self.forces = (np.zeros(3 * self.natoms) if forces is None
                       else np.ascontiguousarray(forces, dtype=np.float64).reshape(-1))
        self.stress = np.zeros(6) if stress is None else np.asarray(stress, dtype=np.float64)
        self.energy = float(energy)
        self.index = int(index)

    def __setattr__(self, name, value):
        object.__setattr__(self, name, value)
        if name != "_cache":
            object.__setattr__(self, "_cache", None)

    def c(self):
        """The cg_frame of this frame.  Coordinates, types and forces are passed by pointer; the
        struct is built once and reused until an attribute is set or the cell or stress values
        change (marshalling 100 frames field by field took ~1.5 ms per call)."""
        sig = self.cell.tobytes() + self.stress.tobytes()
        cache = self._cache
        if cache is not None and cache[0] == sig:
            return cache[1]
        f = Frame()
        f.natoms = self.natoms
        f.xyz = dptr(self.xyz)
        f.type = iptr(self.types)
        for i in range(9):
            f.cell[i] = self.cell[i]
        f.forces = dptr(self.forces)
        for i in range(6):
            f.stress[i] = self.stress[i]
        f.energy = self.energy
        f.index = self.index
        object.__setattr__(self, "_cache", (sig, f))
        return f


class Setup:
    """Parsed fm_setup.in (host C++ parser behind the C ABI)."""

    def __init__(self, path, params=False):
        """path: fm_setup.in, or a params.txt when params=True (model + coefficients)."""
        self.L = lib()
        h = C.c_void_p()
        reader = self.L.cg_params_read if params else self.L.cg_setup_read
        check(reader(path.encode(), C.byref(h)))
        self.h = h
        self.path = path
        self.model_ptr = self.L.cg_setup_model(h)
        self.model = self.model_ptr.contents

    @classmethod
    def from_params(cls, path):
        return cls(path, params=True)

    def coeffs(self):
        """(x in A column order, penalty_dist, penalty_scale) of a params.txt setup."""
        x = np.zeros(self.model.ncols)
        d, sc = C.c_double(), C.c_double()
        check(self.L.cg_params_coeffs(self.h, dptr(x), C.byref(d), C.byref(sc)))
        return x, d.value, sc.value

    def set_chebyfix(self, fix_type, smooth_dist=0.01):
        """Out-of-range mode below S_MINIM (chimes_md's CHEBYFIX) for contexts created afterwards."""
        check(self.L.cg_setup_set_chebyfix(self.h, fix_type, smooth_dist))

    def type_names(self):
        raw = C.string_at(self.model.type_names, 8 * self.model.ntypes)
        return [raw[8 * i:8 * i + 8].split(b"\0")[0].decode() for i in range(self.model.ntypes)]

    def __del__(self):
        try:
            self.L.cg_setup_free(self.h)
        except Exception:
            pass

    @property
    def nframes(self):
        return self.L.cg_setup_nframes(self.h)

    def frame_rows(self, natoms, index):
        return self.L.cg_model_frame_rows(self.model_ptr, natoms, index)

    def frames(self, workdir=None, limit=None):
        """Read frames of the trajectory into PyFrame objects."""
        t = C.c_void_p()
        check(self.L.cg_traj_open(self.h, workdir.encode() if workdir else None, C.byref(t)))
        out = []
        try:
            n = self.nframes if limit is None else min(limit, self.nframes)
            for _ in range(n):
                f = Frame()
                check(self.L.cg_traj_next(t, C.byref(f)))
                N = f.natoms
                out.append(PyFrame(N, np.ctypeslib.as_array(f.xyz, (3 * N,)).copy(),
                                   np.ctypeslib.as_array(f.type, (N,)).copy(), list(f.cell),
                                   np.ctypeslib.as_array(f.forces, (3 * N,)).copy(),
                                   list(f.stress), f.energy, f.index))
        finally:
            self.L.cg_traj_close(t)
        return out

    def fit_columns(self):
        """Columns of A a fit keeps under the hierarchical 1B/2B exclusions."""
        n = self.model.ncols
        keep = (C.c_int * n)()
        nk = C.c_int()
        check(self.L.cg_setup_fit_columns(self.h, keep, C.byref(nk)))
        return list(keep[:nk.value])

    def constraint_rows(self):
        """FITCOUL charge-constraint rows and their b values ([n x ncols], [n])."""
        n = C.c_int()
        check(self.L.cg_setup_constraint_rows(self.h, None, None, C.byref(n)))
        rows = np.zeros((n.value, self.model.ncols))
        b = np.zeros(n.value)
        if n.value:
            check(self.L.cg_setup_constraint_rows(self.h, dptr(rows), dptr(b), C.byref(n)))
        return rows, b

    def frame_order(self, nranks):
        """(frames, rank of each) in the order the reference's nranks MPI ranks process them."""
        n = self.nframes
        order = (C.c_int * max(n, 1))()
        rank = (C.c_int * max(n, 1))()
        cnt = C.c_int()
        check(self.L.cg_frame_order(self.h, nranks, order, rank, C.byref(cnt)))
        return list(order[:cnt.value]), list(rank[:cnt.value])

    def write_text(self, outdir, frames, ranks=None, rows=None):
        """Write b/b-labeled/natoms/dim (and A, from rows [per frame], if given) through the
        library's writers, frame k belonging to ranks[k]."""
        check(self.L.cg_write_open_opts(self.h, outdir.encode(), 0 if rows is not None else 1))
        for k, f in enumerate(frames):
            if ranks is not None:
                check(self.L.cg_write_rank(self.h, int(ranks[k])))
            cf = f.c()
            nr = self.frame_rows(f.natoms, f.index)
            r = None if rows is None else np.ascontiguousarray(rows[k], dtype=np.float64)
            check(self.L.cg_write_frame(self.h, C.byref(cf), dptr(r) if r is not None else None, nr))
        check(self.L.cg_write_close(self.h))

    def write_header(self, outdir):
        check(self.L.cg_write_header(self.h, outdir.encode()))


class Device:
    """A device context (cg_create / cg_destroy)."""

    def __init__(self, setup, device=0):
        self.L = lib()
        if not hasattr(self.L, "cg_create"):
            raise OSError("libchimes_gpu.so was built without the HIP entry points")
        self.setup = setup
        h = C.c_void_p()
        check(self.L.cg_create(setup.model_ptr, device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.L.cg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def assemble(self, frames, prec=CG_FP64, copy=True):
        cf = (Frame * len(frames))(*[f.c() for f in frames])
        rows = sum(self.setup.frame_rows(f.natoms, f.index) for f in frames)
        b = np.zeros(rows)
        Ad = C.c_void_p()
        nrow = C.c_int64()
        check(self.L.cg_assemble(self.h, cf, len(frames), prec, C.byref(Ad), dptr(b), C.byref(nrow)), self.h)
        assert nrow.value == rows
        if not copy:
            return None, b
        A = np.zeros((rows, self.setup.model.ncols))
        check(self.L.cg_copy_A(self.h, 0, rows, dptr(A)), self.h)
        return A, b

    def timing(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        check(self.L.cg_last_timing(self.h, C.byref(a), C.byref(b), C.byref(c)), self.h)
        return a.value, b.value, c.value

    def synchronize(self):
        check(self.L.cg_synchronize(self.h), self.h)

    def detail(self):
        ms = (C.c_double * 3)()
        n = (C.c_int64 * 3)()
        check(self.L.cg_last_detail(self.h, ms, n), self.h)
        return list(ms), list(n)

    def append_rows(self, rows):
        """Append host rows (e.g. charge constraints) after the last assembled A."""
        r = np.ascontiguousarray(rows, dtype=np.float64)
        check(self.L.cg_append_rows(self.h, dptr(r), r.shape[0]), self.h)

    def fit(self, b, eps=1e-5, refine=2, gram=CG_GRAM_FP64):
        x = np.zeros(self.setup.model.ncols)
        res = FitResult()
        b = np.ascontiguousarray(b, dtype=np.float64)
        check(self.L.cg_fit(self.h, dptr(b), eps, refine, gram, dptr(x), C.byref(res)), self.h)
        return x, res

    def set_weights(self, w):
        """Row weights of the following CG_FIT_QR fits (None clears)."""
        if w is None:
            check(self.L.cg_set_fit_weights(self.h, None, 0), self.h)
        else:
            self._w = np.ascontiguousarray(w, dtype=np.float64)
            check(self.L.cg_set_fit_weights(self.h, dptr(self._w), self._w.size), self.h)

    def fit_ridge(self, b, alpha):
        x = np.zeros(self.setup.model.ncols)
        res = FitResult()
        b = np.ascontiguousarray(b, dtype=np.float64)
        check(self.L.cg_fit_ridge(self.h, dptr(b), alpha, dptr(x), C.byref(res)), self.h)
        return x, res

    def fit_lassolars(self, b, alpha, max_iter=100000):
        x = np.zeros(self.setup.model.ncols)
        res = FitResult()
        b = np.ascontiguousarray(b, dtype=np.float64)
        check(self.L.cg_fit_lassolars(self.h, dptr(b), alpha, max_iter, dptr(x), C.byref(res)), self.h)
        return x, res

    def gram(self, b=None, mode=CG_GRAM_FP64):
        """(G, A^T b) of the last assembled A (A^T b is None without b)."""
        n = self.setup.model.ncols
        G = np.zeros((n, n))
        atb = None
        if b is not None:
            b = np.ascontiguousarray(b, dtype=np.float64)
            atb = np.zeros(n)
        check(self.L.cg_gram(self.h, mode, dptr(b) if b is not None else None, dptr(G),
                             dptr(atb) if atb is not None else None), self.h)
        return G, atb

    def matvec(self, x, rows):
        y = np.zeros(rows)
        x = np.ascontiguousarray(x, dtype=np.float64)
        check(self.L.cg_matvec(self.h, dptr(x), dptr(y)), self.h)
        return y

    def force(self, x, frame):
        x = np.ascontiguousarray(x, dtype=np.float64)
        F = np.zeros(3 * frame.natoms)
        E = C.c_double()
        V = np.zeros(9)
        f = frame.c()
        check(self.L.cg_force(self.h, dptr(x), C.byref(f), dptr(F), C.byref(E), dptr(V)), self.h)
        return F, E.value, V

    def force_noalloc(self, x, frame):
        """cg_force_noalloc: the same results as force(), every device buffer taken
        from the library's fixed per-GPU scratch array."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        F = np.zeros(3 * frame.natoms)
        E = C.c_double()
        V = np.zeros(9)
        f = frame.c()
        check(self.L.cg_force_noalloc(self.h, dptr(x), C.byref(f), dptr(F), C.byref(E), dptr(V)), self.h)
        return F, E.value, V

    def copy_probe(self, nbytes):
        check(self.L.cg_copy_probe(self.h, nbytes), self.h)

    def set_penalty(self, dist, scale):
        check(self.L.cg_set_penalty(self.h, dist, scale), self.h)

    def set_fit_columns(self, keep=None):
        """Fit only the given (increasing) columns, the others held at zero; None: all."""
        keep = [] if keep is None else list(keep)
        arr = (C.c_int * max(len(keep), 1))(*keep)
        check(self.L.cg_set_fit_columns(self.h, arr, len(keep)), self.h)

    TILE_CONFIGS = ("512 threads, widest tile", "256 threads, widest tile", "2 x 256 threads per CU, half tile",
                    "512 threads, at most 16 columns", "2 x 256 threads per CU, half tile, at most 16 columns")
    TILE_CLASSES = ("<128 atoms", "128-511 atoms", "512-1023 atoms", ">=1024 atoms")

    def autotune(self, frames, prec=CG_FP64):
        """Pick the assembly tile configuration per (family, frame-size class) by timing
        (see cg_autotune); returns (choice [family][class], -1 for a class without frames,
        kernel ms [config][family][class])."""
        n = len(frames)
        cf = (Frame * n)(*[f.c() for f in frames])
        nk, nc = len(self.TILE_CONFIGS), len(self.TILE_CLASSES)
        choice = (C.c_int * (3 * nc))()
        ms = np.zeros(nk * 3 * nc)
        check(self.L.cg_autotune(self.h, cf, n, prec, choice, dptr(ms)), self.h)
        return np.array(list(choice)).reshape(3, nc), ms.reshape(nk, 3, nc)

    def tile_set(self, choice):
        """Use the given tile configuration [family][frame-size class]."""
        flat = [int(v) for v in np.asarray(choice).reshape(-1)]
        check(self.L.cg_tile_set(self.h, (C.c_int * len(flat))(*flat)), self.h)

    def tile_choice(self):
        """Tile configuration in use, [family][frame-size class]."""
        nc = len(self.TILE_CLASSES)
        choice = (C.c_int * (3 * nc))()
        check(self.L.cg_tile_choice(self.h, choice), self.h)
        return np.array(list(choice)).reshape(3, nc)

    def assembly_graphs(self, on):
        """Capture repeated-shape assemblies into HIP graphs (on by default) or keep them eager."""
        check(self.L.cg_assembly_graphs(self.h, 1 if on else 0), self.h)

    def assembly_deterministic(self, on):
        """Bitwise-reproducible assembly (waves add to the LDS tile in a fixed order)."""
        check(self.L.cg_assembly_deterministic(self.h, 1 if on else 0), self.h)

    def assembly_pair_tables(self, on):
        """3/4-body pair tables computed once per neighbour entry (on by default) or rebuilt per column
        tile; returns the families whose assemblies use them (bit 1: 3-body, bit 2: 4-body)."""
        act = C.c_int(0)
        check(self.L.cg_assembly_pair_tables(self.h, 1 if on else 0, C.byref(act)), self.h)
        return act.value

    def assembly_global_above(self, natoms):
        """Accumulate frames with more than natoms atoms in A with global atomics (0: only those beyond
        the LDS tile)."""
        check(self.L.cg_assembly_global_above(self.h, int(natoms)), self.h)

    def assembly_stats(self):
        """(assembly calls, calls repeated with exact list sizes, graph replays, graph captures)."""
        v = [C.c_int64() for _ in range(4)]
        check(self.L.cg_assembly_stats(self.h, *[C.byref(x) for x in v]), self.h)
        return tuple(x.value for x in v)

    def force_stats(self):
        """(force calls, calls repeated with exact list sizes, calls served by a graph replay)."""
        calls, rep, gr = C.c_int64(), C.c_int64(), C.c_int64()
        check(self.L.cg_force_stats(self.h, C.byref(calls), C.byref(rep), C.byref(gr)), self.h)
        return calls.value, rep.value, gr.value

    def force_batch(self, x, frames):
        """Forces (list of [natoms*3]), energies [n] and virials [n, 9] of independent configurations."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = len(frames)
        tot = sum(f.natoms for f in frames)
        F = np.zeros(3 * tot)
        E = np.zeros(n)
        V = np.zeros(9 * n)
        cf = (Frame * n)(*[f.c() for f in frames])
        check(self.L.cg_force_batch(self.h, dptr(x), cf, n, dptr(F), dptr(E), dptr(V)), self.h)
        out, o = [], 0
        for f in frames:
            out.append(F[o:o + 3 * f.natoms])
            o += 3 * f.natoms
        return out, E, V.reshape(n, 9)


class Server:
    """Continuous-batching force server (cg_serve_*)."""

    def __init__(self, setup, x, workers=2, budget_bytes=0, max_batch_atoms=65536, device=0):
        self.setup = setup  # keeps the model alive
        self.L = lib()
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.h = C.c_void_p()
        check(self.L.cg_serve_create(setup.model_ptr, device, dptr(self.x), workers, budget_bytes,
                                     max_batch_atoms, C.byref(self.h)))
        self.natoms = {}

    def submit(self, frame):
        t = C.c_int64()
        f = frame.c()
        check(self.L.cg_serve_submit(self.h, C.byref(f), C.byref(t)))
        self.natoms[t.value] = frame.natoms
        return t.value

    def poll(self, ticket, wait=True):
        """(F, E, virial, latency_ms) or None when not done (wait=False)."""
        n = self.natoms[ticket]
        F = np.zeros(3 * n)
        E = C.c_double()
        V = np.zeros(9)
        lat = C.c_double()
        done = C.c_int()
        check(self.L.cg_serve_poll(self.h, ticket, 1 if wait else 0, C.byref(done), dptr(F), C.byref(E), dptr(V),
                                   C.byref(lat)))
        if not done.value:
            return None
        del self.natoms[ticket]
        return F, E.value, V, lat.value

    def hold(self, on=True):
        """Stop (on) or resume starting batches; queued requests are then admitted together."""
        check(self.L.cg_serve_hold(self.h, 1 if on else 0))

    def poll_status(self, ticket):
        """(status code, message, F, E, virial) of a finished request, without raising."""
        n = self.natoms[ticket]
        F = np.zeros(3 * n)
        E = C.c_double()
        V = np.zeros(9)
        lat = C.c_double()
        done = C.c_int()
        st = self.L.cg_serve_poll(self.h, ticket, 1, C.byref(done), dptr(F), C.byref(E), dptr(V), C.byref(lat))
        msg = self.L.cg_last_error(None) if st else b""
        del self.natoms[ticket]
        return st, (msg or b"").decode(), F, E.value, V

    def stats(self):
        b, r = C.c_int64(), C.c_int64()
        check(self.L.cg_serve_stats(self.h, C.byref(b), C.byref(r)))
        return b.value, r.value

    COUNTERS = ("batches", "requests", "isolated", "graph_replays", "graph_captures", "exact_repeats",
                "buffer_moves")

    def counters(self):
        """Server counters (cg_serve_counters) as a dict."""
        out = (C.c_int64 * len(self.COUNTERS))()
        check(self.L.cg_serve_counters(self.h, out, len(self.COUNTERS)))
        return dict(zip(self.COUNTERS, list(out)))

    def close(self):
        if self.h:
            self.L.cg_serve_destroy(self.h)
            self.h = C.c_void_p()
