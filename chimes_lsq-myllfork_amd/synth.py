"""Seeded synthetic inputs for the benchmark configurations.

Each generator returns (fm_setup.in text, list of frame dicts) and can write a
reference-format .xyzf trajectory, so the same inputs drive the drop-in CLI and
the in-memory benchmark.  Forces are drawn in kcal/mol/A (the .xyzf writer
converts them to Hartree/Bohr as the reference reader expects,
src/ClassDefs.C:2632-2634), stress in GPa, energy in kcal/mol.
"""
import math

import numpy as np

FORCE_CONV = 627.50960803 * 1.889725989


def _min_image_dist2(d, cell, inv):
    f = d @ inv
    f -= np.round(f)
    c = f @ cell
    return np.einsum("ij,ij->i", c, c)


def random_sequential(rng, n, cell, dmin, max_tries=200000):
    """Uniform positions in a periodic cell with a minimum separation."""
    cell = np.asarray(cell, dtype=np.float64).reshape(3, 3)
    inv = np.linalg.inv(cell)
    pts = np.zeros((n, 3))
    k = 0
    tries = 0
    while k < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place %d atoms with dmin=%g" % (n, dmin))
        p = rng.random(3) @ cell
        if k and _min_image_dist2(pts[:k] - p, cell, inv).min() < dmin * dmin:
            continue
        pts[k] = p
        k += 1
    return pts


def fm_setup(types, masses, o2, o3, o4, smin, smax, lam, smax3=None, smax4=None, nframes=1,
             trj="input.xyzf", fitstrs="true", fitener="true", fcut="CUBIC"):
    """fm_setup.in text in the reference keyword format."""
    nt = len(types)
    lines = ["# TRJFILE #", "\t" + trj, "# WRAPTRJ #", "\tfalse", "# NFRAMES #", "\t%d" % nframes,
             "# NLAYERS #", "\t1", "# FITCOUL #", "\tfalse", "# FITSTRS #", "\t" + fitstrs,
             "# FITENER #", "\t" + fitener, "# PAIRTYP #", "\tCHEBYSHEV %d %d %d" % (o2, o3, o4),
             "# CHBTYPE #", "\tMORSE", "# USENEIG #", "\ttrue", "# NATMTYP #", "\t%d" % nt,
             "# TYPEIDX #\t# ATM_TYP #\t# ATMCHRG #\t# ATMMASS #"]
    for i, (t, m) in enumerate(zip(types, masses)):
        lines.append("%d\t\t%s\t\t0.0\t\t%g" % (i + 1, t, m))
    lines.append("# PAIRIDX #\t# ATM_TY1 #\t# ATM_TY1 #\t# S_MINIM #\t# S_MAXIM #\t# S_DELTA #\t# MORSE_LAMBDA #")
    k = 1
    pairs = [(i, i) for i in range(nt)] + [(i, j) for i in range(nt) for j in range(i + 1, nt)]
    for (i, j) in pairs:
        sm = smin[(types[i], types[j])] if isinstance(smin, dict) else smin
        lines.append("%d\t\t%s\t\t%s\t\t%g\t\t%g\t\t0.01\t\t%g" % (k, types[i], types[j], sm, smax, lam))
        k += 1
    lines += ["# FCUTTYP #", "\t" + fcut]
    if smax3 is not None and o3 > 0:
        lines.append("SPECIAL 3B S_MAXIM: ALL %g" % smax3)
    if smax4 is not None and o4 > 0:
        lines.append("SPECIAL 4B S_MAXIM: ALL %g" % smax4)
    lines.append("# ENDFILE #")
    return "\n".join(lines) + "\n"


def _frame(rng, pos, types_idx, cell, natoms, index):
    return dict(natoms=natoms, xyz=pos.reshape(-1), types=np.asarray(types_idx, dtype=np.int32),
                cell=np.asarray(cell, dtype=np.float64).reshape(9),
                forces=rng.normal(0.0, 1.0, 3 * natoms), stress=np.concatenate([rng.normal(0, 1, 3), np.zeros(3)]),
                energy=float(rng.normal(-100.0 * natoms, 10.0)), index=index)


def carbon(nframes=100, natoms=512, density=2.2, seed=102, o2=16, o3=9, o4=5):
    """4-body carbon: cubic box at `density` g/cc, 1.1 A minimum separation."""
    rng = np.random.default_rng(seed)
    L = (natoms * 12.011 / (density * 0.602214076)) ** (1.0 / 3.0)
    cell = np.diag([L, L, L])
    frames = []
    for i in range(nframes):
        pos = random_sequential(rng, natoms, cell, 1.1)
        frames.append(_frame(rng, pos, np.zeros(natoms), cell, natoms, i))
    setup = fm_setup(["C"], [12.011], o2, o3, o4, 1.0, 4.5, 1.4, smax3=4.0, smax4=3.5, nframes=nframes)
    return setup, frames, ["C"]


def water(nframes=20, nmol=64, L=12.42, seed=101, o2=12, o3=4, o4=0):
    """Water-like O/H: O on a jittered simple-cubic lattice, H at 0.96 A / 104.5 deg."""
    rng = np.random.default_rng(seed)
    n = round(nmol ** (1.0 / 3.0))
    a = L / n
    cell = np.diag([L, L, L])
    frames = []
    for i in range(nframes):
        pos, typ = [], []
        for ix in range(n):
            for iy in range(n):
                for iz in range(n):
                    o = (np.array([ix, iy, iz]) + 0.5) * a + rng.normal(0, 0.1, 3)
                    # random orientation: two unit vectors at 104.5 degrees
                    u = rng.normal(size=3); u /= np.linalg.norm(u)
                    v = rng.normal(size=3); v -= v.dot(u) * u; v /= np.linalg.norm(v)
                    th = math.radians(104.5)
                    h1 = o + 0.96 * u + rng.normal(0, 0.1, 3)
                    h2 = o + 0.96 * (math.cos(th) * u + math.sin(th) * v) + rng.normal(0, 0.1, 3)
                    pos += [o, h1, h2]
                    typ += [0, 1, 1]
        pos = np.array(pos)
        frames.append(_frame(rng, pos, typ, cell, len(typ), i))
    smin = {("O", "O"): 2.0, ("O", "H"): 0.5, ("H", "O"): 0.5, ("H", "H"): 0.8}
    setup = fm_setup(["O", "H"], [15.999, 1.008], o2, o3, o4, smin, 5.0, 1.4, smax3=4.0, nframes=nframes)
    return setup, frames, ["O", "H"]


def chno(nframes=8, nmin=32, nmax=2048, seed=104, o2=12, o3=6, o4=3, target_density=0.1):
    """C/H/N/O frames of log-uniform size in random LAMMPS-restricted triclinic cells."""
    rng = np.random.default_rng(seed)
    types = ["C", "H", "N", "O"]
    frames = []
    for i in range(nframes):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        comp = rng.dirichlet([1, 1, 1, 1])
        typ = rng.choice(4, size=n, p=comp)
        lens = rng.uniform(10, 30, 3)
        al, be, ga = np.radians(rng.uniform(70, 110, 3))
        vol_unit = math.sqrt(1 + 2 * math.cos(al) * math.cos(be) * math.cos(ga) - math.cos(al) ** 2
                             - math.cos(be) ** 2 - math.cos(ga) ** 2)
        scale = (n / target_density / (lens.prod() * vol_unit)) ** (1.0 / 3.0)
        lens = np.clip(lens * scale, 10.0, 30.0)
        a, b, c = lens
        ax = a
        bx, by = b * math.cos(ga), b * math.sin(ga)
        cx = c * math.cos(be)
        cy = (b * c * math.cos(al) - bx * cx) / by
        cz = math.sqrt(max(c * c - cx * cx - cy * cy, 1e-6))
        cell = np.array([[ax, 0, 0], [bx, by, 0], [cx, cy, cz]])
        pos = random_sequential(rng, n, cell, 0.9)
        frames.append(_frame(rng, pos, typ, cell, n, i))
    setup = fm_setup(types, [12.011, 1.008, 14.007, 15.999], o2, o3, o4, 0.8, 5.0, 1.4, smax3=4.0, smax4=3.5,
                     nframes=nframes)
    return setup, frames, types


def write_xyzf(path, frames, type_names, nonortho=False):
    """Reference .xyzf: N, comment line with cell [+ stress] [+ energy], atoms with forces in Hartree/Bohr."""
    with open(path, "w") as out:
        for fr in frames:
            n = fr["natoms"]
            cell = fr["cell"]
            out.write("%d\n" % n)
            if nonortho or np.count_nonzero(cell - np.diag(np.diag(cell.reshape(3, 3))).reshape(9)):
                head = "NON_ORTHO " + " ".join("%.10f" % v for v in cell)
            else:
                head = "%.10f %.10f %.10f" % (cell[0], cell[4], cell[8])
            head += " %.10f %.10f %.10f" % tuple(fr["stress"][:3])
            head += " %.10f" % fr["energy"]
            out.write(head + "\n")
            xyz = fr["xyz"].reshape(-1, 3)
            f = fr["forces"].reshape(-1, 3) / FORCE_CONV
            for k in range(n):
                out.write("%s %.12f %.12f %.12f %.12e %.12e %.12e\n" % (
                    type_names[fr["types"][k]], xyz[k, 0], xyz[k, 1], xyz[k, 2], f[k, 0], f[k, 1], f[k, 2]))
