"""Extract the golden fixtures used by the tests from the ChIMES-lsq test suite.

Run once in a tree that has the reference mounted (default /root/reference);
the outputs under tests/golden/ are data only (inputs and the reference's
expected outputs), trimmed to the frames the tests need:

  headers.json             per LSQ test: fm_setup.in text (+ MULTI list), and
                           sha256/length of the golden params.header and
                           ff_groups.map (byte-exact pins of the parser and the
                           cluster power-set tables).
  lsq/nonorth2/            full input (10 triclinic frames), golden A (8640x12),
                           b, force.txt and the fitted params.txt.
  lsq/test_4atoms.2/       last frame (2+3+4-body, C/H, exclusions) and its
                           12 golden A rows (12x1308).
  lsq/fitstrs/             last frame (96 atoms, diagonal stress rows) and the
                           176 golden A rows that survive in A.txt.ab.
  lsq/tatb/                SPLITFI + SKPFRMS 1 run over 36 ranks (36 frames,
                           C/H/O/N, orders 8/4): input, the per-rank dim files
                           and processed frames (frames.NNNN.log), the
                           concatenated b, b-labeled and natoms files, and
                           params.header / ff_groups.map (gzip).
  lsq/water/               FITCOUL cases (O/H, Ewald charge columns and charge-
                           constraint rows): the shared trajectory (gzip), per
                           case its fm_setup.in; h2o-4bcheby and -small the 154
                           golden A rows that survive in A.txt.ab (the last 152
                           force rows and the 2 constraint rows), h2o-4bcheby
                           its b and force.txt; skip_frames (36 ranks, round
                           robin) its b.txt (gzip) and dim.txt; lsq2 (same
                           setup) the row weights and the params.txt of the
                           weighted SVD, eps 1e-4, ridge 1e-3 and LassoLars
                           1e-4 fits.
  lsq/cases/<test>/        end-to-end runs of 19 LSQ tests (and their standard
                           output, fm_setup.out.gz): fm_setup.in, the
                           trajectory (gzip, the frames the run reads; the
                           water cases use lsq/water/input.xyzf.gz), force.txt
                           (gzip), the sha256 of b.txt, the fitted coefficients
                           (test_suite_params.txt) and the fit summary of
                           params.txt (meta.json).
  md/<test>/               force-evaluation cases of test_suite-md (two of
                           them with fitted charges: Ewald terms): params.txt
                           (gzip), the first frame
                           of the run's coordinate file and the reference's
                           forceout.txt (forces, kcal/mol/A, 14 digits).
"""
import gzip
import hashlib
import json
import os
import re
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
LSQ = os.path.join(REF, "test_suite-lsq")

MD_TESTS = ["verify-lsq-forces-special3b", "verify-lsq-forces-special4b", "verify-lsq-forces-stress-and-ener-2b1",
            "verify-lsq-forces-stress-and-ener-4b", "verify-lsq-forces-stress-and-ener-4b2",
            "verify-lsq-forces-h2o-3bcheby", "verify-lsq-forces-h2o-4bcheby"]

HEADER_TESTS = ["nonorth2", "test_4atoms", "test_4atoms.2", "fitstrs", "fitstrs3b", "h2o-invr", "special3b",
                "special4b", "stress-and-ener-2b1", "stress-and-ener-2b2", "stress-and-ener-2b3",
                "stress-and-ener-4b", "stress-and-ener-4b2", "two-traj-files", "one-file-three-trajs",
                "h2o-2bcheby", "h2o-3bcheby", "h2o-3bcheby2", "h2o-3bcheby3", "h2o-4bcheby", "h2o-4bcheby-small"]

# chimes_md runs of one step whose forceout.txt holds the forces of the
# initial configuration, with the run's parameter file (older layout with an
# S_DELTA column, Tersoff cutoffs; one triclinic cell).  Runs of many steps
# print forces after moving the atoms; carbon-penalty's file lists power sets
# of an older cluster enumeration.
MD_RUNS = ["ti-nonortho", "ti-ortho"]

# end-to-end cases: chimes_lsq + chimes_lsq.py runs whose inputs are kept
# (trimmed to the frames the run reads), with their b.txt digest and force.txt
E2E_TESTS = ["fitstrs", "fitstrs3b", "special3b", "special4b", "stress-and-ener-2b1", "stress-and-ener-2b2",
             "stress-and-ener-2b3", "stress-and-ener-4b", "stress-and-ener-4b2", "one-file-three-trajs",
             "test_4atoms", "test_4atoms.2", "h2o-invr", "two-traj-files", "h2o-2bcheby", "h2o-3bcheby",
             "h2o-2bcheby2", "h2o-3bcheby2", "h2o-3bcheby3"]

# FITCOUL (Ewald charge columns) cases; they share one 99-frame water trajectory
WATER_TESTS = ["h2o-2bcheby", "h2o-4bcheby", "h2o-4bcheby-small", "skip_frames"]


def sha(path):
    data = open(path, "rb").read()
    return hashlib.sha256(data).hexdigest(), len(data)


def frames_of(path):
    """Split an .xyzf file into frame text blocks."""
    lines = open(path).read().split("\n")
    out, i = [], 0
    while i < len(lines) and lines[i].strip():
        n = int(lines[i].split()[0])
        out.append("\n".join(lines[i:i + n + 2]) + "\n")
        i += n + 2
    return out


def tail_rows(path):
    """Rows of a split A.txt.* piece; the first line is a partial row (byte split)."""
    lines = [l for l in open(path).read().split("\n")[1:] if l.strip()]
    return np.array([[float(x) for x in l.split()] for l in lines])


def gz(src, dst):
    """Deterministic gzip copy (no name, mtime 0)."""
    with open(dst, "wb") as raw:
        with gzip.GzipFile(fileobj=raw, mode="wb", compresslevel=9, mtime=0, filename="") as g:
            g.write(open(src, "rb").read())


def setup_of(d):
    """The fm_setup.in the golden outputs were produced from: correct_output's copy
    when the test keeps one (tatb's top-level file has a different 4-body order)."""
    co = os.path.join(d, "correct_output", "fm_setup.in")
    return co if os.path.exists(co) else os.path.join(d, "fm_setup.in")


def set_key(text, key, value):
    return re.sub(r"(# %s #[^\n]*\n)[^\n]*\n" % key, lambda m: m.group(1) + "\t%s\n" % value, text, count=1)


def main():
    headers = {}
    for t in HEADER_TESTS:
        d = os.path.join(LSQ, t)
        if not os.path.exists(os.path.join(d, "correct_output", "params.header")):
            continue
        ent = {"fm_setup": open(setup_of(d)).read()}
        if os.path.exists(os.path.join(d, "traj_list.dat")):
            ent["traj_list"] = open(os.path.join(d, "traj_list.dat")).read()
        for fn in ["params.header", "ff_groups.map"]:
            h, n = sha(os.path.join(d, "correct_output", fn))
            ent[fn] = {"sha256": h, "bytes": n}
        tsp = os.path.join(d, "correct_output", "test_suite_params.txt")
        if os.path.exists(tsp):
            # fitted coefficients of the reference run + its params.txt (Date line
            # dropped) as a digest: pins the params.txt printer for this header
            ent["x"] = [float(l.split()[1]) for l in open(tsp) if l.strip()]
            lines = open(os.path.join(d, "correct_output", "params.txt")).read().split("\n")
            body = "\n".join(lines[1:]).encode()
            ent["params_txt"] = {"sha256": hashlib.sha256(body).hexdigest(), "bytes": len(body),
                                 "summary": [l for l in lines[1:13] if l.startswith("!")]}
        dim = os.path.join(d, "correct_output", "dim.txt")
        if os.path.exists(dim):
            ent["dim"] = open(dim).read().split()
        headers[t] = ent
    json.dump(headers, open(os.path.join(HERE, "headers.json"), "w"), indent=1, sort_keys=True)

    # nonorth2: full
    out = os.path.join(HERE, "lsq", "nonorth2")
    os.makedirs(out, exist_ok=True)
    src = os.path.join(LSQ, "nonorth2")
    open(os.path.join(out, "fm_setup.in"), "w").write(open(os.path.join(src, "fm_setup.in")).read())
    open(os.path.join(out, "input.xyzf"), "w").write(open(os.path.join(src, "input.xyzf")).read())
    A = np.loadtxt(os.path.join(src, "correct_output", "A.txt.aa"))
    b = np.loadtxt(os.path.join(src, "correct_output", "b.txt"))
    y = np.loadtxt(os.path.join(src, "correct_output", "force.txt"))
    np.savez_compressed(os.path.join(out, "golden.npz"), A=A, b=b, force=y, row0=0)
    for fn in ["params.txt", "test_suite_params.txt", "dim.txt"]:
        open(os.path.join(out, fn), "w").write(open(os.path.join(src, "correct_output", fn)).read())
    gz(os.path.join(src, "correct_output", "fm_setup.out"), os.path.join(out, "fm_setup.out.gz"))

    # last frame of test_4atoms.2 and fitstrs
    for t, traj, apiece, frame_rows in [("test_4atoms.2", "dump2.xyzf", "A.txt.ab", 12),
                                        ("fitstrs", "run_5.OUTCAR.xyzf", "A.txt.ab", 291)]:
        out = os.path.join(HERE, "lsq", t)
        os.makedirs(out, exist_ok=True)
        src = os.path.join(LSQ, t)
        fr = frames_of(os.path.join(src, traj))
        text = open(os.path.join(src, "fm_setup.in")).read()
        # the run uses the first NFRAMES frames of the file; the golden rows
        # belong to frame NFRAMES-1
        nframes = int(re.search(r"# NFRAMES #[^\n]*\n\s*(\d+)", text).group(1))
        open(os.path.join(out, "traj.xyzf"), "w").write(fr[nframes - 1])
        text = set_key(set_key(text, "NFRAMES", "1"), "TRJFILE", "traj.xyzf")
        open(os.path.join(out, "fm_setup.in"), "w").write(text)
        rows = tail_rows(os.path.join(src, "correct_output", apiece))
        np.savez_compressed(os.path.join(out, "golden.npz"), A=rows, row0=frame_rows - rows.shape[0])
    # tatb: split A files over 36 ranks, round-robin frame order
    src = os.path.join(LSQ, "tatb")
    out = os.path.join(HERE, "lsq", "tatb")
    os.makedirs(out, exist_ok=True)
    co = os.path.join(src, "correct_output")
    open(os.path.join(out, "fm_setup.in"), "w").write(open(setup_of(src)).read())
    gz(os.path.join(src, "SCAN_all.xyzf"), os.path.join(out, "SCAN_all.xyzf.gz"))
    for fn in ["b.txt", "b-labeled.txt", "natoms.txt", "params.header", "ff_groups.map"]:
        gz(os.path.join(co, fn), os.path.join(out, fn + ".gz"))
    meta = {"nranks": 36, "dims": {}, "frames": {}}
    for r in range(36):
        d = os.path.join(co, "dim.%04d.txt" % r)
        if os.path.exists(d):
            meta["dims"][str(r)] = open(d).read()
        log = open(os.path.join(co, "frames.%04d.log" % r)).read()
        meta["frames"][str(r)] = [int(v) for v in re.findall(r"Processing frame (\d+)", log)]
    json.dump(meta, open(os.path.join(out, "split.json"), "w"), indent=1, sort_keys=True)
    # FITCOUL water cases
    wat = os.path.join(HERE, "lsq", "water")
    os.makedirs(wat, exist_ok=True)
    gz(os.path.join(LSQ, "h2o-4bcheby", "input.xyzf"), os.path.join(wat, "input.xyzf.gz"))
    for t in WATER_TESTS:
        src = os.path.join(LSQ, t)
        assert open(os.path.join(src, "input.xyzf"), "rb").read() == open(os.path.join(LSQ, "h2o-4bcheby", "input.xyzf"), "rb").read()
        out = os.path.join(wat, t)
        os.makedirs(out, exist_ok=True)
        open(os.path.join(out, "fm_setup.in"), "w").write(open(setup_of(src)).read())
        co = os.path.join(src, "correct_output")
        if os.path.exists(os.path.join(co, "A.txt.ab")):
            rows = tail_rows(os.path.join(co, "A.txt.ab"))
            nb = sum(1 for l in open(os.path.join(co, "b.txt")) if l.strip())
            extra = {}
            if t == "h2o-4bcheby":
                extra = {"b": np.loadtxt(os.path.join(co, "b.txt")), "force": np.loadtxt(os.path.join(co, "force.txt"))}
            np.savez_compressed(os.path.join(out, "golden.npz"), A=rows, row0=nb - rows.shape[0], **extra)
        gz(os.path.join(co, "fm_setup.out"), os.path.join(out, "fm_setup.out.gz"))
        if t == "skip_frames":
            gz(os.path.join(co, "b.txt"), os.path.join(out, "b.txt.gz"))
            open(os.path.join(out, "dim.txt"), "w").write(open(os.path.join(co, "dim.txt")).read())
    # lsq2 (the h2o-4bcheby setup): chimes_lsq.py runs with row weights, another
    # eps and ridge regression (its Makefile), their params.txt and the weights
    src = os.path.join(LSQ, "lsq2")
    assert open(os.path.join(src, "fm_setup.in")).read() == open(os.path.join(LSQ, "h2o-4bcheby", "fm_setup.in")).read()
    out = os.path.join(wat, "lsq2")
    os.makedirs(out, exist_ok=True)
    gz(os.path.join(src, "weights.txt"), os.path.join(out, "weights.txt.gz"))
    for fn in ["params.svd.weight.txt", "params.svd.1.0e-04.txt", "params.ridge.1.0e-03.txt",
               "params.lassolars.1.0e-04.txt"]:
        gz(os.path.join(src, "correct_output", fn), os.path.join(out, fn + ".gz"))
    # end-to-end cases
    water_md5 = hashlib.sha256(open(os.path.join(LSQ, "h2o-4bcheby", "input.xyzf"), "rb").read()).hexdigest()
    for t in E2E_TESTS:
        src = os.path.join(LSQ, t)
        out = os.path.join(HERE, "lsq", "cases", t)
        os.makedirs(out, exist_ok=True)
        co = os.path.join(src, "correct_output")
        text = open(setup_of(src)).read()
        open(os.path.join(out, "fm_setup.in"), "w").write(text)
        nframes = int(re.search(r"# NFRAMES #[^\n]*\n\s*(\d+)", text).group(1))
        trj = re.search(r"# TRJFILE #[^\n]*\n\s*(\S+)", text).group(1)
        meta = {"b_sha256": sha(os.path.join(co, "b.txt"))[0], "b_bytes": sha(os.path.join(co, "b.txt"))[1],
                "trajectories": [],
                "x": [float(l.split()[1]) for l in open(os.path.join(co, "test_suite_params.txt")) if l.strip()],
                "summary": [l for l in open(os.path.join(co, "params.txt")).read().split("\n")[1:13]
                            if l.startswith("!")]}
        if trj == "MULTI":
            lst = re.search(r"# TRJFILE #[^\n]*\n\s*MULTI\s+(\S+)", text).group(1)
            open(os.path.join(out, lst), "w").write(open(os.path.join(src, lst)).read())
            files = [l.split()[1] for l in open(os.path.join(src, lst)).read().split("\n")[1:] if l.strip()]
        else:
            files = [trj]
        for fn in files:
            path = os.path.join(src, fn)
            if hashlib.sha256(open(path, "rb").read()).hexdigest() == water_md5:
                meta["trajectories"].append({"name": fn, "water": True})
                continue
            frs = frames_of(path)
            keep = frs[:nframes] if trj != "MULTI" else frs
            tmp = os.path.join(out, fn + ".tmp")
            open(tmp, "w").write("".join(keep))
            gz(tmp, os.path.join(out, fn + ".gz"))
            os.remove(tmp)
            meta["trajectories"].append({"name": fn, "water": False})
        gz(os.path.join(co, "force.txt"), os.path.join(out, "force.txt.gz"))
        # the run's standard output (36 MPI ranks, rank 0 printing)
        gz(os.path.join(co, "fm_setup.out"), os.path.join(out, "fm_setup.out.gz"))
        json.dump(meta, open(os.path.join(out, "meta.json"), "w"), indent=1, sort_keys=True)
    # force-evaluation cases: params.txt + first coordinate frame + forceout.txt
    md = os.path.join(REF, "test_suite-md")
    for t in MD_TESTS:
        src = os.path.join(md, t)
        out = os.path.join(HERE, "md", t.replace("verify-lsq-forces-", ""))
        os.makedirs(out, exist_ok=True)
        run = open(os.path.join(src, "run_md.in")).read()
        crd = re.search(r"# CRDFILE #[^\n]*\n\s*(\S+)", run).group(1)
        open(os.path.join(out, "frame.xyzf"), "w").write(frames_of(os.path.join(src, crd))[0])
        with open(os.path.join(out, "params.txt.gz"), "wb") as raw:
            with gzip.GzipFile(fileobj=raw, mode="wb", compresslevel=9, mtime=0, filename="") as g:
                g.write(open(os.path.join(src, "correct_output", "params.txt"), "rb").read())
        open(os.path.join(out, "forceout.txt"), "w").write(
            open(os.path.join(src, "correct_output", "forceout.txt")).read())
        # |E - E_QM| / N printed by the reference run (6 decimals), when it fits energies
        m = re.search(r"Absolute energy error = (\S+) kcal/mol/atom",
                      open(os.path.join(src, "correct_output", "run_md.out")).read())
        meta = {"abs_energy_error_per_atom": float(m.group(1)) if m else None}
        if m:
            # the run compares with the energy row of its CMPRFRC file (force.txt, the
            # fit's A x at 7 digits): frame 0's first value repeated on three rows
            # after the 3N force rows and the stress rows
            n = int(open(os.path.join(out, "frame.xyzf")).readline().split()[0])
            ft = os.path.join(src, "force.txt")
            if not os.path.exists(ft):
                ft = os.path.join(src, "correct_output", "force.txt")
            y = [float(v) for v in open(ft).read().split()[:3 * n + 12]]
            for k in (3 * n, 3 * n + 3, 3 * n + 9):
                if y[k] == y[k + 1] == y[k + 2]:
                    meta["forcetxt_energy"] = y[k]
                    break
        json.dump(meta, open(os.path.join(out, "expected.json"), "w"), indent=1, sort_keys=True)
    for t in MD_RUNS:
        src = os.path.join(md, t)
        out = os.path.join(HERE, "md", "run-" + t)
        os.makedirs(out, exist_ok=True)
        run = open(os.path.join(src, "run_md.in")).read()
        crd = re.search(r"# CRDFILE #[^\n]*\n\s*(\S+)", run).group(1)
        prm = re.search(r"# PRMFILE #[^\n]*\n\s*(\S+)", run).group(1)
        first = frames_of(os.path.join(src, crd))[0]
        open(os.path.join(out, "frame.xyzf"), "w").write(first)
        gz(os.path.join(src, prm), os.path.join(out, "params.txt.gz"))
        n = int(first.split()[0])
        vals = open(os.path.join(src, "correct_output", "forceout.txt")).read().split("\n")
        open(os.path.join(out, "forceout.txt"), "w").write("\n".join(vals[:3 * n]) + "\n")
        # the step line printed at time 0 (zero temperature: potential terms only)
        # and the stress tensor the run printed
        log = open(os.path.join(src, "correct_output", "run_md.out")).read()
        st = re.search(r"\n\s+1\s+0\.00\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)", log)
        sig = {k: float(v) for k, v in re.findall(r"sigma_(\w\w): (\S+) GPa", log)}
        json.dump({"abs_energy_error_per_atom": None, "ktot_per_atom": float(st.group(1)),
                   "vtot_per_atom": float(st.group(2)), "pressure_gpa": float(st.group(5)), "sigma_gpa": sig},
                  open(os.path.join(out, "expected.json"), "w"), indent=1, sort_keys=True)
    # chimes_md's CHEBYFIX case: the published HN3 2+3+4-body force field
    # (Tersoff cutoffs, special 3B/4B cutoffs, energy offsets, a trimmed
    # parameter file) with its run's CHEBYFIX setting and first frame.  The run
    # starts from velocities its own generator draws at 20000 K, so none of its
    # printed numbers is a function of this frame alone: the case pins the model
    # and frame for GPU-vs-oracle parity of every out-of-range mode.
    src = os.path.join(md, "chebyfix")
    out = os.path.join(HERE, "md", "chebyfix")
    os.makedirs(out, exist_ok=True)
    run = open(os.path.join(src, "run_md.in")).read()
    crd = re.search(r"# CRDFILE #[^\n]*\n\s*(\S+)", run).group(1)
    prm = re.search(r"# PRMFILE #[^\n]*\n\s*(\S+)", run).group(1)
    fix = re.search(r"# CHEBYFIX #[^\n]*\n\s*(\S+)\s+(\S+)", run)
    lines = open(os.path.join(src, crd)).read().split("\n")
    n = int(lines[0].split()[0])
    open(os.path.join(out, "frame.xyzf"), "w").write("\n".join(lines[:n + 2]) + "\n")
    gz(os.path.join(src, prm), os.path.join(out, "params.txt.gz"))
    json.dump({"chebyfix": fix.group(1), "smooth_distance": float(fix.group(2)), "nlayers": 1},
              open(os.path.join(out, "expected.json"), "w"), indent=1, sort_keys=True)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
