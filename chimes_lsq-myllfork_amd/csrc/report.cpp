// The standard output of chimes_lsq (what the reference's test suite keeps as
// fm_setup.out and compares line by line) and its per-rank frames.NNNN.log
// files, for a run whose design matrix is assembled on the GPU.
//
// Reference behaviour reproduced (paths relative to the ChIMES-lsq tree):
//   input echo, cluster tables     INPUT::PARSE_INFILE_LSQ, CLUSTER_LIST::build_all/print
//                                  src/input.C:172-1338, src/Cluster.C:368-426,771-954,1293-1556
//                                  (text gathered by the setup parser)
//   run outline, parameter counts  src/chimes_lsq.C:164-368
//   cutoff warnings, layering      FRAME::READ_XYZF, BOX::IS_RCUT_SAFE src/ClassDefs.C:2166-2190,2531-2600,2655-2680
//   per-frame lines                process_frame, CLUSTER_LIST::update_minmax_cutoffs,
//                                  NEIGHBORS::UPDATE_3B/4B_INTERACTION, ZCalc_Ewald_Deriv
//                                  src/chimes_lsq.C:718-823, src/Cluster.C:957-1017,
//                                  src/ClassDefs.C:1513-1625, src/Ewald.C:428-605
//   closing statistics             print_bond_stats, CLUSTER_LIST::print_min_distances
//                                  src/chimes_lsq.C:474-531, src/Cluster.C:1569-1633
//   frames.NNNN.log                src/chimes_lsq.C:283-338,718-823
//
// The reference runs its frames on MPI ranks and only rank 0 prints; the
// lines that depend on which frames rank 0 processed (per-frame counts, the
// atom counts of its last frame) take the rank layout the caller states.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "setup.hpp"

using namespace cg;

namespace {

cg_status fail(const std::exception &e) {
  set_error(e.what());
  return CG_ERR_INPUT;
}

void put(FILE *out, const std::string &text) {
  if (!text.empty()) std::fwrite(text.data(), 1, text.size(), out);
  std::fflush(out);
}

// Cell quantities the report prints (BOX::UPDATE_CELL / UPDATE_EXTENT /
// UPDATE_VOLUME, src/ClassDefs.C:1740-2021,2194-2211).
struct Cell {
  double a[3], b[3], c[3];
  double vol, extent[3];
  bool ortho;
};

Cell cell_of(const double *h) {
  Cell k;
  for (int q = 0; q < 3; ++q) { k.a[q] = h[q]; k.b[q] = h[3 + q]; k.c[q] = h[6 + q]; }
  k.ortho = std::fabs(h[1]) + std::fabs(h[2]) + std::fabs(h[3]) + std::fabs(h[5]) + std::fabs(h[6]) + std::fabs(h[7]) <= 1.0e-10;
  auto mag = [](const double *v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); };
  auto ang = [&](const double *u, const double *v) {
    return std::acos((u[0] * v[0] + u[1] * v[1] + u[2] * v[2]) / (mag(u) * mag(v)));
  };
  const double la = mag(k.a), lb = mag(k.b), lc = mag(k.c);
  const double al = ang(k.b, k.c), be = ang(k.c, k.a), ga = ang(k.a, k.b);
  if (k.ortho) {
    k.vol = h[0] * h[4] * h[8];
  } else {
    k.vol = la * lb * lc *
            std::sqrt(1 + 2 * std::cos(al) * std::cos(be) * std::cos(ga) - std::cos(al) * std::cos(al) -
                      std::cos(be) * std::cos(be) - std::cos(ga) * std::cos(ga));
  }
  // LAMMPS-restricted tilts and the extents of the tilted box
  double xy = lb * std::cos(ga);
  if (std::fabs(xy) < 1e-12) xy = 0.0;
  double xz = lc * std::cos(be);
  if (std::fabs(xz) < 1e-12) xz = 0.0;
  const double ly = std::sqrt(lb * lb - xy * xy);
  double yz = (lb * lc * std::cos(al) - xy * xz) / ly;
  if (std::fabs(yz) < 1e-12) yz = 0.0;
  const double lz = std::sqrt(lc * lc - xz * xz - yz * yz);
  if (k.ortho) {
    k.extent[0] = h[0]; k.extent[1] = h[4]; k.extent[2] = h[8];
  } else {
    double xlo = std::min({0.0, xy, xz, xy + xz}), xhi = la + std::max({0.0, xy, xz, xy + xz});
    k.extent[0] = xhi - xlo;
    k.extent[1] = ly + std::max(0.0, yz) - std::min(0.0, yz);
    k.extent[2] = lz;
  }
  return k;
}

std::string tuplet(int natoms, bool plural) {
  std::string t = natoms == 3 ? "triplet" : "quadruplet";
  return plural ? t + "s" : t;
}

// optimal_ewald_params with accuracy 1e-6 (src/Ewald.C:428-475) and the
// k-vector count of ZCalc_Ewald_Deriv (0 < |n|^2 < 50 over [0, kmax]^3,
// src/Ewald.C:560-600).
void ewald_report(int natoms, const double *h, std::ostringstream &o) {
  const double ax = h[0], by = h[4], cz = h[8];
  const double accuracy = 1.0e-6, effort = 5.5, balance = 1.1, accfac = 0.8;
  const double V = ax * by * cz, minl = std::min(ax, std::min(by, cz));
  const double p = -std::log(accuracy) * accfac;
  double alpha = std::sqrt(M_PI) * std::pow(effort * natoms / (V * V), 1.0 / 6.0);
  double rc = balance * std::sqrt(p) / alpha;
  if (rc > 0.5 * minl) {
    rc = 0.5 * minl;
    alpha = std::sqrt(p) / rc;
  }
  double rkc = 2.0 * alpha * std::sqrt(p) / balance;
  const double lavg = std::pow(V, 1.0 / 3.0);
  const int kc = (int)std::ceil(rkc * lavg / (2.0 * M_PI));
  rkc = kc * 2.0 * M_PI / lavg;
  const double k_acc = std::exp(-rkc * rkc / (4.0 * alpha * alpha)) / (rkc * rkc);
  const double r_acc = std::erfc(alpha * rc) / rc;
  int totk = 0;
  for (int x = 0; x <= kc; ++x)
    for (int y = 0; y <= kc; ++y)
      for (int z = 0; z <= kc; ++z) {
        const int k2 = x * x + y * y + z * z;
        if (k2 != 0 && k2 < 50) ++totk;
      }
  char buf[256];
  std::snprintf(buf, sizeof buf,
                "\tEwald_Deriv:\n\tR-Space Ewald cutoff      = %13.6e\n\tR-Space accuracy estimate = %13.6e\n"
                "\tK-space accuracy estimate = %13.6e\n",
                rc, r_acc, k_acc);
  o << buf << "\tNumber of Ewald K-vectors = " << totk << "\n";
}

void minmax_cutoffs(const ClusterFamily &F, std::ostringstream &o) {
  o << "\tLatest " << F.natoms << "-body outer cutoff values: \n";
  double v = -1.0;
  for (size_t i = 0; i < F.types.size(); ++i) {
    o << "\t" << i << ":";
    for (double x : F.types[i].smax) { v = std::max(v, x); o << "  " << x; }
    o << "\n";
  }
  o << "     Max " << F.natoms << "-body cutoff = " << v << "\n";
  o << "\tLatest " << F.natoms << "-body inner cutoff values: \n";
  v = 1.0e10;
  for (size_t i = 0; i < F.types.size(); ++i) {
    o << "\t" << i << ":";
    for (double x : F.types[i].smin) { v = std::min(v, x); o << "  " << x; }
    o << "\n";
  }
  o << "     Min " << F.natoms << "-body cutoff = " << v << "\n\n";
}

void interaction_counts(const Setup &s, long long n3, long long n4, std::ostringstream &o) {
  if (s.o3 > 0) o << "Number of 3-body interactions = " << n3 << "\n";
  if (s.o4 > 0) o << "Number of 4-body interactions = " << n4 << "\n";
}

void cluster_stats(const ClusterFamily &F, const double *mins, const int64_t *counts, std::ostringstream &o) {
  o << "\tMinimum distances between atoms " << tuplet(F.natoms, false) << " pairs: (Angstr.)\n";
  for (size_t k = 0; k < F.types.size(); ++k) {
    const auto &c = F.types[k];
    if (c.excluded) continue;
    o << "\t\t" << k << "\t";
    for (const auto &p : c.pair_names) o << p << " ";
    for (int m = 0; m < F.npairs; ++m) {
      const double v = mins ? mins[k * F.npairs + m] : 1.0e10;
      if (v < 1.0e9) o << v << " ";
      else o << "NOT FOUND  ";
    }
    o << "\n";
  }
  o << "\tTotal number of configurations contributing to each " << tuplet(F.natoms, false) << " type:\n";
  for (size_t k = 0; k < F.types.size(); ++k) {
    const auto &c = F.types[k];
    o << "\t\t" << k << "\t";
    for (const auto &p : c.pair_names) o << p << " ";
    if (!c.excluded) o << (counts ? counts[k] : 0) << "\n";
    else o << "excluded\n";
  }
}

}  // namespace

extern "C" {

cg_status cg_report_setup(const cg_setup *cs, FILE *out) {
  if (!cs || !out) { set_error("null argument"); return CG_ERR_ARG; }
  try {
    const Setup &s = *cs->s;
    std::ostringstream o;
    o << "\nReading input file...\n";
    o << "Will read from file: " << s.path << "\n";
    o << "Input file successfully opened.\nInput file has been rewound.\nInput file contents successfully stored.\n";
    o << s.echo.str();
    o << "...input file read successful: \n\n";
    o << "The number of two-body non-coulomb parameters is: " << s.n2b << "\n";
    if (s.o3 > 0) o << "The number of three-body Chebyshev parameters is: " << s.n3b << "\n";
    if (s.o4 > 0) o << "The number of four-body  Chebyshev parameters is: " << s.n4b << "\n";
    if (s.nframes > 0)
      o << "\nSuccefully opened the trajectory file...\nSetting up the matrices for A, Coulomb forces, and overbonding...\n";
    else
      o << "No. frames set to 0; skipping A-mat construction, jumping directly to params.header/ff_groups.map print\n";
    put(out, o.str());
    return CG_OK;
  } catch (const std::exception &e) { return fail(e); }
}

cg_status cg_report_frame_read(const cg_setup *cs, const cg_frame *f, FILE *out) {
  if (!cs || !f || !out) { set_error("null argument"); return CG_ERR_ARG; }
  try {
    const Setup &s = *cs->s;
    const Cell k = cell_of(f->cell);
    const int L = s.nlayers, rep = 2 * L + 1;
    std::ostringstream o;
    // BOX::IS_RCUT_SAFE with the reference's warning (it continues the run)
    for (const auto &p : s.pairs) {
      const double *vec[3] = {k.a, k.b, k.c};
      const char *lab[3] = {"AX", "BY", "CZ"};
      bool safe = true;
      for (int q = 0; q < 3 && safe; ++q) {
        const double len = std::sqrt(vec[q][0] * vec[q][0] + vec[q][1] * vec[q][1] + vec[q][2] * vec[q][2]);
        if (p.smax >= 0.5 * len * rep) {
          o << lab[q] << ": " << len * rep << " cut: " << p.smax << "\n";
          safe = false;
        }
      }
      if (safe) continue;
      o << "WARNING: Outer cutoff greater than half of at least one layered cell vector at least one box length: "
        << p.smax << "\n";
      o << "\tFrame:                      " << f->index << "\n";
      o << "\tPair type:                  " << p.a1 << " " << p.a2 << "\n";
      o << "\tcell vectors (a)            " << k.a[0] << " " << k.a[1] << " " << k.a[2] << "\n";
      o << "\tcell vectors (b)            " << k.b[0] << " " << k.b[1] << " " << k.b[2] << "\n";
      o << "\tcell vectors (c)            " << k.c[0] << " " << k.c[1] << " " << k.c[2] << "\n";
      o << "\tExtent (x,y,z)              " << k.extent[0] << " " << k.extent[1] << " " << k.extent[2] << "\n";
      o << "\tLayers:                     " << L << "\n";
      for (const double *v : {k.a, k.b, k.c})
        o << "\tEffective cell vectors (a): " << v[0] * rep << " " << v[1] * rep << " " << v[2] * rep << "\n";
    }
    if (f->index == 0) {
      o << "\tReporting outcome of layering for first frame ONLY: \n";
      o << "\tReal atoms:                   " << f->natoms << "\n";
      o << "\tTotal atoms (ghost):          " << (long long)f->natoms * rep * rep * rep << "\n";
      o << "\tcell vectors (a)            " << k.a[0] << " " << k.a[1] << " " << k.a[2] << "\n";
      o << "\tcell vectors (b)            " << k.b[0] << " " << k.b[1] << " " << k.b[2] << "\n";
      o << "\tcell vectors (c)            " << k.c[0] << " " << k.c[1] << " " << k.c[2] << "\n";
      o << "\tcell volume (A^3)           " << k.vol << "\n";
      o << "\tExtent (x,y,z)              " << k.extent[0] << " " << k.extent[1] << " " << k.extent[2] << "\n";
      o << "\tLayers:                     " << L << "\n";
      for (const double *v : {k.a, k.b, k.c})
        o << "\tEffective cell vectors (a): " << v[0] * rep << " " << v[1] * rep << " " << v[2] * rep << "\n";
    }
    put(out, o.str());
    return CG_OK;
  } catch (const std::exception &e) { return fail(e); }
}

cg_status cg_report_frame_done(const cg_setup *cs, const cg_frame *f, int first, int nranks, int64_t flat3,
                               int64_t flat4, FILE *out) {
  if (!cs || !f || !out || nranks < 1) { set_error("bad argument"); return CG_ERR_ARG; }
  try {
    Setup &s = *cs->s;
    std::ostringstream o;
    if (first) {
      o << "...matrix setup complete: \n\n...Populating the matrices for A, Coulomb forces, and overbonding...\n\n";
      s.report_ewald_cell[0] = s.report_ewald_cell[1] = s.report_ewald_cell[2] = 0.0;
    }
    if (nranks == 1) o << "\tProcessing frame: " << std::setw(5) << f->index + 1 << " of: " << s.nframes << "\n";
    if (first) {
      if (s.o3 > 0) minmax_cutoffs(s.trips, o);
      if (s.o4 > 0) minmax_cutoffs(s.quads, o);
    }
    interaction_counts(s, flat3, flat4, o);
    if (s.fitcoul) {
      // ZCalc_Ewald_Deriv rebuilds the lists and prints its parameters when the
      // cell changed, and for every frame in a parallel run
      const bool changed = nranks > 1 || f->cell[0] != s.report_ewald_cell[0] ||
                           f->cell[4] != s.report_ewald_cell[1] || f->cell[8] != s.report_ewald_cell[2];
      if (changed) {
        s.report_ewald_cell[0] = f->cell[0];
        s.report_ewald_cell[1] = f->cell[4];
        s.report_ewald_cell[2] = f->cell[8];
        interaction_counts(s, flat3, flat4, o);
        ewald_report(f->natoms, f->cell, o);
      }
    }
    put(out, o.str());
    return CG_OK;
  } catch (const std::exception &e) { return fail(e); }
}

cg_status cg_report_finish(const cg_setup *cs, int64_t force_rows, const int *last_type_counts,
                           const cg_lsq_stats *st, FILE *out) {
  if (!cs || !out) { set_error("null argument"); return CG_ERR_ARG; }
  try {
    const Setup &s = *cs->s;
    std::ostringstream o;
    if (s.nframes > 0) {
      o << "...matrix population complete: \n\nPrinting matrices...\n";
      o << "\t...A matrix length (forces): " << force_rows << "\n\n";
      if (s.fit_ener) {
        o << "Per-atom type energies will be provided at the end of the parameter file.\n";
        o << "Values will correspond to the following atom types, in order: \n";
        for (int t = 0; t < s.ntypes; ++t)
          o << "\t" << s.type_names[t] << " " << (last_type_counts ? last_type_counts[t] : 0) << "\n";
      }
    }
    o << "\tMinimum distances between atoms: (Angstr.)\n";
    for (size_t k = 0; k < s.pairs.size(); ++k) {
      const double v = st && st->pair_min ? st->pair_min[k] : 1.0e10;
      o << "\t\t" << k << "\t" << s.pairs[k].a1 << " " << s.pairs[k].a2 << "\t" << std::fixed << std::setprecision(8)
        << v << std::defaultfloat << std::setprecision(6) << "\n";
    }
    o << "\tTotal number of configurations contributing to each pair type:\n";
    for (size_t k = 0; k < s.pairs.size(); ++k)
      o << "\t\t" << k << "\t" << s.pairs[k].a1 << " " << s.pairs[k].a2 << " "
        << (st && st->pair_count ? st->pair_count[k] : 0) << "\n";
    o << std::fixed << std::setprecision(8);
    if (s.o3 > 0) cluster_stats(s.trips, st ? st->trip_min : nullptr, st ? st->trip_count : nullptr, o);
    if (s.o4 > 0) cluster_stats(s.quads, st ? st->quad_min : nullptr, st ? st->quad_count : nullptr, o);
    o << "...matrix printing complete: \n\n";
    put(out, o.str());
    return CG_OK;
  } catch (const std::exception &e) { return fail(e); }
}

// frames.NNNN.log of every rank: the file switches every rank sees while it
// reads the whole trajectory, and per processed frame its index, the atom
// counts per type and the time it took.
cg_status cg_write_frame_logs(const cg_setup *cs, const char *outdir, int nranks, const cg_frame *frames,
                              const double *seconds) {
  if (!cs || !frames || nranks < 1) { set_error("bad argument"); return CG_ERR_ARG; }
  try {
    const Setup &s = *cs->s;
    const int nf = s.nframes;
    std::vector<int> order(std::max(nf, 1)), rank(std::max(nf, 1));
    int n = 0;
    cg_status st = cg_frame_order(cs, nranks, order.data(), rank.data(), &n);
    if (st != CG_OK) return st;
    std::vector<int> rank_of(nf, -1);
    for (int i = 0; i < n; ++i) rank_of[order[i]] = rank[i];
    const std::string dir = outdir ? outdir : ".";
    for (int r = 0; r < nranks; ++r) {
      char name[64];
      std::snprintf(name, sizeof name, "/frames.%04d.log", r);
      std::ofstream log(dir + name);
      if (!log) throw Error(CG_ERR_INPUT, std::string("cannot write ") + name);
      int file_idx = 0, offset = s.infile_frames.empty() ? nf : s.infile_frames[0];
      for (int i = 0; i < nf; ++i) {
        if (i + 1 > offset && file_idx + 1 < (int)s.infiles.size()) {
          log << "Closing file " << s.infiles[file_idx] << " and opening file " << s.infiles[file_idx + 1]
              << " to read frame " << i + 1 << "\n";
          ++file_idx;
          offset += s.infile_frames[file_idx];
        }
        if (rank_of[i] != r) continue;
        log << "Processing frame " << i << "\n";
        std::vector<int> cnt(s.ntypes, 0);
        for (int a = 0; a < frames[i].natoms; ++a) ++cnt[frames[i].type[a]];
        for (int t = 0; t < s.ntypes; ++t) log << s.type_names[t] << " " << cnt[t] << " ";
        log << "\n";
        log << "Finished frame " << i << " in " << std::fixed << std::setprecision(1) << (seconds ? seconds[i] : 0.0)
            << std::defaultfloat << std::setprecision(6) << " seconds\n";
      }
    }
    return CG_OK;
  } catch (const Error &e) {
    set_error(e.what());
    return e.code;
  } catch (const std::exception &e) { return fail(e); }
}

}  // extern "C"
