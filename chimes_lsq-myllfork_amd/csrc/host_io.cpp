// Host I/O of the drop-in boundary: trajectory reader (.xyzf), A/b/b-labeled/
// natoms/dim text writers, params.header and ff_groups.map writers, and the
// C ABI wrappers around the setup parser.
//
// Formats follow FRAME::READ_XYZF (src/ClassDefs.C:2411-2683), A_MAT::PRINT_FRAME
// and CLEANUP_FILES (src/A_Matrix.C:210-760), print_param_header / print_map_file
// (src/chimes_lsq.C:562-716), FCUT::print_header (src/Fcut.C:169-178) and
// CLUSTER[_LIST]::print_header / print_special (src/Cluster.C:423-508,1636-1697).
#include <algorithm>
#include <charconv>
#include <cmath>
#include <ctime>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "setup.hpp"

using namespace cg;

namespace cg {
thread_local std::string g_err;
void set_error(const std::string &m) { g_err = m; }
}  // namespace cg

namespace {
cg_status fail(const Error &e) { g_err = e.what(); return e.code; }
cg_status fail(const std::exception &e) { g_err = e.what(); return CG_ERR_INPUT; }

constexpr double kForceConv = 627.50960803 * 1.889725989;  // Hartree/Bohr -> kcal/mol/A
constexpr double kGPa = 6.9476955;                          // src/functions.h:45

std::vector<std::string> split_ws(const std::string &line) {
  std::vector<std::string> out;
  std::istringstream is(line);
  std::string t;
  while (is >> t) out.push_back(t);
  return out;
}
}  // namespace

struct cg_traj {
  const Setup *s;
  std::string workdir;
  size_t file_idx = 0;
  int frame = 0;
  int offset = 0;
  std::ifstream in;
  std::vector<double> xyz, forces;
  std::vector<int> type;
};

extern "C" {

const char *cg_last_error(const void *) { return g_err.c_str(); }

cg_status cg_setup_read(const char *path, cg_setup **out) {
  if (!path || !out) { g_err = "null argument"; return CG_ERR_ARG; }
  try {
    *out = new cg_setup{read_setup(path)};
    return CG_OK;
  } catch (const Error &e) { return fail(e); } catch (const std::exception &e) { return fail(e); }
}

cg_status cg_params_read(const char *path, cg_setup **out) {
  if (!path || !out) { g_err = "null argument"; return CG_ERR_ARG; }
  try {
    *out = new cg_setup{read_params(path)};
    return CG_OK;
  } catch (const Error &e) { return fail(e); } catch (const std::exception &e) { return fail(e); }
}

cg_status cg_params_coeffs(const cg_setup *s, double *x, double *penalty_dist, double *penalty_scale) {
  if (!s || !x) { g_err = "null argument"; return CG_ERR_ARG; }
  if (s->s->coeffs.empty()) { g_err = "this setup was not read from a params file"; return CG_ERR_ARG; }
  std::copy(s->s->coeffs.begin(), s->s->coeffs.end(), x);
  if (penalty_dist) *penalty_dist = s->s->penalty_dist;
  if (penalty_scale) *penalty_scale = s->s->penalty_scale;
  return CG_OK;
}

cg_status cg_setup_set_chebyfix(cg_setup *s, int fix_type, double smooth_dist) {
  if (!s) { g_err = "null argument"; return CG_ERR_ARG; }
  if (fix_type != CG_FIX_ZERO_DERIV && fix_type != CG_FIX_CONSTANT_DERIV && fix_type != CG_FIX_SMOOTH) {
    g_err = "ERROR: CHEBYFIX was not recognized";
    return CG_ERR_ARG;
  }
  if (fix_type == CG_FIX_SMOOTH && !(smooth_dist > 0.0)) { g_err = "CHEBYFIX SMOOTH needs a positive distance"; return CG_ERR_ARG; }
  Setup &st = *s->s;
  st.fix_type = fix_type;
  st.model.fix_type = fix_type;
  if (fix_type == CG_FIX_SMOOTH) {
    st.smooth_dist = smooth_dist;
    st.model.smooth_dist = smooth_dist;
  }
  return CG_OK;
}

void cg_setup_free(cg_setup *s) {
  if (!s) return;
  delete s->s;
  delete s;
}

const cg_model *cg_setup_model(const cg_setup *s) { return s ? &s->s->model : nullptr; }
int cg_setup_nframes(const cg_setup *s) { return s ? s->s->nframes : 0; }

int64_t cg_model_frame_rows(const cg_model *m, int natoms, int fi) {
  int64_t r = 3LL * natoms;
  if (m->fit_stress != CG_STRESS_NONE && (m->nstress < 0 || fi < m->nstress))
    r += m->fit_stress == CG_STRESS_ALL ? 9 : 3;
  if (m->fit_energy && (m->nener < 0 || fi < m->nener)) r += 3;
  return r;
}

cg_status cg_traj_open(const cg_setup *cs, const char *workdir, cg_traj **out) {
  if (!cs || !out) { g_err = "null argument"; return CG_ERR_ARG; }
  auto t = new cg_traj;
  t->s = cs->s;
  t->workdir = workdir ? workdir : cs->s->dir;
  if (t->s->infiles.empty()) { delete t; g_err = "no trajectory file"; return CG_ERR_INPUT; }
  t->in.open(t->workdir + "/" + t->s->infiles[0]);
  if (!t->in) { g_err = "cannot open trajectory " + t->s->infiles[0]; delete t; return CG_ERR_INPUT; }
  t->offset = t->s->infile_frames.empty() ? t->s->nframes : t->s->infile_frames[0];
  *out = t;
  return CG_OK;
}

void cg_traj_close(cg_traj *t) { delete t; }

// FRAME::READ_XYZF.  Cell vectors are returned as given; cell-derived
// quantities are recomputed by the consumers exactly as BOX::UPDATE_CELL does.
cg_status cg_traj_next(cg_traj *t, cg_frame *f) {
  const Setup &s = *t->s;
  if (t->frame >= s.nframes) { g_err = "end of trajectory"; return CG_ERR_EOF; }
  try {
    if (t->frame + 1 > t->offset) {
      ++t->file_idx;
      if (t->file_idx >= s.infiles.size()) throw Error(CG_ERR_INPUT, "ran out of trajectory files");
      t->in.close();
      t->in.clear();
      t->in.open(t->workdir + "/" + s.infiles[t->file_idx]);
      if (!t->in) throw Error(CG_ERR_INPUT, "cannot open trajectory " + s.infiles[t->file_idx]);
      t->offset += s.infile_frames[t->file_idx];
    }
    const int fi = t->frame;
    int natoms = 0;
    if (!(t->in >> natoms)) throw Error(CG_ERR_INPUT, "Error: Reading frame " + std::to_string(fi) + ": missing atom count");
    std::string header;
    std::getline(t->in, header);
    std::getline(t->in, header);
    auto tk = split_ws(header);
    if (tk.size() < 3) throw Error(CG_ERR_INPUT, "Error:  Reading frame " + std::to_string(fi) + "\n        Missing box length components.\n" + header);
    bool ortho = tk[0] != "NON_ORTHO";
    std::memset(f, 0, sizeof(*f));
    if (ortho) {
      f->cell[0] = std::stod(tk[0]); f->cell[4] = std::stod(tk[1]); f->cell[8] = std::stod(tk[2]);
    } else {
      if (tk.size() < 10) throw Error(CG_ERR_INPUT, "Error: Reading frame " + std::to_string(fi) + ": NON_ORTHO needs 9 cell values");
      for (int i = 0; i < 9; ++i) f->cell[i] = std::stod(tk[1 + i]);
    }
    if (s.nstress < 0 || fi < s.nstress) {
      if (s.fit_stress == CG_STRESS_DIAG) {
        size_t b = ortho ? 3 : 10;
        if (tk.size() < b + 3) throw Error(CG_ERR_INPUT, "Error:  Reading frame " + std::to_string(fi) + "\n        Missing diagonal stress tensor components.\n" + header);
        for (int i = 0; i < 3; ++i) f->stress[i] = std::stod(tk[b + i]);
      } else if (s.fit_stress == CG_STRESS_ALL) {
        size_t b = ortho ? 3 : 10;
        if (tk.size() < b + 6) throw Error(CG_ERR_INPUT, "Error:  Reading frame " + std::to_string(fi) + "\n        Missing full stress tensor components.\n" + header);
        for (int i = 0; i < 6; ++i) f->stress[i] = std::stod(tk[b + i]);
      }
    }
    if ((s.nener < 0 || fi < s.nener) && s.fit_ener) f->energy = std::stod(tk.back());
    t->xyz.resize(3 * natoms);
    t->forces.resize(3 * natoms);
    t->type.resize(natoms);
    for (int j = 0; j < natoms; ++j) {
      std::string nm;
      t->in >> nm;
      int ty = type_of(s, nm);
      if (ty < 0) throw Error(CG_ERR_INPUT, "Did not recognize atom type " + nm);
      t->type[j] = ty;
      for (int c = 0; c < 3; ++c) t->in >> t->xyz[3 * j + c];
      for (int c = 0; c < 3; ++c) {
        t->in >> t->forces[3 * j + c];
        t->forces[3 * j + c] *= kForceConv;
      }
      if (!t->in) throw Error(CG_ERR_INPUT, "Error: Reading frame " + std::to_string(fi) + ": truncated atom line");
    }
    f->natoms = natoms;
    f->xyz = t->xyz.data();
    f->type = t->type.data();
    f->forces = t->forces.data();
    f->index = fi;
    ++t->frame;
    return CG_OK;
  } catch (const Error &e) { return fail(e); } catch (const std::exception &e) { return fail(e); }
}

// ------------------------------------------------------------ A/b text files
cg_status cg_write_open_opts(const cg_setup *cs, const char *outdir, int flags) {
  Setup &s = *cs->s;
  s.outdir = outdir ? outdir : ".";
  s.write_A = !(flags & CG_WRITE_NO_A);
  s.cur_rank = 0;
  s.rank_open = false;
  s.rank_rows.assign(std::max(s.nranks, 1), 0);
  s.abuf.clear();
  s.cons_written = false;
  if (s.write_A && !s.split_files) {
    s.fA.open(s.outdir + "/A.txt");
    if (!s.fA) { g_err = "Could not open A.txt in " + s.outdir; return CG_ERR_INPUT; }
  }
  s.fb.open(s.outdir + "/b.txt");
  s.fbl.open(s.outdir + "/b-labeled.txt");
  s.fna.open(s.outdir + "/natoms.txt");
  if (!s.fb || !s.fbl || !s.fna) { g_err = "Could not open output files in " + s.outdir; return CG_ERR_INPUT; }
  s.fb.precision(16); s.fb << std::scientific;
  s.data_count = 0;
  return CG_OK;
}

cg_status cg_write_open(const cg_setup *cs, const char *outdir) { return cg_write_open_opts(cs, outdir, 0); }

namespace {
std::string rank_file(const Setup &s, const char *stem, int rank) {
  char name[32];
  std::snprintf(name, sizeof name, "%s.%04d.txt", stem, rank);
  return s.outdir + "/" + name;
}

void flush_A(Setup &s) {
  if (!s.abuf.empty()) {
    s.fA.write(s.abuf.data(), (std::streamsize)s.abuf.size());
    s.abuf.clear();
  }
}

// "%.16e" of x (what the reference's `fileA << x` prints with precision 16
// and std::scientific), appended to the A buffer
void put_e16(std::string &out, double x) {
  char tmp[40];
  const auto r = std::to_chars(tmp, tmp + sizeof tmp, x, std::chars_format::scientific, 16);
  out.append(tmp, r.ptr);
}
}  // namespace

namespace {
// FITCOUL charge-constraint rows (A_MAT::PRINT_CONSTRAINTS,
// src/A_Matrix.C:567-624): zeros for the short-range columns, then the
// weights in pair-index order; b gets the constraint's value.  They go to
// `rank`'s part of the output: the first rank without frames in split mode,
// else the last rank.
void write_constraints(Setup &s, int rank) {
  if (s.charge_cons.empty() || s.cons_written) return;
  s.cons_written = true;
  const int np = (int)s.pairs.size();
  std::string t;
  for (const auto &cc : s.charge_cons) {
    for (int n = 0; n < s.nshort; ++n) t += "0.0 ";
    for (int j = 0; j < np; ++j)
      for (size_t k = 0; k < cc.pair_idx.size(); ++k)
        if (cc.pair_idx[k] == j) {
          put_e16(t, cc.weight[k]);
          t += ' ';
        }
    t += '\n';
    s.fb << cc.force << "\n";
    s.data_count += 1;
  }
  if ((int)s.rank_rows.size() <= rank) s.rank_rows.resize(rank + 1, 0);
  s.rank_rows[rank] += (int64_t)s.charge_cons.size();
  if (!s.write_A) return;
  if (!s.split_files || (rank == s.cur_rank && s.rank_open)) {
    s.abuf += t;
    return;
  }
  std::ofstream f(rank_file(s, "A", rank));  // a rank without frames: a new file
  f << t;
}
}  // namespace

cg_status cg_write_rank(const cg_setup *cs, int rank) {
  Setup &s = *cs->s;
  if (rank < 0 || rank < s.cur_rank) { g_err = "cg_write_rank: ranks must not decrease"; return CG_ERR_ARG; }
  if ((int)s.rank_rows.size() <= rank) s.rank_rows.resize(rank + 1, 0);
  if (s.split_files && s.write_A && rank != s.cur_rank && s.rank_open) {
    flush_A(s);
    s.fA.close();
    s.rank_open = false;
  }
  // ranks passed over without frames: the first takes the constraint rows
  if (s.split_files)
    for (int j = s.cur_rank; j < rank; ++j)
      if (s.rank_rows[j] == 0) { write_constraints(s, j); break; }
  s.cur_rank = rank;
  return CG_OK;
}

namespace {
// 2-body column n of an excluded pair type (A_MAT::skip_2b, src/A_Matrix.C:192-208)
bool skip_2b(const Setup &s, int n) {
  for (int q : s.excl_2b)
    if (n >= q * s.o2 && n < (q + 1) * s.o2) return true;
  return false;
}

void ones(const Setup &s, std::string &o, bool energy_row, const std::vector<int> &count) {
  if (!s.fit_ener) return;
  for (int i = 0; i < s.ntypes; ++i) {
    if (std::find(s.excl_1b.begin(), s.excl_1b.end(), i) != s.excl_1b.end()) continue;  // add_col_of_ones
    if (!energy_row) o += " 0.0";
    else { o += ' '; o += std::to_string(count[i]); }
  }
}
}  // namespace

cg_status cg_write_frame(const cg_setup *cs, const cg_frame *f, const double *rows, int nrows) {
  Setup &s = *cs->s;
  const cg_model &m = s.model;
  const int nc = m.ncols, ns = m.ncols_short;
  int64_t expect = cg_model_frame_rows(&m, f->natoms, f->index);
  if (nrows != expect) { g_err = "row count mismatch in cg_write_frame"; return CG_ERR_ARG; }
  if (s.write_A && !rows) { g_err = "cg_write_frame: rows are NULL"; return CG_ERR_ARG; }
  if (s.write_A && s.split_files && !s.rank_open) {
    s.fA.open(rank_file(s, "A", s.cur_rank));
    if (!s.fA) { g_err = "Could not open " + rank_file(s, "A", s.cur_rank); return CG_ERR_INPUT; }
    s.rank_open = true;
  }
  if ((int)s.rank_rows.size() <= s.cur_rank) s.rank_rows.resize(s.cur_rank + 1, 0);
  // which input file the frame came from (for the b-labeled flags)
  size_t file = 0;
  int check = 0;
  for (size_t i = 0; i < s.infile_frames.size(); ++i) {
    check += s.infile_frames[i];
    if (f->index < check) break;
    ++file;
  }
  file = std::min(file, s.force_flags.size() - 1);
  std::vector<int> count(s.ntypes, 0);
  for (int i = 0; i < f->natoms; ++i) count[f->type[i]]++;
  int total = 0;
  for (int c : count) total += c;
  int r = 0;
  auto row_out = [&](const char *sep, bool energy_row, bool force_row) {
    if (s.write_A) {
      const double *a = rows + (int64_t)r * nc;
      for (int n = 0; n < ns; ++n) {
        if (!s.excl_2b.empty() && skip_2b(s, n)) continue;
        put_e16(s.abuf, a[n]);
        s.abuf += sep;
      }
      // FITCOUL: the Ewald columns, on force rows only (src/A_Matrix.C:249-252)
      if (force_row)
        for (int q = 0; q < m.n_charge_cols; ++q) {
          put_e16(s.abuf, a[ns + q]);
          s.abuf += sep;
        }
      ones(s, s.abuf, energy_row, count);
      s.abuf += '\n';
      if (s.abuf.size() > (1u << 20)) flush_A(s);
    }
    s.fna << total << "\n";
    ++r;
  };
  for (int a = 0; a < f->natoms; ++a) {
    for (int c = 0; c < 3; ++c) row_out("   ", false, true);
    for (int c = 0; c < 3; ++c) s.fb << f->forces[3 * a + c] << "\n";
    s.data_count += 3;
    const std::string &nm = s.type_names[f->type[a]];
    for (int c = 0; c < 3; ++c) s.fbl << s.force_flags[file] << nm << " " << f->forces[3 * a + c] << "\n";
  }
  bool st = s.fit_stress != CG_STRESS_NONE && (s.nstress < 0 || f->index < s.nstress);
  if (st && s.fit_stress == CG_STRESS_DIAG) {
    for (int c = 0; c < 3; ++c) row_out(" ", false, false);
    const char *lab[3] = {"s_xx ", "s_yy ", "s_zz "};
    for (int c = 0; c < 3; ++c) s.fb << f->stress[c] / kGPa << "\n";
    s.data_count += 3;
    for (int c = 0; c < 3; ++c) s.fbl << s.stress_flags[file] << lab[c] << f->stress[c] / kGPa << "\n";
  } else if (st && s.fit_stress == CG_STRESS_ALL) {
    for (int c = 0; c < 9; ++c) row_out(" ", false, false);
    // xx xy xz / yx yy yz / zx zy zz from (xx yy zz xy xz yz)
    const int idx[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
    const char *lab[9] = {"s_xx ", "s_xy ", "s_xz ", "s_yx ", "s_yy ", "s_yz ", "s_zx ", "s_zy ", "s_zz "};
    for (int c = 0; c < 9; ++c) s.fb << f->stress[idx[c]] / kGPa << "\n";
    s.data_count += 9;
    for (int c = 0; c < 9; ++c) s.fbl << s.stress_flags[file] << lab[c] << f->stress[idx[c]] / kGPa << "\n";
  }
  if (s.fit_ener && (s.nener < 0 || f->index < s.nener)) {
    for (int c = 0; c < 3; ++c) row_out(" ", true, false);
    for (int c = 0; c < 3; ++c) {
      s.fb << f->energy << "\n";
      s.fbl << s.energy_flags[file] << "+1 " << f->energy << "\n";
    }
    s.data_count += 3;
  }
  s.rank_rows[s.cur_rank] += r;
  if ((s.write_A && !s.fA.good()) || !s.fb.good()) { g_err = "Error in A/b file"; return CG_ERR_INPUT; }
  return CG_OK;
}

cg_status cg_write_close(const cg_setup *cs) {
  Setup &s = *cs->s;
  if (!s.charge_cons.empty() && !s.cons_written) {
    int pr = std::max(s.nranks, 1) - 1;
    if (s.split_files)
      for (int j = s.cur_rank; j < std::max(s.nranks, 1); ++j)
        if (j >= (int)s.rank_rows.size() || s.rank_rows[j] == 0) { pr = j; break; }
    if (!s.split_files) pr = s.cur_rank;  // one A.txt: the rows go last
    write_constraints(s, pr);
  }
  if (s.write_A) {
    flush_A(s);
    if (s.fA.is_open()) s.fA.close();
  }
  s.rank_open = false;
  s.fb.close(); s.fbl.close(); s.fna.close();
  if (s.split_files) {
    // per rank with rows: columns, first and last row (global), total rows
    int64_t start = 0;
    for (size_t r = 0; r < s.rank_rows.size(); ++r) {
      const int64_t n = s.rank_rows[r];
      if (n > 0) {
        std::ofstream dim(rank_file(s, "dim", (int)r));
        if (!dim) { g_err = "Could not open " + rank_file(s, "dim", (int)r); return CG_ERR_INPUT; }
        dim << s.ncols - s.n_excl_lt3b << " " << start << " " << start + n - 1 << " " << s.data_count << "\n";
      }
      start += n;
    }
    return CG_OK;
  }
  std::ofstream dim(s.outdir + "/dim.txt");
  if (!dim) { g_err = "Could not open dim.txt"; return CG_ERR_INPUT; }
  dim << s.ncols - s.n_excl_lt3b << " " << s.data_count << std::endl;
  return CG_OK;
}

cg_status cg_setup_constraint_rows(const cg_setup *cs, double *rows, double *b, int *nrows) {
  if (!cs || !nrows) { g_err = "cg_setup_constraint_rows: bad argument"; return CG_ERR_ARG; }
  const Setup &s = *cs->s;
  const cg_model &m = s.model;
  *nrows = (int)s.charge_cons.size();
  for (size_t i = 0; i < s.charge_cons.size(); ++i) {
    const auto &cc = s.charge_cons[i];
    if (rows) {
      double *r = rows + i * (size_t)m.ncols;
      std::fill(r, r + m.ncols, 0.0);
      for (size_t k = 0; k < cc.pair_idx.size(); ++k) r[m.ncols_short + cc.pair_idx[k]] = cc.weight[k];
    }
    if (b) b[i] = cc.force;
  }
  return CG_OK;
}

cg_status cg_setup_fit_columns(const cg_setup *cs, int *keep, int *nkeep) {
  if (!cs || !keep || !nkeep) { g_err = "cg_setup_fit_columns: bad argument"; return CG_ERR_ARG; }
  const Setup &s = *cs->s;
  const cg_model &m = s.model;
  int n = 0;
  for (int c = 0; c < m.ncols; ++c) {
    bool drop = false;
    if (c < m.ncols_short) drop = !s.excl_2b.empty() && c < m.n2b_cols && skip_2b(s, c);
    else drop = std::find(s.excl_1b.begin(), s.excl_1b.end(), c - m.ncols_short) != s.excl_1b.end();
    if (!drop) keep[n++] = c;
  }
  *nkeep = n;
  return CG_OK;
}

cg_status cg_frame_order(const cg_setup *cs, int nranks, int *order, int *rank, int *n_out) {
  if (!cs || nranks < 1 || !order || !rank || !n_out) { g_err = "cg_frame_order: bad argument"; return CG_ERR_ARG; }
  Setup &s = *cs->s;
  const int nf = s.nframes;
  s.nranks = nranks;
  int n = 0;
  for (int r = 0; r < nranks; ++r) {
    // divide_atoms: ranks take ceil-divided contiguous blocks
    int i0 = 1, i1 = 0;
    if (nf > 0) {
      const int used = std::min(nranks, nf);
      i0 = (int)std::ceil((double)r * nf / used);
      if (r > nf) { i0 = nf + 1; i1 = nf - 1; }
      else if (r == used - 1) i1 = nf - 1;
      else i1 = std::min(nf - 1, (int)std::ceil((double)(r + 1) * nf / used) - 1);
    }
    for (int i = 0; i < nf; ++i) {
      bool take;
      if (s.skip_frames <= 0) take = i >= i0 && i <= i1;
      else take = (i + r) % nranks == 0 && ((i + r) / nranks) % s.skip_frames == 0;
      if (take) { order[n] = i; rank[n] = r; ++n; }
    }
  }
  *n_out = n;
  return CG_OK;
}

cg_status cg_write_header(const cg_setup *cs, const char *outdir) {
  try {
    std::string d = outdir ? outdir : ".";
    write_header(*cs->s, d);
    write_map(*cs->s, d);
    return CG_OK;
  } catch (const Error &e) { return fail(e); } catch (const std::exception &e) { return fail(e); }
}

}  // extern "C"

namespace cg {

static const char *trans_name(int t) {
  return t == CG_MORSE ? "MORSE" : t == CG_INVRSE_R ? "INVRSE_R" : "NONE";
}

static void print_special_family(std::ostream &h, const ClusterFamily &F) {
  if (!F.used || F.types.empty()) return;
  for (int mode = 0; mode < 2; ++mode) {  // S_MINIM then S_MAXIM
    int found = 0;
    for (auto &c : F.types)
      if ((mode == 0 ? c.special_min : c.special_max) && !c.excluded) ++found;
    if (!found) continue;
    h << std::endl << "SPECIAL " << F.natoms << "B " << (mode == 0 ? "S_MINIM" : "S_MAXIM") << ": SPECIFIC " << found << std::endl;
    for (size_t i = 0; i < F.types.size(); ++i) {
      const auto &c = F.types[i];
      if (!(mode == 0 ? c.special_min : c.special_max)) continue;
      if (c.excluded) continue;
      auto it = F.name_rev.find((int)i);
      std::string rev = it == F.name_rev.end() ? "" : it->second;
      if (mode == 0) h << " " << rev << " ";
      else h << rev << " ";
      for (auto &p : c.pair_names) h << p << " ";
      h << std::fixed << std::setprecision(5);
      for (double v : (mode == 0 ? c.smin : c.smax)) h << v << " ";
      h << std::endl;
    }
  }
}

static void print_family(std::ostream &h, const ClusterFamily &F, int natoms, int order) {
  std::string tup = natoms == 3 ? "TRIPLETS" : "QUADRUPLETS";
  h << std::endl << "ATOM PAIR " << tup << ": ";
  if (order == 0 || F.types.empty()) { h << 0 << std::endl; return; }
  h << F.types.size() << std::endl << std::endl;
  for (const auto &c : F.types) {
    h << " INDEX: " << c.index << " ATOMS: ";
    for (int m = 0; m < c.natoms; ++m) {
      h << c.atom_names[m];
      if (m < c.npairs - 1) h << " ";
    }
    h << std::endl << " PAIRS: ";
    for (int m = 0; m < c.npairs; ++m) {
      h << c.pair_names[m];
      if (m < c.npairs - 1) h << " ";
    }
    if (c.excluded) {
      h << " EXCLUDED:\n";
    } else {
      h << "  " << c.unique_powers.size() << " parameters, " << c.allowed.size() << " total parameters " << std::endl;
      h << "     index  |  powers  |  equiv index  |  param index  " << std::endl;
      h << "   ----------------------------------------------------" << std::endl;
      for (size_t j = 0; j < c.allowed.size(); ++j) {
        h << "      " << std::setw(6) << std::fixed << std::left << j << " ";
        h << " ";
        for (int m = 0; m < c.npairs; ++m) h << std::setw(2) << std::fixed << std::left << c.allowed[j][m] << " ";
        h << "       " << std::setw(8) << c.equiv_idx[j] << " ";
        h << "       " << std::setw(8) << c.param_idx[j] << std::endl;
      }
    }
    h << std::endl;
  }
}

std::string header_text(const Setup &s) {
  std::ostringstream h;
  h << "USECOUL: " << (s.fitcoul || s.use_partial_charges ? "true" : "false") << std::endl;
  h << "FITCOUL: " << (s.fitcoul ? "true" : "false") << std::endl;
  h << "USE3BCH: " << (s.o3 > 0 ? "true" : "false") << std::endl;
  h << "USE4BCH: " << (s.o4 > 0 ? "true" : "false") << std::endl;
  // EXCLD lines: excluded type / pair indices, each followed by a space
  // (src/chimes_lsq.C:592-610)
  if (!s.excl_1b.empty()) {
    h << "EXCLD1B: ";
    for (int t : s.excl_1b) h << t << " ";
    h << std::endl;
  } else {
    h << "EXCLD1B: false" << std::endl;
  }
  if (!s.excl_2b.empty()) {
    h << "EXCLD2B: ";
    for (int q : s.excl_2b) h << q << " ";
    h << std::endl;
  } else {
    h << "EXCLD2B: false" << std::endl;
  }
  h << std::endl << "PAIRTYP: " << "CHEBYSHEV" << " ";
  h << " " << s.o2 << " " << s.o3 << " " << s.o4 << " " << double(s.cheby_low) << " " << double(s.cheby_high) << std::endl;
  h << std::endl << "ATOM TYPES: " << s.ntypes << std::endl << std::endl;
  h << "# TYPEIDX #\t# ATM_TYP #\t# ATMCHRG #\t# ATMMASS #" << std::endl;
  // FITCOUL prints the charge sign of the input, else the fixed charge
  for (int i = 0; i < s.ntypes; ++i) {
    h << i << "\t\t" << s.pairs[i].a1 << "\t\t";
    if (s.fitcoul) h << s.pairs[i].chrgsgn;
    else h << s.pairs[i].chg1;
    h << "\t\t" << s.pairs[i].mass1 << std::endl;
  }
  h << std::endl << "ATOM PAIRS: " << s.pairs.size() << std::endl << std::endl;
  h << "# PAIRIDX #\t# ATM_TY1 #\t# ATM_TY1 #\t# S_MINIM #\t# S_MAXIM #\t# CHBDIST #\t# MORSE_LAMBDA #" << std::endl;
  for (const auto &p : s.pairs) {
    h << "\t" << std::setw(16) << std::left << p.idx << std::setw(16) << std::left << p.a1
      << std::setw(16) << std::left << p.a2 << std::setw(16) << std::left << p.smin
      << std::setw(16) << std::left << p.smax << std::setw(16) << std::left << trans_name(s.transform);
    if (s.transform == CG_MORSE) h << std::setw(16) << std::left << p.lambda << std::endl;
    else h << std::endl;
  }
  h << std::endl << "FCUT TYPE: " << (s.fcut_type == CG_FCUT_CUBIC ? "CUBIC" : "TERSOFF");
  if (s.fcut_type == CG_FCUT_TERSOFF) h << " " << s.fcut_offset;
  h << std::endl;
  print_special_family(h, s.trips);
  print_special_family(h, s.quads);
  print_family(h, s.trips, 3, s.o3);
  print_family(h, s.quads, 4, s.o4);
  h << std::endl;
  return h.str();
}

void write_header(const Setup &s, const std::string &dir) {
  std::ofstream h(dir + "/params.header");
  if (!h) throw Error(CG_ERR_INPUT, "cannot write params.header");
  h << header_text(s);
}

std::string map_text(const Setup &s) {
  std::ostringstream o;
  o << "PAIRMAPS: " << s.pair_map.size() << std::endl;
  for (auto &kv : s.pair_map) o << kv.second << " " << kv.first << std::endl;
  for (const ClusterFamily *F : {&s.trips, &s.quads}) {
    if (F->name_map.empty()) continue;
    o << std::endl << (F->natoms == 3 ? "TRIPMAPS: " : "QUADMAPS: ") << F->name_map.size() << std::endl;
    for (auto &kv : F->name_map) o << kv.second << " " << kv.first << std::endl;
  }
  return o.str();
}

void write_map(const Setup &s, const std::string &dir) {
  std::ofstream o(dir + "/ff_groups.map");
  if (!o) throw Error(CG_ERR_INPUT, "cannot write ff_groups.map");
  o << map_text(s);
}


// params.txt as printed by chimes_lsq.py (src/chimes_lsq.py:155-162,303-309,315-672):
// the fit summary, the header up to the cluster blocks, the 2-body coefficients,
// the cluster tables with a parameter column, the maps and the energy offsets.
// The cluster blocks are located by the same fixed line offsets the script uses.
static std::vector<std::string> split_lines(const std::string &t) {
  std::vector<std::string> v;
  size_t p = 0;
  while (p < t.size()) {
    size_t q = t.find('\n', p);
    if (q == std::string::npos) q = t.size();
    v.push_back(t.substr(p, q - p));
    p = q + 1;
  }
  return v;
}

static std::string rstrip(const std::string &x) {
  size_t e = x.find_last_not_of(" \t\r\n");
  return e == std::string::npos ? "" : x.substr(0, e + 1);
}

std::string params_text(const Setup &s, const double *x, int nx, long long nlines, const cg_fit_result &r,
                        const std::string &date) {
  std::ostringstream o;
  char buf[512];
  auto P = [&](const std::string &t) { o << t << "\n"; };
  auto fx = [&](const char *fmt, double v) { std::snprintf(buf, sizeof buf, fmt, v); return std::string(buf); };
  P("! Date  " + date);
  P("!");
  P("! Number of variables            =  " + std::to_string(nx));
  P("! Number of equations            =  " + std::to_string(nlines));
  if (r.algorithm == CG_ALG_RIDGE) {  // src/chimes_lsq.py:216-225
    P("! ridge regression used");
    P(fx("! Ridge alpha = %11.4e", r.alpha));
  } else if (r.algorithm == CG_ALG_LASSOLARS) {  // src/chimes_lsq.py:248-253
    P("! LARS implementation of LASSO used");
    P(fx("! LASSO alpha = %11.4e", r.alpha));
  } else {
    P("! svd algorithm used");
    P(fx("! eps (= args.eps*dmax)          =  %11.4e", r.eps_abs));
    P(fx("! SVD regularization factor      = %11.4e", r.eps));
  }
  P(fx("! RMS force error                = %11.4e", r.rms));
  P(fx("! max abs variable               = %11.4e", r.max_abs_x));
  P("! number of fitting vars         =  " + std::to_string(r.nvars));
  P(fx("! Bayesian Information Criterion = %11.4e", r.bic));
  const std::string wf(r.weights_file, strnlen(r.weights_file, sizeof r.weights_file));
  if (!wf.empty()) P("! Using weighting file:             " + wf);  // src/chimes_lsq.py:307-308
  P("!");
  std::vector<std::string> hf = split_lines(header_text(s));
  auto tok = [](const std::string &l) { return split_ws(l); };
  int trips_line = 0, quads_line = 0;
  long total_trips = 0, total_quads = 0;
  std::vector<int> excl2, excl1;  // EXCLD2B / EXCLD1B lists of the header
  for (size_t i = 0; i < hf.size(); ++i) {
    P(hf[i]);
    auto t = tok(hf[i]);
    bool brk = false;
    if (hf[i].find("EXCLD2B") != std::string::npos) {
      excl2.clear();
      if (!(t.size() == 2 && t[1] == "false"))
        for (size_t q = 1; q < t.size(); ++q) excl2.push_back(std::stoi(t[q]));
    }
    if (hf[i].find("EXCLD1B") != std::string::npos) {
      excl1.clear();
      if (!(t.size() == 2 && t[1] == "false"))
        for (size_t q = 1; q < t.size(); ++q) excl1.push_back(std::stoi(t[q]));
    }
    if (t.size() > 3 && t[2] == "TRIPLETS:") {
      total_trips = std::stol(t[3]);
      trips_line = (int)i;
      for (size_t j = i; j < hf.size(); ++j) {
        auto u = tok(hf[j]);
        if (u.size() > 3 && u[2] == "QUADRUPLETS:") {
          P(hf[j]);
          total_quads = std::stol(u[3]);
          quads_line = (int)j;
          brk = true;
          break;
        }
      }
    }
    if (brk) break;
  }
  std::string pot = tok(hf[7])[1];
  P("");
  P("PAIR " + pot + " PARAMS \n");
  int snum2 = 0;
  {
    auto t = tok(hf[7]);
    if (t.size() >= 4) snum2 = std::stoi(t[2]);
  }
  int ntypes = std::stoi(tok(hf[9])[2]);
  int pairs_line = 9 + 2 + ntypes + 2;
  int total_pairs = std::stoi(tok(hf[pairs_line])[2]);
  int effective_pair = 0, counted_coul = 0;
  const bool fitcoul = tok(hf[1]).size() > 1 && tok(hf[1])[1] == "true";
  if (fitcoul && !excl2.empty()) throw Error(CG_ERR_ARG, "ERROR: cannot use hierarchical learning with charge fitting.");
  for (int i = 0; i < total_pairs; ++i) {
    auto t = tok(hf[pairs_line + 2 + i + 1]);
    P("PAIRTYPE PARAMS: " + std::to_string(i) + " " + t[1] + " " + t[2] + "\n");
    if (std::find(excl2.begin(), excl2.end(), i) != excl2.end()) {
      // excluded pair type: zeros, no entry of x
      for (int j = 0; j < snum2; ++j) {
        std::snprintf(buf, sizeof buf, "%3d %21.13e", j, 0.0);
        P(buf);
      }
    } else {
      for (int j = 0; j < snum2; ++j) {
        std::snprintf(buf, sizeof buf, "%3d %21.13e", j, x[effective_pair * snum2 + j]);
        P(buf);
      }
      ++effective_pair;
    }
    if (fitcoul) {
      // q_A x q_B: the pair's Ewald column, after every short-range column
      std::snprintf(buf, sizeof buf, "q_%s x q_%s %21.13e", t[1].c_str(), t[2].c_str(),
                    x[total_pairs * snum2 + s.n3b + s.n4b + i]);
      P(buf);
      ++counted_coul;
    }
    P(" ");
  }
  int counted_trip = 0, counted_quad = 0;
  if (total_trips > 0) {
    P("TRIPLET " + pot + " PARAMS \n");
    int add = 0, par = 0;
    for (long t = 0; t < total_trips; ++t) {
      P("TRIPLETTYPE PARAMS:");
      P("  " + rstrip(hf[trips_line + 2 + add]));
      auto p1 = tok(hf[trips_line + 3 + add]);
      if (p1[4] == "EXCLUDED:") {
        P("   PAIRS: " + p1[1] + " " + p1[2] + " " + p1[3] + " EXCLUDED:");
        add += 1;
      } else {
        std::string uniq = p1[4], totl = p1[6];
        P("   PAIRS: " + p1[1] + " " + p1[2] + " " + p1[3] + " UNIQUE: " + uniq + " TOTAL: " + totl);
        P("     index  |  powers  |  equiv index  |  param index  |       parameter       ");
        P("   ----------------------------------------------------------------------------");
        add += 3;
        for (int i = 0; i < std::stoi(totl); ++i) {
          add += 1;
          const std::string &L = hf[trips_line + 2 + add];
          auto ls = tok(L);
          std::snprintf(buf, sizeof buf, "%s %21.13e", L.c_str(), x[effective_pair * snum2 + par + std::stoi(ls[5])]);
          P(buf);
        }
        par += std::stoi(uniq);
        counted_trip += std::stoi(uniq);
      }
      P("");
      add += 2;
    }
  }
  if (total_quads > 0) {
    P("QUADRUPLET " + pot + " PARAMS \n");
    int add = 0, par = 0;
    for (long t = 0; t < total_quads; ++t) {
      P("QUADRUPLETYPE PARAMS: ");
      P("  " + rstrip(hf[quads_line + 2 + add]));
      auto p1 = tok(hf[quads_line + 3 + add]);
      std::string v = p1[1] + " " + p1[2] + " " + p1[3] + " " + p1[4] + " " + p1[5] + " " + p1[6];
      if (p1[7] == "EXCLUDED:") {
        P("   PAIRS: " + v + " EXCLUDED: ");
        add += 1;
      } else {
        std::string uniq = p1[7], totl = p1[9];
        P("   PAIRS: " + v + " UNIQUE: " + uniq + " TOTAL: " + totl);
        P("     index  |  powers  |  equiv index  |  param index  |       parameter       ");
        P("   ----------------------------------------------------------------------------");
        add += 3;
        for (int i = 0; i < std::stoi(totl); ++i) {
          add += 1;
          const std::string &L = hf[quads_line + 2 + add];
          auto ls = tok(L);
          std::snprintf(buf, sizeof buf, "%s %21.13e", L.c_str(),
                        x[effective_pair * snum2 + counted_trip + par + std::stoi(ls[8])]);
          P(buf);
        }
        par += std::stoi(uniq);
        counted_quad += std::stoi(uniq);
      }
      P("");
      add += 2;
    }
  }
  P("");
  for (auto &l : split_lines(map_text(s))) P(l);
  P("");
  int total = effective_pair * snum2 + counted_trip + counted_quad + counted_coul;
  const int n_offsets = ntypes - (int)excl1.size();
  if (total != nx && nx != total + n_offsets) throw Error(CG_ERR_ARG, "Error in counting parameters");
  if (nx == total + n_offsets) {
    P("NO ENERGY OFFSETS:  " + std::to_string(ntypes));
    int eff = 0;
    for (int i = 0; i < ntypes; ++i) {
      const bool ex = std::find(excl1.begin(), excl1.end(), i) != excl1.end();
      std::snprintf(buf, sizeof buf, "ENERGY OFFSET %d %21.13e", i + 1, ex ? 0.0 : x[total + eff]);
      if (!ex) ++eff;
      P(buf);
    }
  }
  P("ENDFILE");
  return o.str();
}

}  // namespace cg

extern "C" cg_status cg_write_params(const cg_setup *cs, const char *outdir, const double *x, int ncols,
                                     int64_t nrows, const cg_fit_result *res, const double *y, const double *b) {
  try {
    std::string d = outdir ? outdir : ".";
    std::time_t now = std::time(nullptr);
    char date[32];
    std::strftime(date, sizeof date, "%Y-%m-%d", std::localtime(&now));
    std::ofstream p(d + "/params.txt");
    if (!p) throw Error(CG_ERR_INPUT, "cannot write params.txt");
    p << params_text(*cs->s, x, ncols, nrows, *res, date);
    if (y) {
      std::FILE *f = std::fopen((d + "/force.txt").c_str(), "w");
      if (!f) throw Error(CG_ERR_INPUT, "cannot write force.txt");
      for (int64_t i = 0; i < nrows; ++i) std::fprintf(f, "%13.6e\n", y[i]);
      std::fclose(f);
    }
    (void)b;
    return CG_OK;
  } catch (const Error &e) { return fail(e); } catch (const std::exception &e) { return fail(e); }
}
