// params.txt reader: the force-field file that chimes_lsq.py writes and
// chimes_md reads (src/chimes_md.C:1474-2179; cluster blocks in
// src/Cluster.C:511-603,705-747).
//
// The model is rebuilt with the fm_setup.in parser from an equivalent set of
// keywords (species, pair table with cutoffs and transform, Chebyshev orders
// and range, FCUT type, special cutoffs, excluded cluster types), so the
// column layout is the one chimes_lsq would use for the same model.  Each
// coefficient block is then checked against the rebuilt tables (pair names,
// cluster atoms, unique/total counts, every allowed power set and its param
// index) before the coefficients are placed in A column order.
#include <cmath>
#include <fstream>
#include <sstream>
#include <map>
#include <string>
#include <vector>

#include "setup.hpp"

namespace cg {

namespace {

std::vector<std::string> tokenize(const std::string &line) { return split_tokens(line); }
int to_int(const std::string &v) { return parse_int(v); }
double to_double(const std::string &v) { return parse_double(v); }

bool starts(const std::string &l, const char *p) { return l.compare(0, std::string(p).size(), p) == 0; }

std::string after_colon(const std::string &l) {
  const size_t c = l.find(':');
  return c == std::string::npos ? std::string() : l.substr(c + 1);
}

struct ParamsText {
  std::vector<std::string> lines;
  size_t find(const char *prefix, size_t from = 0) const {
    for (size_t i = from; i < lines.size(); ++i)
      if (starts(lines[i], prefix)) return i;
    return std::string::npos;
  }
};

[[noreturn]] void bad(const std::string &path, const std::string &what) {
  throw Error(CG_ERR_INPUT, "params file " + path + ": " + what);
}

}  // namespace

Setup *read_params(const std::string &path) {
  std::ifstream in(path);
  if (!in) throw Error(CG_ERR_INPUT, "ERROR: Cannot open parameter file: " + path);
  ParamsText P;
  std::string line;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    P.lines.push_back(line);
  }
  auto flag = [&](const char *key) {
    const size_t i = P.find(key);
    if (i == std::string::npos) return false;
    auto t = tokenize(after_colon(P.lines[i]));
    return !t.empty() && (t[0] == "true" || t[0] == "TRUE" || t[0] == "True");
  };
  // FITCOUL: charge products q_A x q_B after each pair's 2-body block
  // (chimes_md.C:2009-2016); fixed charges (USECOUL without FITCOUL) are not
  // supported
  const bool fitcoul = flag("FITCOUL:");
  if (flag("USECOUL:") && !fitcoul)
    throw Error(CG_ERR_UNSUPPORTED, "fixed Coulomb charges in " + path + " are not supported");
  if (flag("USEPOVR:")) throw Error(CG_ERR_UNSUPPORTED, "over-coordination terms in " + path + " are not supported");

  // ---- header
  size_t i = P.find("PAIRTYP:");
  if (i == std::string::npos) bad(path, "no PAIRTYP line");
  auto pt = tokenize(after_colon(P.lines[i]));
  if (pt.empty() || pt[0] != "CHEBYSHEV") throw Error(CG_ERR_UNSUPPORTED, "only CHEBYSHEV force fields are supported");
  const int o2 = pt.size() > 1 ? to_int(pt[1]) : 0;
  const int o3 = pt.size() > 2 ? to_int(pt[2]) : 0;
  const int o4 = pt.size() > 3 ? to_int(pt[3]) : 0;
  const std::string lo = pt.size() > 4 ? pt[4] : "-1", hi = pt.size() > 5 ? pt[5] : "1";

  i = P.find("ATOM TYPES:");
  if (i == std::string::npos) bad(path, "no ATOM TYPES line");
  const int nt = to_int(tokenize(after_colon(P.lines[i]))[0]);
  i = P.find("# TYPEIDX #", i);
  if (i == std::string::npos) bad(path, "no TYPEIDX table");
  std::vector<std::vector<std::string>> types;
  for (size_t k = i + 1; (int)types.size() < nt && k < P.lines.size(); ++k) {
    auto t = tokenize(P.lines[k]);
    if (t.empty()) continue;
    if (t.size() < 4) bad(path, "bad atom type line: " + P.lines[k]);
    types.push_back(t);
  }
  if ((int)types.size() != nt) bad(path, "truncated TYPEIDX table");

  i = P.find("ATOM PAIRS:");
  if (i == std::string::npos) bad(path, "no ATOM PAIRS line");
  const int np = to_int(tokenize(after_colon(P.lines[i]))[0]);
  i = P.find("# PAIRIDX #", i);
  if (i == std::string::npos) bad(path, "no PAIRIDX table");
  std::vector<std::vector<std::string>> pairs;
  for (size_t k = i + 1; (int)pairs.size() < np && k < P.lines.size(); ++k) {
    auto t = tokenize(P.lines[k]);
    if (t.empty()) continue;
    if (t.size() < 6) bad(path, "bad atom pair line: " + P.lines[k]);
    // older files carry an S_DELTA column before CHBDIST: a number where the
    // transform name is expected is skipped (chimes_md.C:1846-1858)
    char *end = nullptr;
    std::strtod(t[5].c_str(), &end);
    if (end && *end == '\0' && !t[5].empty()) t.erase(t.begin() + 5);
    if (t.size() < 6 || (t[5] == "MORSE" && t.size() < 7)) bad(path, "bad atom pair line: " + P.lines[k]);
    if (t.size() < 7) t.push_back("0");
    pairs.push_back(t);
  }
  if ((int)pairs.size() != np) bad(path, "truncated PAIRIDX table");
  for (auto &p : pairs)
    if (p[5] != pairs[0][5])
      throw Error(CG_ERR_UNSUPPORTED, "pair types with different distance transforms are not supported");

  std::string fcut = "CUBIC";
  if ((i = P.find("FCUT TYPE:")) != std::string::npos) {
    auto t = tokenize(after_colon(P.lines[i]));
    fcut.clear();
    for (auto &w : t) fcut += (fcut.empty() ? "" : " ") + w;
  }

  // excluded cluster types: "INDEX: k ATOMS: a b c" followed by "PAIRS: ... EXCLUDED:"
  std::vector<std::vector<std::string>> excl3, excl4;
  for (size_t k = 0; k + 1 < P.lines.size(); ++k) {
    const bool t3 = starts(P.lines[k], "TRIPLETTYPE PARAMS:"), t4 = starts(P.lines[k], "QUADRUPLETYPE PARAMS:");
    if (!t3 && !t4) continue;
    auto idx = tokenize(P.lines[k + 1]);
    auto pl = k + 2 < P.lines.size() ? P.lines[k + 2] : std::string();
    if (pl.find("EXCLUDED") == std::string::npos) continue;
    std::vector<std::string> atoms;
    for (size_t q = 0; q < idx.size(); ++q)
      if (idx[q] == "ATOMS:")
        for (size_t r = q + 1; r < idx.size(); ++r) atoms.push_back(idx[r]);
    (t3 ? excl3 : excl4).push_back(atoms);
  }
  double pen_dist = 0.01, pen_scale = 1.0e4;
  if ((i = P.find("PAIR CHEBYSHEV PENALTY DIST:")) != std::string::npos)
    pen_dist = to_double(tokenize(after_colon(P.lines[i]))[0]);
  if ((i = P.find("PAIR CHEBYSHEV PENALTY SCALING:")) != std::string::npos)
    pen_scale = to_double(tokenize(after_colon(P.lines[i]))[0]);
  std::vector<double> offsets;
  if ((i = P.find("NO ENERGY OFFSETS:")) != std::string::npos) {
    const int no = to_int(tokenize(after_colon(P.lines[i]))[0]);
    for (int q = 0; q < no; ++q) {
      const size_t k = P.find("ENERGY OFFSET ", i + 1);
      if (k == std::string::npos) bad(path, "missing ENERGY OFFSET line");
      auto t = tokenize(P.lines[k]);
      if (t.size() < 4 || to_int(t[2]) != q + 1) bad(path, "bad ENERGY OFFSET line: " + P.lines[k]);
      offsets.push_back(to_double(t[3]));
      i = k;
    }
  }

  // ---- equivalent fm_setup.in keywords
  std::ostringstream fm;
  fm << "# TRJFILE #\n\tnone.xyzf\n# NFRAMES #\n\t1\n# NLAYERS #\n\t1\n# FITCOUL #\n\t" << (fitcoul ? "true" : "false") << "\n"
     << "# FITSTRS #\n\tfalse\n# FITENER #\n\t" << (offsets.empty() ? "false" : "true") << "\n"
     << "# PAIRTYP #\n\tCHEBYSHEV " << o2 << " " << o3 << " " << o4 << " " << lo << " " << hi << "\n"
     << "# CHBTYPE #\n\t" << pairs[0][5] << "\n"
     << "# NATMTYP #\n\t" << nt << "\n# TYPEIDX # # ATM_TYP # # ATMCHRG # # ATMMASS #\n";
  for (int q = 0; q < nt; ++q)
    fm << "\t" << q + 1 << " " << types[q][1] << " " << (fitcoul ? types[q][2] : std::string("0")) << " " << types[q][3]
       << "\n";
  fm << "# PAIRIDX # # ATM_TY1 # # ATM_TY1 # # S_MINIM # # S_MAXIM # # CHBDIST # # MORSE_LAMBDA #\n";
  for (auto &p : pairs)
    fm << "\t" << p[0] << " " << p[1] << " " << p[2] << " " << p[3] << " " << p[4] << " " << p[5] << " " << p[6]
       << "\n";
  for (int fam = 3; fam <= 4; ++fam) {
    const auto &ex = fam == 3 ? excl3 : excl4;
    if (ex.empty()) continue;
    fm << "EXCLUDE " << fam << "B INTERACTION: " << ex.size() << "\n";
    for (auto &e : ex) {
      for (auto &a : e) fm << a << " ";
      fm << "\n";
    }
  }
  for (const char *key : {"SPECIAL 3B S_MINIM:", "SPECIAL 3B S_MAXIM:", "SPECIAL 4B S_MINIM:", "SPECIAL 4B S_MAXIM:"}) {
    if ((i = P.find(key)) == std::string::npos) continue;
    fm << P.lines[i] << "\n";
    auto t = tokenize(P.lines[i]);
    if (t.size() >= 5 && t[3] == "SPECIFIC")
      for (int q = 1; q <= to_int(t[4]); ++q) fm << P.lines.at(i + q) << "\n";
  }
  fm << "# FCUTTYP #\n\t" << fcut << "\n# ENDFILE #\n";
  std::istringstream fin(fm.str());
  Setup *s = read_setup_from(fin, path);
  try {
    s->penalty_dist = pen_dist;
    s->penalty_scale = pen_scale;
    s->coeffs.assign(s->model.ncols, 0.0);
    std::vector<char> set(s->model.ncols, 0);
    auto put = [&](int col, double v, const std::string &where) {
      if (col < 0 || col >= s->model.ncols) bad(path, "column out of range in " + where);
      if (set[col] && s->coeffs[col] != v) bad(path, "inconsistent parameters for one column in " + where);
      s->coeffs[col] = v;
      set[col] = 1;
    };
    // 2-body blocks: "PAIRTYPE PARAMS: p A B", then "power coefficient" lines
    const cg_model &m = s->model;
    size_t k = 0;
    for (int p = 0; p < m.npairs; ++p) {
      k = P.find("PAIRTYPE PARAMS:", k);
      if (k == std::string::npos) bad(path, "missing PAIRTYPE PARAMS block " + std::to_string(p));
      auto h = tokenize(after_colon(P.lines[k]));
      if (h.size() < 3 || to_int(h[0]) != p) bad(path, "bad PAIRTYPE PARAMS header: " + P.lines[k]);
      const std::string nm = h[1] + h[2];
      if (s->pairs[p].proper != nm && s->pairs[p].a2 + s->pairs[p].a1 != nm)
        bad(path, "pair " + std::to_string(p) + " is " + nm + ", expected " + s->pairs[p].proper);
      int got = 0;
      for (size_t q = k + 1; q < P.lines.size() && got < m.order2b; ++q) {
        auto t = tokenize(P.lines[q]);
        if (t.empty()) continue;
        if (t.size() != 2) break;
        const int pw = to_int(t[0]);
        if (pw < 0 || pw >= m.order2b) bad(path, "bad 2-body power in " + P.lines[k]);
        put(p * m.order2b + pw, to_double(t[1]), P.lines[k]);
        ++got;
      }
      if (got != m.order2b) bad(path, "truncated PAIRTYPE PARAMS block " + std::to_string(p));
      if (fitcoul) {
        // "q_A x q_B value", the pair's Ewald column
        size_t q = k + 1;
        while (q < P.lines.size() && !starts(P.lines[q], "q_")) ++q;
        auto t = q < P.lines.size() ? tokenize(P.lines[q]) : std::vector<std::string>();
        if (t.size() != 4 || t[1] != "x") bad(path, "missing charge product after PAIRTYPE PARAMS block " + std::to_string(p));
        put(m.ncols_short + p, to_double(t[3]), P.lines[k]);
      }
      ++k;
    }
    // cluster blocks
    for (int fam = 3; fam <= 4; ++fam) {
      const cg_cluster_set &cs = fam == 3 ? m.c3 : m.c4;
      if (!cs.nbody) continue;
      const char *head = fam == 3 ? "TRIPLETTYPE PARAMS:" : "QUADRUPLETYPE PARAMS:";
      k = 0;
      for (int t = 0; t < cs.ntypes; ++t) {
        k = P.find(head, k);
        if (k == std::string::npos) bad(path, std::string("missing ") + head + " block " + std::to_string(t));
        auto idl = tokenize(P.lines.at(k + 1));
        if (idl.size() < 4 || idl[0] != "INDEX:" || to_int(idl[1]) != t) bad(path, "bad cluster index line: " + P.lines[k + 1]);
        const std::string pl = P.lines.at(k + 2);
        if (pl.find("EXCLUDED") != std::string::npos) {
          if (!cs.excluded[t]) bad(path, "cluster type " + std::to_string(t) + " excluded in the file only");
          ++k;
          continue;
        }
        if (cs.excluded[t]) bad(path, "cluster type " + std::to_string(t) + " is not marked excluded");
        auto pt2 = tokenize(pl);
        int uniq = -1, total = -1;
        for (size_t q = 0; q + 1 < pt2.size(); ++q) {
          if (pt2[q] == "UNIQUE:") uniq = to_int(pt2[q + 1]);
          if (pt2[q] == "TOTAL:") total = to_int(pt2[q + 1]);
        }
        const int b = cs.perm_offset[t], e = cs.perm_offset[t + 1];
        // UNIQUE: -1 marks a trimmed file (published force fields): the block
        // lists only the power sets that carry a parameter, with their powers;
        // chimes_md sums over exactly those sets (CLUSTER::read_ff_params,
        // src/Cluster.C:511-603).  Each listed set is found in the rebuilt
        // table by its powers; a column must be listed with every one of its
        // permutations, as a fit produces them.
        const bool trimmed = uniq == -1;
        if (trimmed) {
          if (total < 0 || total > e - b)
            bad(path, "cluster type " + std::to_string(t) + ": TOTAL exceeds the rebuilt tables");
        } else if (uniq != cs.n_unique[t] || total != e - b) {
          bad(path, "cluster type " + std::to_string(t) + ": UNIQUE/TOTAL differ from the rebuilt tables");
        }
        std::map<std::vector<int>, int> by_powers;
        std::vector<int> listed;
        if (trimmed) {
          for (int i = b; i < e; ++i)
            by_powers[std::vector<int>(cs.perm_powers + (size_t)i * cs.npairs,
                                       cs.perm_powers + (size_t)(i + 1) * cs.npairs)] = i;
          listed.assign(cs.n_unique[t], 0);
        }
        size_t q = k + 3;
        while (q < P.lines.size() && P.lines[q].find("----") == std::string::npos) ++q;
        int row = 0;
        for (++q; q < P.lines.size() && row < total; ++q) {
          auto r = tokenize(P.lines[q]);
          if (r.empty()) continue;
          if ((int)r.size() != cs.npairs + 4) bad(path, "bad cluster parameter row: " + P.lines[q]);
          if (to_int(r[0]) != row) bad(path, "cluster parameter rows out of order: " + P.lines[q]);
          if (trimmed) {
            std::vector<int> pw(cs.npairs);
            for (int c = 0; c < cs.npairs; ++c) pw[c] = to_int(r[1 + c]);
            auto it = by_powers.find(pw);
            if (it == by_powers.end()) bad(path, "power set not in the rebuilt tables: " + P.lines[q]);
            const int col = cs.perm_col[it->second];
            ++listed[col];
            put(cs.col_offset[t] + col, to_double(r[cs.npairs + 3]), P.lines[k + 1]);
            ++row;
            continue;
          }
          for (int c = 0; c < cs.npairs; ++c)
            if (to_int(r[1 + c]) != cs.perm_powers[(size_t)(b + row) * cs.npairs + c])
              bad(path, "power set differs from the rebuilt tables: " + P.lines[q]);
          const int pidx = to_int(r[cs.npairs + 2]);
          if (pidx != cs.perm_col[b + row]) bad(path, "param index differs from the rebuilt tables: " + P.lines[q]);
          put(cs.col_offset[t] + pidx, to_double(r[cs.npairs + 3]), P.lines[k + 1]);
          ++row;
        }
        if (row != total) bad(path, "truncated cluster parameter block " + std::to_string(t));
        if (trimmed) {
          std::vector<int> nperm(cs.n_unique[t], 0);
          for (int i = b; i < e; ++i) ++nperm[cs.perm_col[i]];
          for (int col = 0; col < cs.n_unique[t]; ++col)
            if (listed[col] && listed[col] != nperm[col])
              bad(path, "cluster type " + std::to_string(t) + " lists " + std::to_string(listed[col]) + " of the " +
                            std::to_string(nperm[col]) + " permutations of one power set");
        }
        k = q;
      }
    }
    // energy offsets (one per atom type, after the short-range columns)
    if (!offsets.empty()) {
      if ((int)offsets.size() != m.n_offset_cols) bad(path, "energy offsets do not match the atom types");
      for (int q = 0; q < m.n_offset_cols; ++q) s->coeffs[m.ncols_short + q] = offsets[q];
    }
  } catch (...) {
    delete s;
    throw;
  }
  return s;
}

}  // namespace cg
