// fm_setup.in parsing and the model tables of the ChIMES least-squares path.
//
// Keyword semantics follow INPUT::PARSE_INFILE_LSQ (src/input.C:172-1338);
// the pair tables follow PARSE_TOPOLOGY_TYPEIDX / PAIRIDX (src/input.C:918-1146)
// and build_int_pair_map (src/functions.C:1055-1095); the cluster tables follow
// CLUSTER_LIST::build_all / build_pairs_loop / CLUSTER::build
// (src/Cluster.C:27-366, 771-954, 1211-1484, 1746-1878).
#include "setup.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>

namespace cg {
namespace {

// Strip '!' and '##' comments (src/util.C:54-67) and split on whitespace.
std::vector<std::string> tokenize(std::string line) {
  size_t p = line.find('!');
  if (p != std::string::npos) line.erase(p);
  p = line.find("##");
  if (p != std::string::npos) line.erase(p);
  std::vector<std::string> out;
  std::istringstream is(line);
  std::string t;
  while (is >> t) out.push_back(t);
  return out;
}

struct Lines {
  std::vector<std::vector<std::string>> v;
  const std::string &at(size_t i, size_t j) const {
    if (i >= v.size() || j >= v[i].size())
      throw Error(CG_ERR_INPUT, "fm_setup.in: missing value after line " + std::to_string(i + 1));
    return v[i][j];
  }
  size_t width(size_t i) const { return i < v.size() ? v[i].size() : 0; }
  bool keyword(size_t i, const char *key) const {  // "# KEY #"
    const auto &l = v[i];
    return l.size() >= 3 && l[0] == "#" && l[2] == "#" && l[1] == key;
  }
  std::string head3(size_t i) const {
    const auto &l = v[i];
    if (l.size() < 4) return "";
    return l[0] + " " + l[1] + " " + l[2];
  }
};

const char *bool_str(bool b) { return b ? "true" : "false"; }

bool to_bool(const std::string &s) {
  if (s.empty()) throw Error(CG_ERR_INPUT, "expected true/false");
  char c = s[0];
  if (c == 't' || c == 'T') return true;
  if (c == 'f' || c == 'F') return false;
  throw Error(CG_ERR_INPUT, "String found where true/false expected: " + s);
}
int to_int(const std::string &s) {
  size_t pos = 0;
  int v = 0;
  try { v = std::stoi(s, &pos); } catch (...) {
    throw Error(CG_ERR_INPUT, "String found where integer expected: " + s);
  }
  return v;
}
double to_double(const std::string &s) {
  try { return std::stod(s); } catch (...) {
    throw Error(CG_ERR_INPUT, "String found where floating point expected: " + s);
  }
}

int transform_of(const std::string &s) {
  if (s == "MORSE") return CG_MORSE;
  if (s == "INVRSE_R") return CG_INVRSE_R;
  if (s == "DEFAULT" || s == "NONE") return CG_NONE;
  throw Error(CG_ERR_INPUT, "Bad Cheby transformation type: " + s);
}

// Cheby::set_cheby_params (src/Cheby.C:102-139).
void cheby_params(double smin, double smax, double lambda, int tr, double &xmin, double &xmax,
                  double &xdiff, double &xavg) {
  switch (tr) {
    case CG_MORSE: xmin = std::exp(-smax / lambda); xmax = std::exp(-smin / lambda); break;
    case CG_INVRSE_R: xmin = 1.0 / smax; xmax = 1.0 / smin; break;
    default: xmin = smin; xmax = smax; break;
  }
  xavg = 0.5 * (xmin + xmax);
  xdiff = 0.5 * (xmax - xmin);
}

int make_id(const std::vector<int> &idx) {  // CLUSTER_LIST::make_id_int
  int sum = 0, w = 1;
  for (int v : idx) { sum += w * (v + 1); w *= CG_MAX_ATOM_TYPES; }
  return sum;
}

// ---------------------------------------------------------- power sets
// Enumerate the atom permutations that keep identical names in their blocks
// and record each distinct permuted power vector once (CLUSTER::store_permutations,
// CLUSTER::permute_atom_indices).
void permute_blocks(ClusterType &c, size_t idx, std::vector<int> perm, const std::vector<int> &pw,
                    int unique, int equiv) {
  const int n = c.natoms;
  if ((int)idx < n) {
    size_t cnt = 1;
    for (size_t i = idx; i + 1 < (size_t)n && c.atom_names[i + 1] == c.atom_names[i]; ++i) ++cnt;
    do {
      permute_blocks(c, idx + cnt, perm, pw, unique, equiv);
    } while (std::next_permutation(perm.begin() + idx, perm.begin() + idx + cnt));
    return;
  }
  std::vector<std::vector<int>> mat(n, std::vector<int>(n, 0));
  int m = 0;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) { mat[i][j] = mat[j][i] = pw[m++]; }
  std::vector<int> out(c.npairs);
  m = 0;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) out[m++] = mat[perm[i]][perm[j]];
  auto it = c.allowed_lookup.find(out);
  if (it == c.allowed_lookup.end()) {
    c.allowed.push_back(out);
    c.power_count.push_back(1);
    c.allowed_lookup.emplace(out, (int)c.allowed.size() - 1);
    c.param_idx.push_back(unique);
    c.equiv_idx.push_back(equiv);
  } else {
    ++c.power_count[it->second];
  }
}

void store_powers(ClusterType &c, const std::vector<int> &pw) {
  if (c.allowed_lookup.count(pw)) return;
  c.unique_powers.push_back(pw);
  std::vector<int> perm(c.natoms);
  for (int i = 0; i < c.natoms; ++i) perm[i] = i;
  permute_blocks(c, 0, perm, pw, (int)c.unique_powers.size() - 1, (int)c.allowed.size());
}

void power_loop(ClusterType &c, int order, int k, std::vector<int> &pw) {
  if (k < c.npairs) {
    for (int p = 0; p < order; ++p) { pw[k] = p; power_loop(c, order, k + 1, pw); }
    return;
  }
  // every atom must carry a non-zero power (src/Cluster.C:129-154)
  std::vector<bool> rep(c.natoms, false);
  int m = 0;
  for (int i = 0; i < c.natoms; ++i)
    for (int j = i + 1; j < c.natoms; ++j)
      if (pw[m++] > 0) rep[i] = rep[j] = true;
  for (bool r : rep) if (!r) return;
  store_powers(c, pw);
}

// Canonical-slot assignment of an enumerated tuple of atom types
// (CLUSTER::map_indices_int, src/Cluster.C:1768-1878): actual atom m takes the
// first unused canonical slot of its type.
void slot_maps(const ClusterType &c, const std::vector<int> &types, std::vector<int> &pair_map,
               std::vector<int> &atom_perm) {
  const int n = c.natoms;
  std::vector<int> rev(n, -1);
  std::vector<bool> used(n, false);
  for (int m = 0; m < n; ++m)
    for (int s = 0; s < n; ++s)
      if (!used[s] && c.atom_types[s] == types[m]) { rev[m] = s; used[s] = true; break; }
  for (int m = 0; m < n; ++m)
    if (rev[m] < 0) throw Error(CG_ERR_INPUT, "cluster slot map failed");
  std::vector<std::vector<int>> pidx(n, std::vector<int>(n, -1));
  int cnt = 0;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) { pidx[i][j] = pidx[j][i] = cnt++; }
  pair_map.assign(c.npairs, 0);
  int m = 0;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) pair_map[m++] = pidx[rev[i]][rev[j]];
  atom_perm.assign(n, 0);
  for (int i = 0; i < n; ++i) atom_perm[rev[i]] = i;
}

// CLUSTER_LIST::tuplet_name (src/Cluster.C:1699-1745).
std::string tuplet_name(int natoms, bool plural, bool caps) {
  std::string t = natoms == 2 ? "pair" : natoms == 3 ? "triplet" : natoms == 4 ? "quadruplet" : std::to_string(natoms) + "-tuplet";
  if (plural) t += "s";
  if (caps) t[0] = (char)std::toupper(t[0]);
  return t;
}

bool is_excluded(const ClusterFamily &f, std::vector<std::string> names) {
  std::sort(names.begin(), names.end());
  for (const auto &e : f.exclude)
    if (e == names) return true;
  return false;
}

void build_types(Setup &s, ClusterFamily &f, int idx, std::vector<int> &tup) {
  const int n = f.natoms;
  if (idx < n) {
    for (int t = 0; t < s.ntypes; ++t) { tup[idx] = t; build_types(s, f, idx + 1, tup); }
    return;
  }
  std::vector<std::string> names(n);
  for (int i = 0; i < n; ++i) names[i] = s.type_names[tup[i]];
  std::sort(names.begin(), names.end());
  std::vector<std::string> pn, pu;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      pn.push_back(s.pairs[s.pair_map.at(names[i] + names[j])].proper);
      pu.push_back(s.pairs[s.pair_map.at(s.type_names[tup[i]] + s.type_names[tup[j]])].proper);
    }
  int map_index = -1;
  bool found = false;
  for (size_t i = 0; i < f.types.size(); ++i)
    if (f.types[i].pair_names == pn) { found = true; map_index = (int)i; break; }
  bool excl = is_excluded(f, names);
  if (!found) {
    // the reference prints the new entry before filling in its pair names
    s.echo << "\n\tMade the following " << tuplet_name(n, true, false) << ": " << f.types.size() << " ";
    for (int m = 0; m < f.npairs; ++m) s.echo << " ";
    s.echo << "\n\n";
    ClusterType c;
    c.index = (int)f.types.size();
    c.natoms = n;
    c.npairs = f.npairs;
    c.atom_names = names;
    for (auto &nm : names) c.atom_types.push_back(type_of(s, nm));
    c.pair_names = pn;
    c.excluded = excl;
    if (!excl) map_index = c.index;
    f.types.push_back(c);
  }
  std::string all;
  for (auto &p : pu) all += p;
  if (excl) map_index = -1;
  f.name_map.insert({all, map_index});
  f.name_rev.insert({map_index, all});
  int id = make_id(tup);
  f.id_map[id] = map_index;
  if (map_index >= 0) {
    slot_maps(f.types[map_index], tup, f.id_pair_map[id], f.id_atom_perm[id]);
  }
  s.echo << "\t\tAtom type idxs: ";
  for (int i = 0; i < n; ++i) s.echo << std::setw(2) << tup[i];
  s.echo << "  " << tuplet_name(n, false, true) << " name: " << std::setw(12) << all << " Unique "
         << tuplet_name(n, false, false) << " ID " << id << " index: " << std::setw(4) << map_index << "\n";
}

// CLUSTER_LIST::process_cutoff_params (src/Cluster.C:771-954).
void special_cutoffs(Setup &s, ClusterFamily &f, bool is_max) {
  const auto &in = is_max ? f.smax_input : f.smin_input;
  const char *what = is_max ? "S_MAXIM" : "S_MINIM";
  if (!in.empty()) {
    auto tk = tokenize(in[0]);
    if (tk.size() < 5) throw Error(CG_ERR_INPUT, "Not enough tokens: " + in[0]);
    if (tk[3] == "ALL") {
      double v = to_double(tk[4]);
      for (auto &c : f.types) {
        auto &vec = is_max ? c.smax : c.smin;
        for (auto &x : vec) x = v;
        (is_max ? c.special_max : c.special_min) = true;
      }
      s.echo << "\tNote: Setting all " << f.natoms << "-body " << what << " values to " << v << "\n";
    } else if (tk[3] == "SPECIFIC") {
      s.echo << "\tNote: Setting specific " << f.natoms << "-body " << what << " values: \n";
      int nspec = to_int(tk[4]);
      for (int i = 0; i < nspec; ++i) {
        auto t = tokenize(in.at(1 + i));
        if ((int)t.size() != 2 * f.npairs + 1) throw Error(CG_ERR_INPUT, "Wrong number of arguments: " + in[1 + i]);
        auto it = f.name_map.find(t[0]);
        if (it == f.name_map.end()) throw Error(CG_ERR_INPUT, "Did not find a map for " + t[0]);
        int ci = it->second;
        if (ci < 0) continue;
        auto &c = f.types[ci];
        for (int k = 0; k < f.npairs; ++k) {
          std::string pname = s.pairs[s.pair_map.at(t[1 + k])].proper;
          double v = to_double(t[1 + k + f.npairs]);
          for (int j = 0; j < f.npairs; ++j) {
            if (c.pair_names[j] == pname) (is_max ? c.smax : c.smin)[j] = v;
            (is_max ? c.special_max : c.special_min) = true;
          }
        }
      }
      for (const auto &c : f.types) {
        s.echo << "\t";
        for (const auto &nm : c.atom_names) s.echo << nm << " ";
        for (int j = 0; j < f.npairs; ++j) s.echo << c.pair_names[j] << " " << (is_max ? c.smax : c.smin)[j] << " ";
        s.echo << "\n";
      }
      s.echo << "\n";
    }
  }
  if (is_max) {
    double v = -1.0;
    for (auto &c : f.types) for (double x : c.smax) v = std::max(v, x);
    f.max_cutoff = v;
    s.echo << "     Max " << f.natoms << "-body cutoff = " << v << "\n\n";
  } else {
    double v = 1.0e10;
    for (auto &c : f.types) for (double x : c.smin) v = std::min(v, x);
    f.min_cutoff = v;
    s.echo << "     Min " << f.natoms << "-body cutoff = " << v << "\n";
  }
}

// CLUSTER_LIST::print / CLUSTER::print (src/Cluster.C:368-426,1529-1556)
// without the MD parameters: the power-set tables of every cluster type.
void print_family(Setup &s, const ClusterFamily &f) {
  s.echo << "\n\tThe following unique " << tuplet_name(f.natoms, true, false) << " of atoms have been identified:\n";
  if (!f.exclude.empty()) s.echo << "\tNote: The following types have been removed, if present: \n";
  for (const auto &e : f.exclude) {
    for (const auto &nm : e) s.echo << "\t\t" << nm;
    s.echo << "\n";
  }
  s.echo << "\n";
  for (const auto &c : f.types) {
    s.echo << "    Cluster index       : " << c.index << "\n";
    s.echo << "    Atoms in the cluster: ";
    for (const auto &nm : c.atom_names) s.echo << nm << " ";
    s.echo << "\n    Pairs in the cluster: ";
    for (const auto &pn : c.pair_names) s.echo << pn << " ";
    s.echo << "\n";
    if (c.excluded) {
      s.echo << "   Interaction is excluded by the user\n";
    } else {
      s.echo << "    Number of unique sets of powers: " << c.unique_powers.size() << " (" << c.allowed.size()
             << " total)...\n";
      s.echo << "\t\t     index  |  powers  |  equiv index  |  param index  | count \n";
      s.echo << "\t\t   ----------------------------------------------------\n";
      for (size_t j = 0; j < c.allowed.size(); ++j) {
        s.echo << "\t\t      " << std::left << std::setw(6) << j << " ";
        for (int p : c.allowed[j]) s.echo << " " << std::setw(2) << p << " ";
        s.echo << "       " << std::setw(8) << c.equiv_idx[j] << " ";
        s.echo << "       " << std::setw(8) << c.param_idx[j] << " ";
        s.echo << "       " << std::setw(8) << c.power_count[j] << " " << std::right << "\n";
      }
    }
    s.echo << "\n";
  }
}

void build_family(Setup &s, ClusterFamily &f) {
  f.npairs = f.natoms * (f.natoms - 1) / 2;
  int dim = 1;
  for (int i = 0; i < f.natoms; ++i) dim *= CG_MAX_ATOM_TYPES;
  f.id_map.assign(dim, -1);
  f.id_pair_map.assign(dim, std::vector<int>(f.npairs, -1));
  f.id_atom_perm.assign(dim, std::vector<int>(f.natoms, -1));
  std::vector<int> tup(f.natoms);
  build_types(s, f, 0, tup);
  for (auto &c : f.types) {
    if (!c.excluded) {
      std::vector<int> pw(c.npairs);
      power_loop(c, f.order, 0, pw);
    }
    // default cutoffs from the 2-body pair of the same name
    c.smin.assign(c.npairs, 0);
    c.smax.assign(c.npairs, 0);
    for (int j = 0; j < c.npairs; ++j) {
      int p = s.pair_map.at(c.pair_names[j]);
      c.smin[j] = s.pairs[p].smin;
      c.smax[j] = s.pairs[p].smax;
    }
  }
  special_cutoffs(s, f, true);
  special_cutoffs(s, f, false);
  print_family(s, f);
  for (auto &c : f.types) {
    c.xmin.assign(c.npairs, 0); c.xmax = c.xavg = c.xdiff = c.xmin;
    for (int j = 0; j < c.npairs; ++j) {
      const auto &p = s.pairs[s.pair_map.at(c.pair_names[j])];
      cheby_params(c.smin[j], c.smax[j], p.lambda, s.transform, c.xmin[j], c.xmax[j], c.xdiff[j],
                   c.xavg[j]);
    }
  }
}

void parse_fcut(Setup &s) {
  auto tk = tokenize(s.fcut_line);
  if (tk.empty()) throw Error(CG_ERR_INPUT, "empty FCUTTYP");
  if (tk[0] == "CUBIC") s.fcut_type = CG_FCUT_CUBIC;
  else if (tk[0] == "TERSOFF") {
    s.fcut_type = CG_FCUT_TERSOFF;
    if (tk.size() < 2) throw Error(CG_ERR_INPUT, "TERSOFF needs an offset");
    s.fcut_offset = to_double(tk[1]);
  } else throw Error(CG_ERR_INPUT, "unknown cutoff type " + tk[0]);
}

}  // namespace

// Token helpers shared with the params.txt reader.
std::vector<std::string> split_tokens(const std::string &line) { return tokenize(line); }
int parse_int(const std::string &v) { return to_int(v); }
double parse_double(const std::string &v) { return to_double(v); }

int type_of(const Setup &s, const std::string &name) {
  for (int i = 0; i < s.ntypes; ++i)
    if (s.type_names[i] == name) return i;
  return -1;
}

Setup *read_setup(const std::string &path) {
  std::ifstream in(path);
  if (!in) throw Error(CG_ERR_INPUT, "ERROR: Cannot open input file: " + path);
  return read_setup_from(in, path);
}

Setup *read_setup_from(std::istream &in, const std::string &path) {
  auto s = new Setup;
  try {
    s->path = path;
    size_t slash = path.find_last_of('/');
    s->dir = slash == std::string::npos ? "." : path.substr(0, slash);
    Lines L;
    std::string line;
    int kick = 0;
    bool ended = false;
    while (std::getline(in, line)) {
      auto tk = tokenize(line);
      if (tk.empty()) continue;
      L.v.push_back(tk);
      if (line.find("# ENDFILE #") != std::string::npos) { ended = true; break; }
      if (++kick == 1000)
        throw Error(CG_ERR_INPUT, "Read 1000 lines and have not found \"# ENDFILE # \" in file: " + path);
    }
    if (!ended) throw Error(CG_ERR_INPUT, "no \"# ENDFILE #\" in file: " + path);
    const size_t N = L.v.size();
    auto find_kw = [&](const char *k) -> long {
      for (size_t i = 0; i < N; ++i) if (L.keyword(i, k)) return (long)i;
      return -1;
    };
    // TRJFILE (src/input.C:292-428)
    for (size_t i = 0; i < N; ++i) {
      if (L.width(i) > 1 && L.v[i][1] == "TRJFILE") {
        if (L.width(i + 1) == 1) {
          s->infiles.push_back(L.at(i + 1, 0));
          s->force_flags.push_back(""); s->stress_flags.push_back(""); s->energy_flags.push_back("");
          s->echo << "\t# TRJFILE #: " << s->infiles[0] << "\n";
        } else if (L.width(i + 1) == 2) {
          s->multi_file = L.at(i + 1, 1);
          s->echo << "\t# TRJFILE #: MULTI " << s->multi_file << "\n";
          std::ifstream mf(s->dir + "/" + L.at(i + 1, 1));
          if (!mf) throw Error(CG_ERR_INPUT, "ERROR: Cannot open MULTI file: " + L.at(i + 1, 1));
          std::getline(mf, line);
          auto t = tokenize(line);
          if (t.size() != 1) throw Error(CG_ERR_INPUT, "ERROR: Expected to read <nfiles>, got: " + line);
          s->echo << "\t\t...Will read from the following files: \n";
          int nf = to_int(t[0]);
          for (int j = 0; j < nf; ++j) {
            std::getline(mf, line);
            t = tokenize(line);
            if (t.size() != 2 && t.size() != 3 && t.size() != 5 && t.size() != 6)
              throw Error(CG_ERR_INPUT, "ERROR: bad MULTI line: " + line);
            s->infile_frames.push_back(to_int(t[0]));
            s->infiles.push_back(t[1]);
            if (t.size() >= 5) {
              auto f = [](const std::string &x) { return x == "None" ? std::string() : x; };
              s->force_flags.push_back(f(t[t.size() - 3]));
              s->stress_flags.push_back(f(t[t.size() - 2]));
              s->energy_flags.push_back(f(t[t.size() - 1]));
            } else {
              s->force_flags.push_back(""); s->stress_flags.push_back(""); s->energy_flags.push_back("");
            }
            s->echo << "\t\t-   " << s->infiles[j] << " with " << s->infile_frames[j] << " frames.\n";
          }
          s->echo << "Removing the following files with zero frames:\n";
          for (int j = (int)s->infile_frames.size() - 1; j >= 0; --j)
            if (s->infile_frames[j] < 1) {
              s->echo << "\t\t-   " << s->infiles[j] << " with " << s->infile_frames[j] << " frames.\n";
              s->infile_frames.erase(s->infile_frames.begin() + j);
              s->infiles.erase(s->infiles.begin() + j);
              s->force_flags.erase(s->force_flags.begin() + j);
              s->stress_flags.erase(s->stress_flags.begin() + j);
              s->energy_flags.erase(s->energy_flags.begin() + j);
            }
        } else throw Error(CG_ERR_INPUT, "ERROR: Unrecognized TRAJFILE option");
        break;
      }
    }
    long k;
    if ((k = find_kw("WRAPTRJ")) >= 0 && k + 1 < (long)N && L.width(k + 1) == 1) {
      s->wrap = to_bool(L.at(k + 1, 0));
      s->echo << "\t# WRAPTRJ #: " << bool_str(s->wrap) << "\n";
    }
    if ((k = find_kw("SPLITFI")) >= 0) {
      s->split_files = to_bool(L.at(k + 1, 0));
      s->echo << "\t# SPLITFI #: " << bool_str(s->split_files) << "\n";
    }
    if ((k = find_kw("HIERARC")) >= 0) {
      s->hierarchical = to_bool(L.at(k + 1, 0));
      s->echo << "\t# HIERARC #: " << bool_str(s->hierarchical) << "\n";
    }
    if ((k = find_kw("NFRAMES")) >= 0) {
      s->nframes = to_int(L.at(k + 1, 0));
      s->echo << "\t# NFRAMES #: " << s->nframes << "\n";
    }
    if ((k = find_kw("NLAYERS")) >= 0) {
      s->nlayers = to_int(L.at(k + 1, 0));
      s->echo << "\t# NLAYERS #: " << s->nlayers << "\n";
      if (L.width(k + 1) == 3) throw Error(CG_ERR_UNSUPPORTED, "NLAYERS REPLICATE is not supported");
    }
    if ((k = find_kw("FITCOUL")) >= 0) {
      s->fitcoul = to_bool(L.at(k + 1, 0));
      s->echo << "\t# FITCOUL #: " << bool_str(s->fitcoul) << "\n";
    }
    if ((k = find_kw("FITSTRS")) >= 0) {
      const std::string &v = L.at(k + 1, 0);
      if (v == "first" || v == "First" || v == "FIRST") {
        s->fit_stress = CG_STRESS_DIAG; s->nstress = to_int(L.at(k + 1, 1));
      } else if (v == "all" || v == "All" || v == "ALL" || v == "A" || v == "a") {
        s->fit_stress = CG_STRESS_ALL;
      } else if (v == "firstall" || v == "FirstAll" || v == "FIRSTALL") {
        s->fit_stress = CG_STRESS_ALL; s->nstress = to_int(L.at(k + 1, 1));
      } else s->fit_stress = to_bool(v) ? CG_STRESS_DIAG : CG_STRESS_NONE;
      s->echo << "\t# FITSTRS #: ";
      if (s->fit_stress == CG_STRESS_ALL) s->echo << "true ...will fit to all tensor components\n";
      else if (s->nstress > 0)
        s->echo << bool_str(s->fit_stress != CG_STRESS_NONE) << " ...will only fit tensors for first " << s->nstress
                << " frames.\n";
      else s->echo << bool_str(s->fit_stress != CG_STRESS_NONE) << "\n";
    }
    for (size_t i = 0; i < N; ++i)   // FITENER: last occurrence wins (no break in reference)
      if (L.keyword(i, "FITENER")) {
        const std::string &v = L.at(i + 1, 0);
        if (v == "first" || v == "First" || v == "FIRST") { s->fit_ener = true; s->nener = to_int(L.at(i + 1, 1)); }
        else s->fit_ener = to_bool(v);
        s->echo << "\t# FITENER #: " << bool_str(s->fit_ener) << "\n";
        if (s->nener >= 0) s->echo << "    \t\t\t ...will fit energies for first " << s->nener << " frames.\n";
      }
    if ((k = find_kw("PAIRTYP")) >= 0) {
      s->ff_type = L.at(k + 1, 0);
      if (s->ff_type != "CHEBYSHEV") throw Error(CG_ERR_UNSUPPORTED, "ERROR: Unrecognized pair type " + s->ff_type);
      s->o2 = to_int(L.at(k + 1, 1));
      s->o3 = L.width(k + 1) >= 3 ? to_int(L.at(k + 1, 2)) : 0;
      s->o4 = L.width(k + 1) >= 4 ? to_int(L.at(k + 1, 3)) : 0;
      s->cheby_low = L.width(k + 1) > 4 ? to_int(L.at(k + 1, 4)) : -1;
      s->cheby_high = L.width(k + 1) > 5 ? to_int(L.at(k + 1, 5)) : 1;
      if (s->cheby_low < -1 || s->cheby_low > 1 || s->cheby_high < -1 || s->cheby_high > 1 ||
          s->cheby_high < s->cheby_low)
        throw Error(CG_ERR_INPUT, "ERROR: bad Chebyshev range");
      s->echo << "\t# PAIRTYP #: " << s->ff_type
              << " ....NOTE: Forces reported in units of kcal/(mol.A), potential energy in kcal/mol.\n"
              << "\t             Will use 2-body order: " << s->o2 << "\n"
              << "\t             Will use 3-body order: " << s->o3 << "\n"
              << "\t             Will use 4-body order: " << s->o4 << "\n"
              << "\t             Will transform Chebyshev pair distances to range " << s->cheby_low << " to "
              << s->cheby_high << "\n";
    } else throw Error(CG_ERR_INPUT, "missing # PAIRTYP #");
    for (size_t i = 0; i < N; ++i)
      if (L.keyword(i, "CHBTYPE")) {
        s->transform = transform_of(L.at(i + 1, 0));
        s->echo << "\t# CHBTYPE #: " << L.at(i + 1, 0) << "\n";
      }
    for (size_t i = 0; i < N; ++i)
      if (L.keyword(i, "CHEBYFIX")) {
        const std::string &v = L.at(i + 1, 0);
        if (v == "ZERO_DERIV") s->fix_type = CG_FIX_ZERO_DERIV;
        else if (v == "CONSTANT_DERIV") s->fix_type = CG_FIX_CONSTANT_DERIV;
        else if (v == "SMOOTH") { s->fix_type = CG_FIX_SMOOTH; s->smooth_dist = to_double(L.at(i + 1, 1)); }
        else throw Error(CG_ERR_INPUT, "ERROR: CHEBYFIX was not recognized");
        s->echo << "\t# CHEBYFIX #: " << v << "\n";
        if (v == "SMOOTH") s->echo << "\t\tCheby smoothing distance = " << s->smooth_dist << "\n";
      }
    if ((k = find_kw("USENEIG")) >= 0) {
      s->use_neigh = to_bool(L.at(k + 1, 0));
      s->echo << "\t# USENEIG #: " << bool_str(s->use_neigh) << "\n";
      if (L.width(k + 1) == 2 && s->use_neigh) {
        if (L.at(k + 1, 1) == "SMALL") {
          s->small_mode = true;
          s->echo << "\t\tWill update the neighbor list through the \"small\" method \n";
        } else throw Error(CG_ERR_INPUT, "ERROR: Unrecognized # USENEIG # option");
      }
    }
    if ((k = find_kw("SKPFRMS")) >= 0) {
      s->skip_frames = to_int(L.at(k + 1, 0));
      if (s->skip_frames >= 1)
        s->echo << "\t# SKPFRMS #: " << s->skip_frames
                << " .. will skip through frames in parallel execution (round-robin).\n";
      else
        s->echo << "\t# SKPFRMS #: " << s->skip_frames << " ... will process all frames in contiguous order\n";
    }

    // NATMTYP / TYPEIDX
    if ((k = find_kw("NATMTYP")) < 0) throw Error(CG_ERR_INPUT, "missing # NATMTYP #");
    s->ntypes = to_int(L.at(k + 1, 0));
    s->echo << "\t# NATMTYP #: " << s->ntypes << "\n";
    if (s->ntypes < 1 || s->ntypes > CG_MAX_ATOM_TYPES - 1)
      throw Error(CG_ERR_UNSUPPORTED, "NATMTYP must be 1..9 (cluster keys are base-10 digits)");
    const int nt = s->ntypes;
    s->pairs.resize(nt * (nt + 1) / 2);
    if ((k = find_kw("TYPEIDX")) < 0) throw Error(CG_ERR_INPUT, "missing # TYPEIDX #");
    s->echo << "\t# TYPEIDX #    # ATM_TYP #    # ATMCHRG #    # ATMMASS #\n";
    double sumq = 0;
    for (int j = 0; j < nt; ++j) {
      auto &p = s->pairs[j];
      int ti = to_int(L.at(k + 1 + j, 0)) - 1;
      if (ti < 0 || ti >= nt) throw Error(CG_ERR_INPUT, "Bad atom index in TYPEIDX section");
      if (ti != j) throw Error(CG_ERR_UNSUPPORTED, "TYPEIDX entries must be listed in index order");
      p.idx = j;
      p.t1 = p.t2 = j;
      p.a1 = p.a2 = L.at(k + 1 + j, 1);
      if (!s->fitcoul) p.chg1 = to_double(L.at(k + 1 + j, 2));
      else p.chrgsgn = L.at(k + 1 + j, 2);
      sumq += std::fabs(p.chg1);
      p.mass1 = p.mass2 = to_double(L.at(k + 1 + j, 3));
      p.chg2 = p.chg1;
      p.proper = p.a1 + p.a2;
      s->type_names.push_back(p.a1);
      s->type_idx.push_back(ti);
      s->echo << " \t" << std::setw(15) << std::left << j + 1 << std::setw(15) << std::left << p.a1;
      if (!s->fitcoul) s->echo << std::setw(15) << std::left << p.chg1;
      else s->echo << p.chrgsgn << "\t\t";
      s->echo << std::setw(15) << std::left << p.mass1 << std::right << "\n";
    }
    if (!s->fitcoul && sumq > 0) s->use_partial_charges = true;
    int t = nt;
    for (int i = 0; i < nt; ++i)
      for (int j = i + 1; j < nt; ++j) {
        auto &p = s->pairs[t];
        p.idx = t; p.a1 = s->pairs[i].a1; p.t1 = i; p.a2 = s->pairs[j].a1; p.t2 = j;
        p.chg1 = s->pairs[i].chg1; p.chg2 = s->pairs[j].chg1;
        p.mass1 = s->pairs[i].mass1; p.mass2 = s->pairs[j].mass1;
        p.proper = p.a1 + p.a2;
        ++t;
      }
    for (int i = 0; i < nt; ++i)
      for (int j = 0; j < nt; ++j)
        for (size_t q = 0; q < s->pairs.size(); ++q) {
          const auto &p = s->pairs[q];
          if ((s->type_names[i] == p.a1 && s->type_names[j] == p.a2) ||
              (s->type_names[j] == p.a1 && s->type_names[i] == p.a2))
            s->pair_map.insert({s->type_names[i] + s->type_names[j], (int)q});
        }
    s->echo << "\n\tThe following unique pair types have been identified:\n";
    for (const auto &p : s->pairs) s->echo << "\t\t" << p.idx << "  " << p.a1 << " " << p.a2 << "\n";
    // EXCLUDE (src/input.C:753-874): 1- and 2-body lists only in a
    // hierarchical fit, the first list of each kind
    if (s->hierarchical) {
      for (size_t i = 0; i < N; ++i)
        if (L.width(i) >= 4 && L.head3(i) == "EXCLUDE 1B INTERACTION:") {
          const int ne = to_int(L.at(i, 3));
          s->echo << "Number of 1-body interactions to exclude: " << ne << "\n";
          for (int j = 1; j <= ne; ++j)
            for (int t = 0; t < nt; ++t)
              if (L.at(i + j, 0) == s->type_names[t]) {
                s->excl_1b.push_back(t);
                s->n_excl_lt3b += 1;
                s->echo << "\tEXCLUDED: " << t << ": " << s->type_names[t] << "\n";
              }
          break;
        }
      for (size_t i = 0; i < N; ++i)
        if (L.width(i) >= 4 && L.head3(i) == "EXCLUDE 2B INTERACTION:") {
          const int ne = to_int(L.at(i, 3));
          s->echo << "Number of 2-body interactions to exclude: " << ne << "\n";
          for (int j = 1; j <= ne; ++j)
            for (size_t q = 0; q < s->pairs.size(); ++q) {
              const auto &p = s->pairs[q];
              const std::string a = L.at(i + j, 0), b = L.at(i + j, 1);
              if ((p.a1 == a && p.a2 == b) || (p.a1 == b && p.a2 == a)) {
                s->excl_2b.push_back((int)q);
                s->n_excl_lt3b += s->o2;
                s->echo << "\tEXCLUDED: " << p.idx << " " << p.a1 << " " << p.a2 << "\n";
              }
            }
          break;
        }
    }
    for (int fam = 3; fam <= 4; ++fam) {
      auto &F = fam == 3 ? s->trips : s->quads;
      F.natoms = fam;
      std::string key = "EXCLUDE " + std::to_string(fam) + "B INTERACTION:";
      for (size_t i = 0; i < N; ++i)
        if (L.head3(i) == key) {
          int ne = to_int(L.at(i, 3));
          for (int j = 1; j <= ne; ++j) {
            if ((int)L.width(i + j) != fam) throw Error(CG_ERR_INPUT, "Wrong number of atoms in exclude command");
            auto v = L.v[i + j];
            std::sort(v.begin(), v.end());
            F.exclude.push_back(v);
          }
          break;
        }
    }
    // PAIRIDX (src/input.C:1077-1146)
    if ((k = find_kw("PAIRIDX")) < 0) throw Error(CG_ERR_INPUT, "missing # PAIRIDX #");
    for (size_t i = 0; i < s->pairs.size(); ++i) {
      std::string nm = L.at(k + 1 + i, 1) + L.at(k + 1 + i, 2);
      auto it = s->pair_map.find(nm);
      if (it == s->pair_map.end()) throw Error(CG_ERR_INPUT, "unknown pair in PAIRIDX: " + nm);
      auto &p = s->pairs[it->second];
      p.smin = to_double(L.at(k + 1 + i, 3));
      p.smax = to_double(L.at(k + 1 + i, 4));
      p.lambda = to_double(L.at(k + 1 + i, 6));
      if (p.smax > s->max_cutoff_2b) s->max_cutoff_2b = p.smax;
    }
    s->echo << "\t# PAIRIDX #     # ATM_TY1 #     # ATM_TY1 #     # S_MINIM #     # S_MAXIM #     # MORSE_LAMBDA #\n";
    for (const auto &p : s->pairs)
      s->echo << "\t" << std::left << std::setw(16) << p.idx << std::setw(16) << p.a1 << std::setw(16) << p.a2
              << std::setw(16) << p.smin << std::setw(16) << p.smax << std::setw(16) << p.lambda << std::right << "\n";
    s->int_pair_map.assign(nt * nt, -1);
    for (int i = 0; i < nt; ++i)
      for (int j = 0; j < nt; ++j) {
        const int idx = s->type_idx[i] * nt + s->type_idx[j];
        s->int_pair_map[idx] = s->pair_map.at(s->type_names[i] + s->type_names[j]);
        // build_int_pair_map (src/functions.C:1055-1095)
        s->echo << "\t\tAtom type idxs: " << std::setw(2) << i << std::setw(2) << j << " Pair name: " << std::setw(4)
                << s->type_names[i] + s->type_names[j] << " Explicit pair index: " << std::setw(4) << idx
                << " Unique pair index: " << std::setw(4) << s->int_pair_map[idx] << "\n";
      }
    for (auto &p : s->pairs) cheby_params(p.smin, p.smax, p.lambda, s->transform, p.xmin, p.xmax, p.xdiff, p.xavg);
    // CHARGE CONSTRAINTS: npairs-1 lines of npairs pair names, npairs
    // weights and the b value (last token), src/input.C:1148-1200; read (and
    // echoed) whatever FITCOUL says, their rows belong to FITCOUL fits only
    // (A_MAT::PRINT_CONSTRAINTS, src/A_Matrix.C:606)
    for (size_t i = 0; i < N; ++i)
      if (L.width(i) >= 2 && L.v[i][0] == "CHARGE" && L.v[i][1] == "CONSTRAINTS:") {
        const int np = (int)s->pairs.size();
        std::vector<Setup::ChargeConstraint> cons;
        s->echo << "\n\tAttempting to read " << np - 1 << " charge constraints...:\n";
        for (int n = 0; n < np - 1; ++n) {
          Setup::ChargeConstraint cc;
          const size_t li = i + 1 + n;
          s->echo << "\t\t" << n + 1 << "\t ";
          for (int j = 0; j < np; ++j) {
            auto it = s->pair_map.find(L.at(li, j));
            if (it == s->pair_map.end()) throw Error(CG_ERR_INPUT, "unknown pair in CHARGE CONSTRAINTS: " + L.at(li, j));
            cc.pair_idx.push_back(it->second);
            s->echo << L.at(li, j) << " (" << it->second << ") ";
          }
          for (int j = 0; j < np; ++j) {
            cc.weight.push_back(to_double(L.at(li, np + j)));
            s->echo << cc.weight.back() << " ";
          }
          cc.force = to_double(L.at(li, L.width(li) - 1));
          s->echo << cc.force << "\n";
          cons.push_back(cc);
        }
        s->echo << "\n";
        if (s->fitcoul) s->charge_cons = cons;
        break;
      }
    // SPECIAL nB S_MINIM / S_MAXIM
    for (int fam = 3; fam <= 4; ++fam) {
      auto &F = fam == 3 ? s->trips : s->quads;
      for (const char *which : {"S_MINIM:", "S_MAXIM:"}) {
        std::string key = "SPECIAL " + std::to_string(fam) + "B " + which;
        for (size_t i = 0; i < N; ++i)
          if (L.head3(i) == key) {
            auto &dst = std::string(which) == "S_MAXIM:" ? F.smax_input : F.smin_input;
            std::string joined;
            for (auto &w : L.v[i]) joined += (joined.empty() ? "" : " ") + w;
            dst.push_back(joined);
            if (L.width(i) < 5) throw Error(CG_ERR_INPUT, "Not enough tokens: " + joined);
            if (L.v[i][3] == "SPECIFIC") {
              int ns = to_int(L.at(i, 4));
              for (int j = 1; j <= ns; ++j) {
                std::string jl;
                for (auto &w : L.v.at(i + j)) jl += (jl.empty() ? "" : " ") + w;
                dst.push_back(jl);
              }
            }
            break;
          }
      }
    }
    if ((k = find_kw("FCUTTYP")) >= 0) {
      std::string joined;
      for (auto &w : L.v[k + 1]) joined += (joined.empty() ? "" : " ") + w;
      s->fcut_line = joined;
      s->echo << "# FCUTTYP #: " << s->fcut_line << "      ... for all Chebyshev interactions\n";
    }
    parse_fcut(*s);
    // the reference writes the Ewald columns on force rows only
    // (src/A_Matrix.C:240-290) and stops on stress fits with Ewald
    // (src/input.C:1336-1337); chimes_lsq.py stops on hierarchical charge fits
    if (s->fitcoul && s->fit_stress != CG_STRESS_NONE)
      throw Error(CG_ERR_INPUT, "ERROR: Inclusion of stress tensors currently not compatible with use of ZCalc_Ewald_Deriv.");
    if (s->fitcoul && s->fit_ener)
      throw Error(CG_ERR_UNSUPPORTED, "FITCOUL true with FITENER: the reference's energy rows carry no Ewald columns");
    if (s->fitcoul && (!s->excl_1b.empty() || !s->excl_2b.empty()))
      throw Error(CG_ERR_UNSUPPORTED, "ERROR: cannot use hierarchical learning with charge fitting.");
    if (s->use_partial_charges) throw Error(CG_ERR_UNSUPPORTED, "fixed non-zero charges (Ewald subtraction) are not supported");
    if (s->wrap && s->nlayers > 0) {  // RUN_SANITY_LSQ
      s->wrap = false;
      s->echo << "WARNING: Coordinate wrapping not supported for ghost atom use. Turning option off\n";
    }
    if (s->nlayers < 1) throw Error(CG_ERR_UNSUPPORTED, "NLAYERS 0 (no ghost atoms) is not supported");
    // image shifts are stored in 4-bit fields (-8..+7) of a neighbour reference
    if (s->nlayers > 7) throw Error(CG_ERR_UNSUPPORTED, "NLAYERS above 7 is not supported");

    // clusters (build_clusters, src/chimes_lsq.C:533-560)
    s->trips.order = s->o3;
    s->quads.order = s->o4;
    if (s->o3 > 0) { s->trips.used = true; build_family(*s, s->trips); }
    if (s->o4 > 0) { s->quads.used = true; build_family(*s, s->quads); }
    if (s->infiles.size() == 1) s->infile_frames.push_back(s->nframes);
    if (s->fit_stress != CG_STRESS_NONE && s->nstress == -1) s->nstress = s->nframes;
    if (s->fit_ener && s->nener == -1) s->nener = s->nframes;
    flatten(*s);
  } catch (...) {
    delete s;
    throw;
  }
  return s;
}

static void flatten_family(const ClusterFamily &F, int col_base, Setup::Flat &o, cg_cluster_set &cs) {
  cs = cg_cluster_set{};
  if (!F.used) return;
  cs.nbody = F.natoms;
  cs.npairs = F.npairs;
  cs.order = F.order;
  cs.ntypes = (int)F.types.size();
  cs.max_cutoff = F.max_cutoff;
  int col = col_base, perm = 0;
  o.perm_offset.push_back(0);
  for (const auto &c : F.types) {
    o.excluded.push_back(c.excluded);
    for (int a : c.atom_types) o.atom_type.push_back(a);
    o.col_offset.push_back(col);
    o.n_unique.push_back((int)c.unique_powers.size());
    col += (int)c.unique_powers.size();
    for (size_t i = 0; i < c.allowed.size(); ++i) {
      for (int p : c.allowed[i]) o.perm_powers.push_back(p);
      o.perm_col.push_back(c.param_idx[i]);
      o.perm_equiv.push_back(c.equiv_idx[i]);
      o.perm_count.push_back(c.power_count[i]);
    }
    perm += (int)c.allowed.size();
    o.perm_offset.push_back(perm);
    for (int j = 0; j < c.npairs; ++j) {
      o.smin.push_back(c.smin[j]); o.smax.push_back(c.smax[j]);
      o.xavg.push_back(c.xavg[j]); o.xdiff.push_back(c.xdiff[j]);
    }
  }
  cs.total_allowed = perm;
  cs.ncols = col - col_base;
  o.id_map = F.id_map;
  for (auto &v : F.id_pair_map) for (int x : v) o.id_pair_map.push_back(x);
  for (auto &v : F.id_atom_perm) for (int x : v) o.id_atom_perm.push_back(x);
  cs.id_size = (int)F.id_map.size();
  cs.excluded = o.excluded.data(); cs.atom_type = o.atom_type.data(); cs.col_offset = o.col_offset.data();
  cs.n_unique = o.n_unique.data(); cs.perm_offset = o.perm_offset.data(); cs.perm_powers = o.perm_powers.data();
  cs.perm_col = o.perm_col.data(); cs.perm_equiv = o.perm_equiv.data(); cs.perm_count = o.perm_count.data();
  cs.s_min = o.smin.data(); cs.s_max = o.smax.data(); cs.x_avg = o.xavg.data(); cs.x_diff = o.xdiff.data();
  cs.id_map = o.id_map.data(); cs.id_pair_map = o.id_pair_map.data(); cs.id_atom_perm = o.id_atom_perm.data();
}

void flatten(Setup &s) {
  // pair types of canonical cluster pairs (needs the pair map)
  auto pair_types = [&](const ClusterFamily &F, std::vector<int> &out) {
    for (const auto &c : F.types)
      for (int j = 0; j < c.npairs; ++j) out.push_back(s.pair_map.at(c.pair_names[j]));
  };
  cg_model &m = s.model;
  m = cg_model{};
  m.ntypes = s.ntypes;
  s.names_flat.assign(8 * s.ntypes, 0);
  for (int i = 0; i < s.ntypes; ++i) std::strncpy(&s.names_flat[8 * i], s.type_names[i].c_str(), 7);
  m.type_names = s.names_flat.data();
  m.npairs = (int)s.pairs.size();
  for (auto &p : s.pairs) {
    s.f_pair_t1.push_back(p.t1); s.f_pair_t2.push_back(p.t2);
    s.f_smin.push_back(p.smin); s.f_smax.push_back(p.smax); s.f_lambda.push_back(p.lambda);
    s.f_xavg.push_back(p.xavg); s.f_xdiff.push_back(p.xdiff);
  }
  m.pair_t1 = s.f_pair_t1.data(); m.pair_t2 = s.f_pair_t2.data();
  m.int_pair_map = s.int_pair_map.data();
  m.pair_smin = s.f_smin.data(); m.pair_smax = s.f_smax.data(); m.pair_lambda = s.f_lambda.data();
  m.pair_xavg = s.f_xavg.data(); m.pair_xdiff = s.f_xdiff.data();
  m.transform = s.transform;
  m.order2b = s.o2;
  m.cheby_low = s.cheby_low; m.cheby_high = s.cheby_high;
  m.deriv_const = (double(s.cheby_high) - double(s.cheby_low)) / 2.0;
  m.fcut_type = s.fcut_type; m.fcut_offset = s.fcut_offset;
  m.fix_type = s.fix_type; m.smooth_dist = s.smooth_dist;
  m.nlayers = s.nlayers;
  m.small_mode = s.small_mode ? 1 : 0;
  m.fit_stress = s.fit_stress; m.fit_energy = s.fit_ener ? 1 : 0;
  m.nstress = s.nstress; m.nener = s.nener;
  s.n2b = s.o2 * (int)s.pairs.size();
  s.n3b = 0; s.n4b = 0;
  for (auto &c : s.trips.types) s.n3b += (int)c.unique_powers.size();
  for (auto &c : s.quads.types) s.n4b += (int)c.unique_powers.size();
  s.nshort = s.n2b + s.n3b + s.n4b;
  m.n2b_cols = s.n2b;
  m.ncols_short = s.nshort;
  m.n_charge_cols = s.fitcoul ? (int)s.pairs.size() : 0;
  m.n_offset_cols = s.fit_ener ? s.ntypes : 0;
  m.ncols = s.nshort + m.n_charge_cols + m.n_offset_cols;
  s.ncols = m.ncols;
  m.max_cutoff_2b = s.max_cutoff_2b;
  s.charge_sign.clear();
  if (s.fitcoul)
    for (int t = 0; t < s.ntypes; ++t) s.charge_sign.push_back(s.pairs[t].chrgsgn == "+" ? 1 : -1);
  m.charge_sign = s.charge_sign.empty() ? nullptr : s.charge_sign.data();
  flatten_family(s.trips, s.n2b, s.f3, m.c3);
  flatten_family(s.quads, s.n2b + s.n3b, s.f4, m.c4);
  pair_types(s.trips, s.f3.pair_type);
  pair_types(s.quads, s.f4.pair_type);
  m.c3.pair_type = s.f3.pair_type.data();
  m.c4.pair_type = s.f4.pair_type.data();
}

}  // namespace cg
