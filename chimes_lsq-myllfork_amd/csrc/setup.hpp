// Host-side model setup: fm_setup.in parsing, 2-body pair tables and the
// 3/4-body cluster power-set tables, flattened into the cg_model POD.
#pragma once
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>
#include <fstream>
#include <istream>

#include "../../include/chimes_gpu.h"

namespace cg {

struct Error : std::runtime_error {
  cg_status code;
  Error(cg_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// One 2-body pair type (reference: class PAIRS, src/Cluster.h:7-82).
struct PairType {
  int idx = 0;
  std::string a1, a2, proper;   // ATM1TYP, ATM2TYP, PRPR_NM
  int t1 = 0, t2 = 0;           // atom-type indices
  double chg1 = 0, chg2 = 0, mass1 = 0, mass2 = 0;
  std::string chrgsgn;
  double smin = 0.8, smax = 5.0, lambda = 0;
  double xmin = 0, xmax = 0, xavg = 0, xdiff = 0;
};

// One n-body cluster type (reference: class CLUSTER, src/Cluster.h).
struct ClusterType {
  int index = 0, natoms = 0, npairs = 0;
  std::vector<std::string> atom_names;  // sorted
  std::vector<int> atom_types;          // type index per canonical slot
  std::vector<std::string> pair_names;  // canonical pair names (proper)
  bool excluded = false;
  bool special_min = false, special_max = false;
  std::vector<std::vector<int>> unique_powers, allowed;
  std::vector<int> param_idx, equiv_idx, power_count;
  std::map<std::vector<int>, int> allowed_lookup;
  std::vector<double> smin, smax, xmin, xmax, xavg, xdiff;
};

struct ClusterFamily {
  int natoms = 0, npairs = 0, order = 0;
  bool used = false;
  std::vector<ClusterType> types;
  std::map<std::string, int> name_map;       // MAP (pair-name string -> index)
  std::map<int, std::string> name_rev;       // MAP_REVERSE
  std::vector<int> id_map;                   // INT_MAP
  std::vector<std::vector<int>> id_pair_map; // PAIR_INDICES
  std::vector<std::vector<int>> id_atom_perm;
  std::vector<std::vector<std::string>> exclude;
  std::vector<std::string> smax_input, smin_input;
  double max_cutoff = 0;
  double min_cutoff = 0;                     // smallest S_MINIM of any type
};

struct Setup {
  std::string path, dir;
  // What chimes_lsq prints while it reads the input and builds the cluster
  // tables (src/input.C:172-1338, src/Cluster.C:368-426,771-954,1293-1556),
  // in its order; printed by the stdout report.
  std::ostringstream echo;
  std::string multi_file;   // TRJFILE MULTI list file ("" for one file)
  double report_ewald_cell[3] = {0, 0, 0};  // stdout report: cell of the last Ewald parameter print
  // controls
  std::vector<std::string> infiles;
  std::vector<int> infile_frames;
  std::vector<std::string> force_flags, stress_flags, energy_flags;
  bool wrap = false, split_files = false, hierarchical = false;
  // hierarchical fit (HIERARC true): excluded atom types (energy-offset
  // columns) and pair types (2-body columns), in input order, and the number
  // of columns they remove (N_EXCL_LT_3B, src/input.C:760-840)
  std::vector<int> excl_1b, excl_2b;
  int n_excl_lt3b = 0;
  int nframes = 0, nlayers = 0, skip_frames = 0;
  // FITCOUL: one Ewald column per pair type, and the CHARGE CONSTRAINTS
  // rows appended after the last frame (src/input.C:1148-1200): per row the
  // pair index and weight of each listed pair, and the row's b value
  struct ChargeConstraint {
    std::vector<int> pair_idx;
    std::vector<double> weight;
    double force = 0.0;
  };
  std::vector<ChargeConstraint> charge_cons;
  std::vector<int> charge_sign;  // FITCOUL: +1 for "+", else -1 (chimes_md.C:1448-1463)
  bool cons_written = false;  // text output: constraint rows already written
  bool fitcoul = false, fit_ener = false, use_partial_charges = false;
  int fit_stress = CG_STRESS_NONE, nstress = -1, nener = -1;
  std::string ff_type;
  int o2 = 0, o3 = 0, o4 = 0, cheby_low = -1, cheby_high = 1;
  int transform = CG_NONE;
  int fix_type = CG_FIX_SMOOTH;
  double smooth_dist = 0.01;
  bool use_neigh = true, small_mode = false;
  std::string fcut_line = "CUBIC";
  int fcut_type = CG_FCUT_CUBIC;
  double fcut_offset = 0.0;
  // topology
  int ntypes = 0;
  std::vector<std::string> type_names;
  std::vector<int> type_idx;
  std::vector<PairType> pairs;
  std::map<std::string, int> pair_map;
  std::vector<int> int_pair_map;
  ClusterFamily trips, quads;
  double max_cutoff_2b = 0;
  int n2b = 0, n3b = 0, n4b = 0, nshort = 0, ncols = 0;

  // flattened model
  cg_model model{};
  std::vector<char> names_flat;
  std::vector<int> f_pair_t1, f_pair_t2;
  std::vector<double> f_smin, f_smax, f_lambda, f_xavg, f_xdiff;
  struct Flat {
    std::vector<int> excluded, atom_type, col_offset, n_unique, perm_offset, perm_powers,
        perm_col, perm_equiv, perm_count, pair_type, id_map, id_pair_map, id_atom_perm;
    std::vector<double> smin, smax, xavg, xdiff;
  } f3, f4;

  // coefficients read from a params.txt (A column order) and the chimes_md
  // 2-body penalty settings found there (defaults of src/chimes_md.C:1747-1748)
  std::vector<double> coeffs;
  double penalty_dist = 0.01, penalty_scale = 1.0e4;

  // text output state
  std::ofstream fA, fb, fbl, fna;
  int64_t data_count = 0;
  std::string outdir;
  bool write_A = true;               // A text written (cg_write_open_opts)
  int nranks = 1;                    // ranks of the frame distribution (cg_frame_order)
  int cur_rank = 0;                  // rank of the frames being written
  bool rank_open = false;            // A.<rank>.txt of cur_rank open (SPLITFI)
  std::vector<int64_t> rank_rows;    // rows written per rank
  std::string abuf;                  // formatted A text awaiting a write
};

void set_error(const std::string &m);  // thread-local message for cg_last_error
Setup *read_setup(const std::string &path);
Setup *read_setup_from(std::istream &in, const std::string &path);
Setup *read_params(const std::string &path);  // params.txt -> model + coefficients
// fm_setup.in token rules ('!' and '##' comments stripped, whitespace split)
std::vector<std::string> split_tokens(const std::string &line);
int parse_int(const std::string &v);
double parse_double(const std::string &v);
void flatten(Setup &s);
int type_of(const Setup &s, const std::string &name); // -1 if unknown
void write_header(const Setup &s, const std::string &dir);
void write_map(const Setup &s, const std::string &dir);

}  // namespace cg

// Handle of the C ABI (include/chimes_gpu.h).
struct cg_setup {
  cg::Setup *s;
};
