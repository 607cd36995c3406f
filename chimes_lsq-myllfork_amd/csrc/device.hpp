// Device-side context shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/chimes_gpu.h"

namespace cg {

void set_error(const std::string &m);

#define CG_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      throw ::cg::Error(CG_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Set on a thread while it captures a graph: a buffer that would have to
// grow then fails instead (allocation is not part of a capture).
extern thread_local bool t_no_grow;
// Buffer moves (device or pinned) made by this thread: each one is an
// allocation, and a device free also waits for the device.
extern thread_local uint64_t t_grow_events;

struct DBuf;

// Bump allocator over a fixed device region (the allocation-free force entry
// point): while one is installed on a thread, DBuf::get hands out pieces of
// it instead of calling hipMalloc.  A buffer asked for again during the same
// call gets the same piece while it fits (DBuf semantics: contents are kept
// only while the buffer does not grow).  Nothing is freed; reset() starts
// over.
struct Arena {
  static constexpr int kMaxPieces = 96;
  char *base = nullptr;
  size_t cap = 0, used = 0;
  size_t wanted = 0;        // bytes the call would have needed (overflow report)
  bool overflow = false;
  int n = 0;
  const DBuf *owner[kMaxPieces];
  void *ptr[kMaxPieces];
  size_t size[kMaxPieces];
  void reset() { used = 0; wanted = 0; overflow = false; n = 0; }
  void *take(const DBuf *b, size_t bytes) {
    for (int i = n - 1; i >= 0; --i)
      if (owner[i] == b) {
        if (bytes <= size[i]) return ptr[i];
        break;
      }
    const size_t off = (used + 255) & ~size_t(255);
    wanted = std::max(wanted, off + bytes);
    if (n >= kMaxPieces || off + bytes > cap) { overflow = true; return nullptr; }
    owner[n] = b; ptr[n] = base + off; size[n] = bytes; ++n;
    used = off + bytes;
    return base + off;
  }
};
// Installed by the allocation-free entry point for the duration of its call.
extern thread_local Arena *t_arena;

// Grow-only device buffer.  A buffer owned by a context grows through the
// stream-ordered allocator on the context's stream (set at context creation):
// moving it waits for nothing but that stream's earlier work, so a context
// that grows does not stall the others sharing the GPU (hipFree waits for the
// whole device).
struct DBuf {
  void *p = nullptr;
  size_t cap = 0;
  hipStream_t s = nullptr;  // stream of stream-ordered growth (null: hipMalloc / hipFree)
  void *get(size_t bytes) {
    if (t_arena) return t_arena->take(this, bytes);
    if (bytes > cap) {
      if (t_no_grow) return nullptr;
      // geometric growth (at least twice the old capacity, at most twice the
      // request): a buffer that follows growing requests moves O(log) times
      const size_t old = cap;
      free_now();
      ++t_grow_events;
      size_t n = bytes + bytes / 4 + 256;
      if (old > 0) n = std::max(n, std::min(2 * old, 2 * bytes + 256));
      const hipError_t e = s ? hipMallocAsync(&p, n, s) : hipMalloc(&p, n);
      if (e != hipSuccess) { p = nullptr; cap = 0; return nullptr; }
      cap = n;
    }
    return p;
  }
  template <class T> T *as(size_t count) { return static_cast<T *>(get(count * sizeof(T))); }
  void free_now() {
    if (p) (void)(s ? hipFreeAsync(p, s) : hipFree(p));
    p = nullptr;
    cap = 0;
  }
  void release() { free_now(); }
};

// Per-frame geometry handed to the kernels (computed on the host exactly as
// BOX::UPDATE_CELL does, src/ClassDefs.C:1940-2211).
struct DevFrame {
  int natoms;
  int atom_off;        // first atom of the frame in the batch arrays
  long long row_off;   // first row of the frame in A
  int stress_rows;     // 0, 3 or 9
  int ener_rows;       // 0 or 3
  int small;           // unordered neighbour lists (NEIGHBORS::DO_UPDATE_SMALL)
  int ortho;
  double H[9];         // HMAT (row-major: H[r*3+c])
  double Hi[9];        // INVR_HMAT
  double lx, ly, lz;   // LAMMPS lengths used by the orthorhombic wrap
  double width[3];     // perpendicular cell widths
  double min_extent;   // smallest LAMMPS extent (small-cell switch)
  double inv_vol;
  double perm_scale[5];
  int cell;            // neighbour search through fractional-coordinate bins
  int nbin[3];         // bins per cell axis (each at least the list cutoff thick)
  int bin_off;         // first bin of the frame in the batch bin arrays
  int nlayers;         // periodic image layers searched along each cell axis
};

// Neighbour image reference: atom index in the frame (20 bits) and the image
// shift along each cell axis (4 bits each, biased by 8: shifts -8..+7).
constexpr int kRefShiftBias = 8;
constexpr int kMaxLayers = 7;  // largest |shift| an image layer search may need
__host__ __device__ constexpr uint32_t make_ref(int a, int sx, int sy, int sz) {
  return uint32_t(a) | (uint32_t(sx + kRefShiftBias) << 20) | (uint32_t(sy + kRefShiftBias) << 24) |
         (uint32_t(sz + kRefShiftBias) << 28);
}
__host__ __device__ constexpr int ref_atom(uint32_t r) { return int(r & 0xFFFFFu); }
__host__ __device__ constexpr int ref_sx(uint32_t r) { return int((r >> 20) & 15u) - kRefShiftBias; }
__host__ __device__ constexpr int ref_sy(uint32_t r) { return int((r >> 24) & 15u) - kRefShiftBias; }
__host__ __device__ constexpr int ref_sz(uint32_t r) { return int((r >> 28) & 15u) - kRefShiftBias; }
// Every shift a search of kMaxLayers layers produces, with the largest atom
// index, decodes to itself.
constexpr bool ref_round_trip() {
  for (int sx = -kMaxLayers; sx <= kMaxLayers; ++sx)
    for (int sy = -kMaxLayers; sy <= kMaxLayers; ++sy)
      for (int sz = -kMaxLayers; sz <= kMaxLayers; ++sz)
        for (int a : {0, 1, 0xFFFFF}) {
          const uint32_t r = make_ref(a, sx, sy, sz);
          if (ref_atom(r) != a || ref_sx(r) != sx || ref_sy(r) != sy || ref_sz(r) != sz) return false;
        }
  return true;
}
static_assert(ref_round_trip(), "image reference encoding");

// Sizes of a call without host reads: capacity of the neighbour list and of
// each family's cluster list.  The kernels write nothing past them; the
// device totals tell afterwards whether the call fitted.
// A work item of the assembly (a column tile of one (frame, type) bucket)
// or force kernels (a chunk of clusters of one bucket).
struct Task {
  int frame;
  int type;
  int col0;    // first column of the tile, relative to the type (assembly)
  int ncols;   // tile width (assembly)
  int c0, c1;  // cluster range in the bucketed arrays (force evaluation)
};
// Tasks [begin, end) of one assembly launch: frame-size classes cls0..cls1
// under one tile configuration, with the largest LDS tile among them.
// gacc: the tasks accumulate in A itself (frames beyond the LDS tile).
struct TaskGroup { size_t begin, end; int cls0, cls1; size_t lds; bool gacc = false; };

struct Caps {
  int nnb = 0;
  int ncl[3] = {0, 0, 0};
};

// Device copy of one cluster family (3B or 4B) or the 2B pair family.
struct DevFamily {
  int np = 0;            // pairs per cluster (1, 3, 6)
  int nb = 0;            // atoms per cluster (2, 3, 4)
  int ntypes = 0;        // cluster (or pair) types
  int ncols = 0;         // columns of the family
  int max_pow = 0;       // table width needed (largest power index + 1)
  int qform = 0;         // 4-body set loops in quotient form (6 or more sets per column on average)
  // 3/4-body pair tables computed once per neighbour entry (FP64 or FP32):
  // set when every cluster slot of a pair type has that type's cutoffs and
  // scaling, so a table depends on the entry alone; ptp holds per 2-body pair
  // type smin, smax, xavg, 1/lambda, 1/xdiff, 1/smax
  int pcache = 0;
  double *ptp = nullptr;
  double cutoff = 0;     // list cutoff of the family
  std::vector<int> h_col_offset;   // absolute first column per type
  std::vector<int> h_ncols;        // columns per type
  std::vector<int> h_excluded;
  // device tables
  int *col_offset = nullptr;  // [2*ntypes] family-relative then absolute first column
  int *ncol_type = nullptr;   // [ntypes] columns per type
  int *colptr = nullptr;      // [ncols+1] first perm of each family column (family-relative)
  uint32_t *perm = nullptr;   // [nperm] packed powers, 4 bits per canonical pair
  int *sets8 = nullptr;       // set records (2 * power per pair): 4-body powers < 5 [nperm + 2][8], 3-body powers < 9 [nperm + 2][4]
  double *smin = nullptr, *smax = nullptr, *xavg = nullptr, *xdiff = nullptr, *lambda = nullptr; // [ntypes*np]
  double *recip = nullptr;  // [ntypes*np*3]: 1/lambda, 1/xdiff, 1/smax (assembly tables)
  int *cct = nullptr;         // [ntypes_atom^nb] cluster type of a type tuple, -1 if none
  double *csmax = nullptr;    // [ntypes_atom^nb * np] S_MAXIM of each enumerated pair slot
  int *cperm = nullptr;       // [ntypes_atom^nb * nb] canonical slot of each enumerated atom
  int *cslot = nullptr;       // [ntypes_atom^nb * np] canonical pair slot of each enumerated pair (PAIR_INDICES)
};

// Statistics of the fit's clusters that chimes_lsq prints at the end of a
// run (print_bond_stats, src/chimes_lsq.C:474-531, CLUSTER_LIST::
// print_min_distances, src/Cluster.C:1569-1633) and the per-frame sizes of
// its flat 3- and 4-body interaction lists (NEIGHBORS::UPDATE_3B/4B_
// INTERACTION, src/ClassDefs.C:1513-1625).  Minima are kept as the bit
// patterns of positive doubles (their order is that of the values).
struct LsqStats {
  bool on = false;
  unsigned long long *d_min = nullptr;  // [npairs | ntypes3*3 | ntypes4*6]
  int n_min = 0, off3 = 0, off4 = 0;
  std::vector<long long> count[3];       // per pair / cluster type
  long long *d_flat = nullptr;           // [2 * frames of the last call]: 3-body, 4-body
  size_t flat_cap = 0;
  std::vector<long long> flat3, flat4;   // of the last call's frames
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // assembly launches of further frame-size classes run on these (default
  // priority), forked from and joined back to `stream`
  // (one per launch after the first: the size classes and the
  // global-accumulation launch)
  hipStream_t aux[CG_TILE_CLASSES] = {};
  hipEvent_t fork = nullptr, join[CG_TILE_CLASSES] = {};
  const cg_model *model = nullptr;
  cg_model m{};               // copy of the POD header
  DevFamily fam[3];           // 0: 2B pairs, 1: 3B, 2: 4B
  int *d_int_pair_map = nullptr;
  int sticky_small = 0;       // NEIGHBORS::UPDATE_WITH_BIG is sticky across frames
  double penalty_scale = 1.0e4;  // PAIR CHEBYSHEV PENALTY SCALING default (chimes_md)
  double penalty_dist = 0.01;    // PAIR CHEBYSHEV PENALTY DIST default
  // batch buffers
  DBuf b_frames, b_xyz, b_type, b_wrap;   // frame table, coordinates, types, wrapped coordinates
  DBuf b_ncount, b_noff, b_nlist, b_nvec, b_nown;  // neighbour lists
  DBuf b_sub[2];                          // 3B / 4B neighbour sublists
  DBuf b_bins;                            // cell-list keys, atoms and bin starts
  DBuf b_ccount, b_coff;                  // per-entry cluster counts / offsets (3 families)
  DBuf b_fam[3];                          // per-family clusters, keys and bucket bounds
  DBuf b_tasks[3];                        // per-family assembly / force task lists
  DBuf b_tcount[3];                       // force tasks: chunks per bucket and their offsets
  DBuf b_A, b_scan;                       // design matrix; hipcub scratch
  DBuf b_x;                               // fit workspace
  DBuf b_atvp;                            // fit: per-slab partial sums of A^T v
  DBuf b_ptab[2];                         // 3/4-body pair tables per neighbour entry (F, then Q or D)
  DBuf b_pidx[2];                         // 3/4-body: neighbour entry of each cluster pair (-1: none)
  DBuf b_fitkeep;                         // kept columns of a fit (device copy)
  DBuf b_qr;                              // QR fit: FP64 copy of A, factors, SVD
  DBuf b_fitw;                            // row weights of a fit (device copy)
  DBuf b_lars;                            // LassoLars: Cholesky factor, coefficients
  std::vector<double> fit_w;              // row weights (empty: none)
  std::vector<int> fit_keep;              // columns a fit keeps (empty: all)
  DBuf b_force;                           // force evaluation: F, energy/virial, call totals
  std::vector<double> h_out;              // host copy of the force results and totals
  double rate_nb = 0.0;                   // largest neighbour entries per atom seen (0: none yet)
  double rate_cl[3] = {0.0, 0.0, 0.0};    // largest clusters per atom seen, per family
  int64_t n_force = 0, n_repeat = 0;      // force calls, and those repeated with exact sizes
  int64_t n_graph = 0;                    // force calls served by replaying a captured graph
  int64_t n_capture = 0;                  // force calls captured into a new graph
  // captured force calls by shape (frame sizes, cell lists, list capacities);
  // a shape is captured the second time in a row it is seen
  // captured force calls per shape key: the buffer addresses they were
  // captured with (fingerprint) and the last use (least recently used graph
  // is evicted when the cache is full)
  struct Graph { hipGraphExec_t exec = nullptr; uint64_t buffers = 0; uint64_t used = 0; };
  uint64_t graph_clock = 0;
  std::map<std::string, Graph> graphs;
  std::string last_key;
  // list capacities of the last capacity-mode call (force calls and
  // assemblies separately) and its atom count
  struct CapPlan { Caps last; int natoms = -1; };
  CapPlan cap_force, cap_asm;
  // assemblies: repeated with exact sizes, replayed from a graph, captured;
  // capture of repeated shapes on/off (cg_assembly_graphs)
  int64_t n_asm = 0, n_asm_repeat = 0, n_asm_graph = 0, n_asm_capture = 0;
  bool asm_graphs = true;
  std::string last_asm_key;               // shape of the last capacity-mode assembly
  bool asm_replayed = false;              // the last assembly was a graph replay
  DBuf b_atot;                            // device totals of an assembly
  // assembly task lists of the last shape (frame sizes, tile choice), kept on
  // the host and in b_atasks while the shape repeats
  DBuf b_atasks[3];
  std::string asm_task_key;
  std::vector<Task> asm_tasks[3];
  std::vector<TaskGroup> asm_groups[3];
  // frames with more atoms than this accumulate in A with global atomics
  // (0: only frames whose force rows do not fit the LDS tile)
  int gacc_above = 0;
  bool asm_det = false;                   // bitwise-reproducible assembly (cg_assembly_deterministic)
  bool asm_pcache = true;                 // 3/4-body pair tables per neighbour entry (cg_assembly_pair_tables)
  double *h_atot = nullptr;               // pinned host copy of them (4 doubles)
  double *h_pin = nullptr;                // pinned host copy of results and totals
  size_t h_pin_cap = 0;
  std::vector<double *> h_pin_retired;    // outgrown pinned blocks (freed with the context)
  bool capturing = false;
  DBuf b_coef[3];                         // per-family coefficient tensors (force evaluation)
  int coef_size[3] = {0, 0, 0};           // tensor entries per cluster type
  bool coef_valid = false;
  std::vector<double> h_x_last;           // coefficients the tensors were built from
  DBuf b_tmp;                             // matvec output
  DBuf b_ewald;                           // Ewald frame table, k-vectors, structure factors
  std::vector<char> h_ewald;              // host source of the Ewald table upload
  std::vector<double> q_type;             // force evaluation with Ewald terms: charge per atom type (e)
  void *A = nullptr;          // last assembled A (device)
  int A_prec = CG_FP64;
  int64_t A_rows = 0;
  int A_ld = 0;
  std::vector<double> h_b;    // last b
  float ms_enum = 0, ms_asm = 0, ms_total = 0;
  hipEvent_t ev[4] = {};
  hipEvent_t fev[3][2] = {};   // around each family's assembly launches
  float ms_fam[3] = {0, 0, 0};
  // per (family, frame-size class): events around the class's launch and its
  // time in the last assembly (0 when the class had no task)
  hipEvent_t cev[3][CG_TILE_CLASSES][2] = {};
  float ms_cls[3][CG_TILE_CLASSES] = {};
  long long ncl[3] = {0, 0, 0};  // clusters per family in the last batch
  int tile_cfg[3][CG_TILE_CLASSES] = {};  // assembly tile configuration per (family, frame-size class)
  bool split_classes = false;             // one launch per frame-size class (timing by cg_autotune)
  LsqStats stats;                // fit statistics (cg_stats_enable)
};

// Defined in assemble.hip.
void ctx_init(Ctx &c, const cg_model *m, int device);
void ctx_free(Ctx &c);
void force_batch(Ctx &c, const double *x_host, const cg_frame *frames, int nframes, double *F, double *E,
                 double *vir);

// Defined in fit.hip / lars.hip.
void matvec(Ctx &c, const double *x_host, double *y_host);
void gram_and_atb(Ctx &c, const double *b_host, double **G, double **Atb);

// Defined in ewald.hip: Ewald parameters of a frame (optimal_ewald_params) and
// the FITCOUL charge columns of an assembled batch.
void ewald_params(int natoms, const double L[3], double *alpha, double *rc, int *kmax);
void ewald_columns(Ctx &c, const cg_frame *frames, const std::vector<DevFrame> &hf, const int *d_af,
                   const int *d_type, const double *d_wrap, int natoms_total, void *A, long long ld, int prec);
// Charges per atom type from the Ewald columns of x, and the Ewald forces,
// energies and virials of a force call added to its device results.
void ewald_type_charges(const cg_model &m, const double *x, std::vector<double> &qt);
void ewald_md(Ctx &c, const cg_frame *frames, const std::vector<DevFrame> &hf, const int *d_af, const int *d_type,
              const double *d_wrap, int natoms_total, double *d_F, double *d_EV);

}  // namespace cg

struct cg_ctx {
  cg::Ctx c;
};
