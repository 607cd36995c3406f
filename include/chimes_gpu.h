/*
 * chimes_gpu.h -- C ABI of the MI355X (gfx950) ChIMES design-matrix / fit /
 * force-evaluation path.
 *
 * Everything here is plain C: POD structs, pointers and sizes, integer status
 * codes.  No exceptions or process exits cross this boundary; the last error
 * message of a context (or of the thread, for calls without one) is returned
 * by cg_last_error().
 *
 * Reference interfaces replaced (paths relative to the ChIMES-lsq tree):
 *   cg_setup_read         INPUT::PARSE_INFILE_LSQ + build_clusters
 *                         (src/input.C:172-225, src/chimes_lsq.C:533-560)
 *   cg_traj_*             FRAME::READ_XYZF (src/ClassDefs.C:2411-2683)
 *   cg_assemble           ZCalc_Deriv -> Cheby::Deriv_2B/3B/4B for a batch of
 *                         frames (src/functions.C:755-804, src/Cheby.C:709-1729)
 *                         plus A_MAT row layout (src/A_Matrix.C:210-567)
 *   cg_write_*            A_MAT::PRINT_FRAME / CLEANUP_FILES, print_param_header,
 *                         print_map_file (src/A_Matrix.C:210-760,
 *                         src/chimes_lsq.C:562-716)
 *   cg_fit                chimes_lsq.py SVD path (src/chimes_lsq.py:169-214,284-306)
 *   cg_write_params       chimes_lsq.py params.txt printer (src/chimes_lsq.py:315-672)
 *   cg_force              ZCalc -> Cheby::Force_all/Force_3B/Force_4B
 *                         (src/functions.C:815-891, src/Cheby.C:1731-2675)
 *   cg_serve_*            no reference counterpart (the reference evaluates one
 *                         configuration per MD step); a request queue over cg_force.
 */
#ifndef CHIMES_GPU_H
#define CHIMES_GPU_H

#include <stdint.h>
#include <stddef.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
typedef int cg_status;
enum {
  CG_OK = 0,
  CG_ERR_ARG = 1,        /* bad argument / inconsistent sizes            */
  CG_ERR_INPUT = 2,      /* malformed fm_setup.in / trajectory           */
  CG_ERR_UNSUPPORTED = 3,/* option outside this path (FITCOUL, ...)      */
  CG_ERR_DEVICE = 4,     /* HIP runtime error                            */
  CG_ERR_MEMORY = 5,     /* allocation / budget                          */
  CG_ERR_NUMERIC = 6,    /* solver failure                               */
  CG_ERR_EOF = 7         /* trajectory exhausted                         */
};

/* Last error message for a context (ctx may be NULL: thread-local message). */
const char *cg_last_error(const void *ctx);

/* ----------------------------------------------------------- model (POD) */
#define CG_MAX_ATOM_TYPES 10   /* src/functions.h:34 (INT_MAP key base)      */

enum { CG_MORSE = 0, CG_INVRSE_R = 1, CG_NONE = 2 };            /* Cheby_trans */
enum { CG_FCUT_CUBIC = 0, CG_FCUT_TERSOFF = 1 };                /* FCUT_TYPE   */
enum { CG_FIX_ZERO_DERIV = 0, CG_FIX_CONSTANT_DERIV = 1, CG_FIX_SMOOTH = 2 };
enum { CG_STRESS_NONE = 0, CG_STRESS_DIAG = 1, CG_STRESS_ALL = 2 };

/* One n-body cluster family (3-body triplets or 4-body quadruplets).
 * Cluster types are in the reference's CLUSTER_LIST::VEC order; canonical
 * atom order inside a type is the sorted-name order of CLUSTER::ATOM_NAMES,
 * canonical pair m enumerates atom pairs (i<j) of that order. */
typedef struct cg_cluster_set {
  int nbody;          /* 3 or 4; 0 when the family is unused             */
  int npairs;         /* 3 or 6                                          */
  int order;          /* powers run 0..order-1 (PAIRTYP token)           */
  int ntypes;         /* cluster types, excluded ones included           */
  int total_allowed;  /* sum of N_ALLOWED_POWERS                         */
  int ncols;          /* sum of N_TRUE_ALLOWED_POWERS                    */
  double max_cutoff;  /* CLUSTER_LIST::MAX_CUTOFF                        */
  const int *excluded;      /* [ntypes]                                  */
  const int *atom_type;     /* [ntypes*nbody] atom-type index per slot   */
  const int *col_offset;    /* [ntypes] absolute first column            */
  const int *n_unique;      /* [ntypes] N_TRUE_ALLOWED_POWERS            */
  const int *perm_offset;   /* [ntypes+1] first allowed set of a type    */
  const int *perm_powers;   /* [total_allowed*npairs] ALLOWED_POWERS     */
  const int *perm_col;      /* [total_allowed] PARAM_INDICES (type-rel.) */
  const int *perm_equiv;    /* [total_allowed] EQUIV_INDICES             */
  const int *perm_count;    /* [total_allowed] POWER_COUNT               */
  const double *s_min;      /* [ntypes*npairs] per canonical pair        */
  const double *s_max;
  const double *x_avg;
  const double *x_diff;
  const int *pair_type;     /* [ntypes*npairs] 2-body pair-type index    */
  int id_size;              /* CG_MAX_ATOM_TYPES^nbody                   */
  const int *id_map;        /* [id_size] INT_MAP (-1: excluded/absent)   */
  const int *id_pair_map;   /* [id_size*npairs] PAIR_INDICES             */
  const int *id_atom_perm;  /* [id_size*nbody] canonical slot -> position
                               of that atom in the enumerated tuple      */
} cg_cluster_set;

typedef struct cg_model {
  int ntypes;                 /* NATMTYP                                   */
  const char *type_names;     /* ntypes * 8 bytes, NUL padded              */
  int npairs;                 /* 2-body pair types                         */
  const int *pair_t1;         /* [npairs] atom type of ATM1                */
  const int *pair_t2;
  const int *int_pair_map;    /* [ntypes*ntypes] INT_PAIR_MAP              */
  const double *pair_smin;    /* [npairs] S_MINIM                          */
  const double *pair_smax;    /* [npairs] S_MAXIM                          */
  const double *pair_lambda;  /* [npairs] MORSE_LAMBDA                     */
  const double *pair_xavg;    /* [npairs] X_AVG                            */
  const double *pair_xdiff;   /* [npairs] X_DIFF                           */
  int transform;              /* CG_MORSE / CG_INVRSE_R / CG_NONE          */
  int order2b;                /* SNUM: 2-body columns per pair type        */
  double deriv_const;         /* (CHEBY_RANGE_HIGH-CHEBY_RANGE_LOW)/2      */
  int cheby_low, cheby_high;  /* PAIRTYP range tokens                      */
  int fcut_type;              /* CG_FCUT_*                                 */
  double fcut_offset;
  int fix_type;               /* CG_FIX_*                                  */
  double smooth_dist;
  int nlayers;
  int small_mode;             /* 1: USENEIG ... SMALL requested            */
  int fit_stress;             /* CG_STRESS_*                               */
  int fit_energy;
  int nstress, nener;         /* -1: every frame                           */
  int n2b_cols;               /* TOT_SNUM                                  */
  int ncols_short;            /* TOT_SHORT_RANGE                           */
  int n_charge_cols;          /* FITCOUL Ewald columns (0 or npairs), right
                                 after the short-range columns             */
  int n_offset_cols;          /* energy-offset columns (0 or ntypes)       */
  int ncols;                  /* columns of A (TOT_ALL_PARAMS)             */
  double max_cutoff_2b;       /* NEIGHBORS::MAX_CUTOFF                     */
  cg_cluster_set c3, c4;
  const int *charge_sign;     /* [ntypes] FITCOUL: +1 for "+" in ATMCHRG,
                                 else -1 (chimes_md's reading), or NULL    */
} cg_model;

/* -------------------------------------------------------------- frames */
typedef struct cg_frame {
  int natoms;
  const double *xyz;      /* [natoms*3] Angstrom, as read                  */
  const int *type;        /* [natoms] atom-type index                      */
  double cell[9];         /* a, b, c cell vectors (row-wise)               */
  const double *forces;   /* [natoms*3] kcal/mol/A (converted) or NULL     */
  double stress[6];       /* xx yy zz xy xz yz, GPa as read                */
  double energy;          /* kcal/mol                                      */
  int index;              /* frame index in the trajectory                 */
} cg_frame;

/* --------------------------------------------------- host-side setup */
typedef struct cg_setup cg_setup;
cg_status cg_setup_read(const char *fm_setup_path, cg_setup **out);
void cg_setup_free(cg_setup *s);
/* Model + coefficients from a params.txt (the file chimes_md reads,
 * src/chimes_md.C:1474-2179): the model is rebuilt as chimes_lsq would for the
 * same species, pair table, orders, cutoffs and exclusions, and every
 * coefficient block is checked against it.  Coulomb, over-coordination and
 * 1B/2B exclusion terms are refused (CG_ERR_UNSUPPORTED). */
cg_status cg_params_read(const char *path, cg_setup **out);
/* Coefficients in A column order (x[ncols]; energy offsets last) and the
 * file's 2-body penalty settings. */
cg_status cg_params_coeffs(const cg_setup *s, double *x, double *penalty_dist, double *penalty_scale);
/* Out-of-range treatment of pair distances below S_MINIM (CG_FIX_*, with the
 * SMOOTH distance): chimes_md's # CHEBYFIX # run option
 * (INPUT::PARSE_CONTROLS_CHEBYFIX, src/input.C:713-747), which a params.txt
 * does not carry.  Applies to contexts created afterwards from this setup. */
cg_status cg_setup_set_chebyfix(cg_setup *s, int fix_type, double smooth_dist);
const cg_model *cg_setup_model(const cg_setup *s);
int cg_setup_nframes(const cg_setup *s);
/* rows of A produced by a frame (3N + stress rows + energy rows) */
int64_t cg_model_frame_rows(const cg_model *m, int natoms, int frame_index);

/* Trajectory reader: frames are owned by the reader until the next call. */
typedef struct cg_traj cg_traj;
cg_status cg_traj_open(const cg_setup *s, const char *workdir, cg_traj **out);
cg_status cg_traj_next(cg_traj *t, cg_frame *f);
void cg_traj_close(cg_traj *t);

/* Text writers (byte layout of the reference's files).  rows: row-major
 * [nrows*ncols] for the frame in reference row order.
 * With SPLITFI true the A rows of each rank go to A.<rank>.txt and
 * cg_write_close writes dim.<rank>.txt ("cols first_row last_row total_rows")
 * for every rank with rows (A_Matrix.C:626-720); otherwise A.txt and dim.txt.
 * b.txt, b-labeled.txt and natoms.txt are single files in rank order, as the
 * reference's concatenation leaves them. */
cg_status cg_write_open(const cg_setup *s, const char *outdir);
enum { CG_WRITE_NO_A = 1 };  /* flags of cg_write_open_opts: skip the A text */
cg_status cg_write_open_opts(const cg_setup *s, const char *outdir, int flags);
/* Following frames belong to `rank` (ranks in increasing order). */
cg_status cg_write_rank(const cg_setup *s, int rank);
cg_status cg_write_frame(const cg_setup *s, const cg_frame *f, const double *rows, int nrows);
cg_status cg_write_close(const cg_setup *s);

/* Which frames the reference's `nranks` MPI ranks process and in what order
 * (src/chimes_lsq.C:263,305-319, divide_atoms in src/functions.C:182-222):
 * contiguous blocks, or with SKPFRMS >= 1 round robin (rank r takes frame i
 * when (i + r) % nranks == 0 and every SKPFRMS-th such frame).  order
 * [nframes] receives the processed frames rank by rank, rank [nframes] the
 * rank of each, n_out their number. */
cg_status cg_frame_order(const cg_setup *s, int nranks, int *order, int *rank, int *n_out);
cg_status cg_write_header(const cg_setup *s, const char *outdir); /* params.header + ff_groups.map */

/* ------------------------------------------------ standard output report
 * The text chimes_lsq prints to standard output (its test suite keeps it as
 * fm_setup.out) for a run of this setup, written to `out` in pieces the
 * caller orders as the reference does (src/chimes_lsq.C:160-424):
 *   cg_report_setup        input echo, cluster tables, parameter counts
 *                          (INPUT::PARSE_INFILE_LSQ, build_clusters)
 *   cg_report_frame_read   per frame read, in trajectory order: the cutoff
 *                          warnings of BOX::IS_RCUT_SAFE (the reference warns
 *                          and continues) and, for frame 0, the layering report
 *   cg_report_frame_done   per frame processed by rank 0 of an `nranks` run
 *                          (first = its first frame): the flat 3/4-body list
 *                          sizes (cg_stats_frames) and the Ewald parameters
 *   cg_report_finish       A length (3 x atoms processed), the atom counts per
 *                          type of rank 0's last frame (FITENER) and the
 *                          statistics of cg_stats_get (print_bond_stats)
 * cg_write_frame_logs writes frames.NNNN.log for each of nranks ranks
 * (src/chimes_lsq.C:283-338,718-823); frames: all frames in file order,
 * seconds: the time each took (may be NULL). */
cg_status cg_report_setup(const cg_setup *s, FILE *out);
cg_status cg_report_frame_read(const cg_setup *s, const cg_frame *f, FILE *out);
cg_status cg_report_frame_done(const cg_setup *s, const cg_frame *f, int first, int nranks, int64_t flat3,
                               int64_t flat4, FILE *out);
struct cg_lsq_stats;
cg_status cg_report_finish(const cg_setup *s, int64_t force_rows, const int *last_type_counts,
                           const struct cg_lsq_stats *stats, FILE *out);
cg_status cg_write_frame_logs(const cg_setup *s, const char *outdir, int nranks, const cg_frame *frames,
                              const double *seconds);

/* ------------------------------------------------------- device context */
typedef struct cg_ctx cg_ctx;
enum { CG_FP64 = 0, CG_FP32 = 1 };

cg_status cg_create(const cg_model *m, int device, cg_ctx **out);
void cg_destroy(cg_ctx *ctx);

/* Assemble the A rows of nframes frames (reference row order, frames in
 * order).  Output lives on the device: *A_dev points to a row-major
 * [rows x ncols] array of double (CG_FP64) or float (CG_FP32) owned by the
 * context and valid until the next cg_assemble / cg_destroy.  b is written
 * to the host array b_host (may be NULL) as doubles.  rows_out receives the
 * total row count. */
cg_status cg_assemble(cg_ctx *ctx, const cg_frame *frames, int nframes, int prec,
                      const void **A_dev, double *b_host, int64_t *rows_out);
/* Copy the last assembled A (rows [r0, r0+nrows)) to host doubles. */
cg_status cg_copy_A(cg_ctx *ctx, int64_t r0, int64_t nrows, double *host);
/* Append nrows host rows (row-major, ncols doubles each) after the last
 * assembled A, so that cg_gram / cg_fit / cg_matvec include them: the
 * FITCOUL charge-constraint rows that the reference prints after the last
 * frame (A_MAT::PRINT_CONSTRAINTS, src/A_Matrix.C:567-624). */
cg_status cg_append_rows(cg_ctx *ctx, const double *rows, int nrows);
/* Ewald parameters the charge columns of a frame use (optimal_ewald_params,
 * src/Ewald.C:428-475, accuracy 1e-6): splitting parameter alpha (1/A),
 * real-space cutoff rc (A) and integer k cutoff kmax; cell as cg_frame.cell
 * (orthorhombic: its diagonal is read). */
cg_status cg_ewald_params(int natoms, const double *cell, double *alpha, double *rc, int *kmax);

/* Tile autotuner of the 3- and 4-body assembly.  Frames are split in
 * CG_TILE_CLASSES size classes (fewer than 128, 512, 1024 atoms, and the
 * rest), each launched separately.  The given frames are assembled once per
 * tile configuration (0: 512-thread workgroups with the widest LDS tile;
 * 1: 256 threads, widest tile; 2: 256 threads, tiles half as wide so that two
 * workgroups share a CU; 3 and 4: 0 and 2 with at most 16 columns per tile),
 * each (family, class) launch is timed, and the fastest configuration of each
 * (family, class) is kept (for a family, only if those choices, launched side
 * by side, assemble faster than the default tiles; otherwise the family keeps
 * configuration 0) for later cg_assemble calls on ctx and, for the
 * life of the process, for contexts created later on a model with the same
 * species, orders and columns.  choice (may be NULL) receives
 * [3 families][CG_TILE_CLASSES] choices (2-body always 0, -1 for a class
 * without frames); ms_table (may be NULL) the
 * [CG_TILE_CONFIGS][3 families][CG_TILE_CLASSES] kernel times in ms.
 * The last assembly left in ctx is that of the kept choices.
 * No counterpart in the reference (a CPU code without tiles). */
#define CG_TILE_CONFIGS 5
#define CG_TILE_CLASSES 4
cg_status cg_autotune(cg_ctx *ctx, const cg_frame *frames, int nframes, int prec, int *choice,
                      double *ms_table);
/* The tile configuration ctx uses now, [3 families][CG_TILE_CLASSES]. */
cg_status cg_tile_choice(cg_ctx *ctx, int *choice);
/* Set it (e.g. a choice saved from cg_tile_choice); configurations
 * 0..CG_TILE_CONFIGS-1, and 0 for the 2-body family. */
cg_status cg_tile_set(cg_ctx *ctx, const int *choice);
/* Kernel timing of the last cg_assemble (ms, measured with HIP events on the
 * context stream): neighbour+enumeration, assembly kernels, total.  After a
 * call replayed from a captured graph (below) ms_enum is 0 and ms_assemble
 * is the whole call. */
cg_status cg_last_timing(cg_ctx *ctx, double *ms_enum, double *ms_assemble, double *ms_total);

/* Per-family detail of the last cg_assemble: kernel time of the 2-body,
 * 3-body and 4-body assembly launches (ms, HIP events on the context stream;
 * -1 after a graph replay, whose kernels carry no timing events) and the
 * number of pairs / triplets / quadruplets assembled. */
cg_status cg_last_detail(cg_ctx *ctx, double *ms_family3, int64_t *clusters3);

/* Graph capture of repeated assemblies (on by default).  The first
 * cg_assemble on a context reads its list sizes back; later ones (without
 * statistics) size the neighbour and cluster lists from the per-atom counts
 * seen so far and read nothing back before the end of the call, where the
 * device totals are checked (a call that did not fit is repeated with exact
 * sizes).  Such a call with the same shape as the previous call on ctx
 * (frame sizes and cells, list capacities, tile choice, precision) is
 * captured into a HIP graph, one launch for the whole assembly, and replayed
 * by later calls of that shape.  Models with FITCOUL charge columns are not
 * captured.  on = 0 keeps every call eager.  No counterpart in the reference
 * (the frame loop of src/chimes_lsq.C:292-328 assembles one frame at a
 * time). */
cg_status cg_assembly_graphs(cg_ctx *ctx, int on);
/* Bitwise-reproducible assembly (off by default).  The assembly kernels add
 * the contributions of a workgroup's waves to a shared LDS tile with FP64
 * atomics, so two runs of one batch agree to rounding (about 1e-16 relative),
 * not bit for bit.  on != 0 makes the waves add each column in a fixed order
 * (a barrier per wave and column), so that A comes out the same bits on
 * every run, at a cost in time; frames accumulated in A itself (above) stay
 * order-dependent.  The reference sums serially (src/A_Matrix.C). */
cg_status cg_assembly_deterministic(cg_ctx *ctx, int on);
/* The 3- and 4-body tile kernels of an assembly (FP64 or FP32 A) load each
 * cluster's pair tables (Chebyshev values and derivatives of its pair
 * distances) from tables computed once per neighbour entry, instead of
 * rebuilding them for every column tile (on by default).  on = 0 rebuilds
 * them per tile; both agree to rounding.  *active (may be null) gets bit 1
 * (3-body) and bit 2 (4-body) for the families whose assemblies will use
 * them: on, 3-body
 * order <= 9 with 3 or more allowed sets per column on average, or 4-body
 * order <= 5 (the hand-scheduled set loops), and every cluster slot of a
 * pair type with that type's cutoffs.  The reference
 * builds them per cluster (src/Cheby.C:885-1277, :1279-1729). */
cg_status cg_assembly_pair_tables(cg_ctx *ctx, int on, int *active);
/* Frames whose force rows do not fit the LDS tile of the assembly kernels
 * (above about 6300 atoms) are accumulated straight into A with global
 * atomics, slower but without a size limit; A is then zeroed first and the
 * call is not captured.  natoms > 0 sends every frame with more atoms than
 * that down the same path (0, the default: only those that do not fit).
 * The reference has no limit either (A_MAT is a per-frame host buffer,
 * src/A_Matrix.C:26-49). */
cg_status cg_assembly_global_above(cg_ctx *ctx, int natoms);
/* Assembly calls on ctx, those repeated with exact list sizes, those
 * replayed from a captured graph and the captures made. */
cg_status cg_assembly_stats(cg_ctx *ctx, int64_t *calls, int64_t *repeats, int64_t *graph_replays,
                            int64_t *captures);

/* Statistics chimes_lsq prints at the end of a run, gathered by the
 * following cg_assemble calls on ctx (print_bond_stats,
 * src/chimes_lsq.C:474-531; CLUSTER_LIST::print_min_distances,
 * src/Cluster.C:1569-1633): per 2-body pair type the smallest distance in its
 * neighbour list and the number of pairs with S_MINIM < r < S_MAXIM; per
 * cluster type the smallest distance of each canonical pair slot and the
 * number of clusters; per frame of the last cg_assemble the sizes of the flat
 * 3- and 4-body interaction lists it prints (NEIGHBORS::UPDATE_3B/4B_
 * INTERACTION, src/ClassDefs.C:1513-1625).  Lists are padded by 0.3 A as the
 * reference's are.  on != 0 resets and starts gathering; minima never found
 * read 1e10.  Members of cg_lsq_stats may be NULL. */
typedef struct cg_lsq_stats {
  double *pair_min;       /* [npairs]          */
  int64_t *pair_count;    /* [npairs]          */
  double *trip_min;       /* [c3.ntypes * 3]   */
  int64_t *trip_count;    /* [c3.ntypes]       */
  double *quad_min;       /* [c4.ntypes * 6]   */
  int64_t *quad_count;    /* [c4.ntypes]       */
} cg_lsq_stats;
cg_status cg_stats_enable(cg_ctx *ctx, int on);
cg_status cg_stats_get(cg_ctx *ctx, cg_lsq_stats *out);
cg_status cg_stats_frames(cg_ctx *ctx, int64_t *flat3, int64_t *flat4, int *nframes);

/* Wait for all work queued on the context's stream (no reference
 * counterpart: the reference is synchronous). */
cg_status cg_synchronize(cg_ctx *ctx);

/* Gram product G = A^T A of the last assembled A (whichever precision A
 * was assembled in).
 *   CG_GRAM_FP64   : FP64 MFMA (v_mfma_f64_16x16x4f64).  With an FP32 A the
 *                    products of the FP32 entries are exact in FP64.
 *   CG_GRAM_BF16X3 : FP32 A only; each entry split into hi + lo bfloat16,
 *                    hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16 with
 *                    FP32 accumulation per row slab, FP64 across slabs.
 *                    Relative error ~1e-6 of max|G|: too coarse to resolve
 *                    singular values below ~1e-3 sigma_max. */
enum { CG_GRAM_FP64 = 0, CG_GRAM_BF16X3 = 1 };
/* Fit mode of cg_fit without a Gram matrix: LQ factorization of A and the SVD
 * of its triangular factor (rocSOLVER), the SVD of A to working precision. */
enum { CG_FIT_QR = 2 };

/* No reference counterpart as an entry point (chimes_lsq.py forms no Gram
 * matrix); exposed so callers can run their own solvers on G and A^T b.
 * G: [ncols*ncols] row-major (symmetric), Atb: [ncols] or NULL (needs b). */
cg_status cg_gram(cg_ctx *ctx, int gram_mode, const double *b_host, double *G, double *Atb);

/* Fit: solve A x ~ b for the last assembled A (on device), the truncated
 * least-squares estimator of chimes_lsq.py (src/chimes_lsq.py:169-214):
 * x = V diag(1/sigma_i for sigma_i > eps*sigma_max) U^T b, obtained from
 * the eigen-decomposition of G = A^T A (FP64, rocSOLVER) with lambda_i kept
 * while lambda_i > eps^2 lambda_max, followed by FP64 iterative refinement
 * x += V Lambda^+ V^T A^T (b - A x) with the residual in FP64.  Refinement
 * stops early when the residual norm stops decreasing (res->iters).
 * eps: relative singular-value cut (chimes_lsq.py --eps), refine: iterations.
 * gram_mode CG_FIT_QR instead takes the SVD of A itself (A^T = L Q, then
 * L = U S V^T; x = U S+ V^T (Q b)): singular directions near the cut are then
 * resolved to ~1e-16 sigma_max / gap rather than ~1e-16 sigma_max^2 / gap^2,
 * at the cost of an FP64 copy of A (needs rows >= columns, else the FP64
 * Gram path runs); refine is unused there.
 * Outputs: x[ncols], rms (over all rows), nvars kept, eps*sigma_max. */
typedef struct cg_fit_result {
  double eps;           /* relative cut used (input)                       */
  double rms;
  double max_abs_x;
  int nvars;
  double eps_abs;       /* eps * sigma_max                                 */
  double bic;
  double ms_gram, ms_solve, ms_refine;
  int iters;            /* refinement iterations applied                   */
  int algorithm;        /* CG_ALG_SVD or CG_ALG_RIDGE (set by the fit)     */
  double alpha;         /* ridge alpha                                     */
  char weights_file[256]; /* row-weight file named in params.txt ("" none) */
} cg_fit_result;
enum { CG_ALG_SVD = 0, CG_ALG_RIDGE = 1, CG_ALG_LASSOLARS = 2 };
cg_status cg_fit(cg_ctx *ctx, const double *b_host, double eps, int refine_iters, int gram_mode,
                 double *x, cg_fit_result *res);
/* Fit only the columns keep[0..nkeep) (increasing column indices) and hold the
 * others at zero; nkeep = 0 restores fits over every column.  Used for the
 * hierarchical fit's excluded 1-body and 2-body columns, which chimes_lsq
 * leaves out of A.txt (src/A_Matrix.C:155-208). */
cg_status cg_set_fit_columns(cg_ctx *ctx, const int *keep, int nkeep);
/* Row weights of the following CG_FIT_QR fits (chimes_lsq.py --weights,
 * src/chimes_lsq.py:75-150): rows of A and b are scaled by w for the solve,
 * the RMS / BIC / A x use the unweighted rows.  nrows must equal the rows of
 * A at fit time ("Wrong number of lines in WEIGHTS file" otherwise); NULL
 * clears.  Ridge fits ignore them, as the script's ridge branch does. */
cg_status cg_set_fit_weights(cg_ctx *ctx, const double *w, int64_t nrows);
/* Ridge regression without intercept (chimes_lsq.py --algorithm=ridge, i.e.
 * sklearn Ridge(alpha, fit_intercept=False), src/chimes_lsq.py:216-225):
 * x = argmin |A x - b|^2 + alpha |x|^2, from the SVD of A (LQ factor, as
 * CG_FIT_QR): x = sum_i s_i / (s_i^2 + alpha) (u_i . b) v_i. */
cg_status cg_fit_ridge(cg_ctx *ctx, const double *b_host, double alpha, double *x, cg_fit_result *res);
/* LASSO by least-angle regression (chimes_lsq.py --algorithm=lassolars:
 * scikit-learn LassoLars(alpha, fit_intercept=False, fit_path=False,
 * max_iter), src/chimes_lsq.py:248-264): minimises
 * |A x - b|^2 / (2 rows) + alpha |x|_1 along the LARS-lasso path on the
 * device (Gram form: rows > columns, at most 2048 columns; no column subsets
 * or row weights).  res->nvars counts |x_i| > 1e-5 as the script does;
 * res->iters the path steps. */
cg_status cg_fit_lassolars(cg_ctx *ctx, const double *b_host, double alpha, int max_iter, double *x,
                           cg_fit_result *res);
/* Columns of A that a fit keeps under the setup's hierarchical exclusions
 * (EXCLUDE 1B/2B INTERACTION with HIERARC true; src/input.C:760-840): keep
 * [ncols] receives them in increasing order, nkeep their number. */
cg_status cg_setup_fit_columns(const cg_setup *s, int *keep, int *nkeep);
/* FITCOUL charge-constraint rows (CHARGE CONSTRAINTS:, src/input.C:1148-1200;
 * A_MAT::PRINT_CONSTRAINTS, src/A_Matrix.C:567-624): *nrows receives their
 * count (npairs - 1 with FITCOUL, else 0); rows (may be NULL) receives them as
 * [nrows x ncols] doubles (zeros but the Ewald columns, which hold the
 * weights in pair-index order) and b (may be NULL) their b values. */
cg_status cg_setup_constraint_rows(const cg_setup *s, double *rows, double *b, int *nrows);
/* y = A x for the last assembled A (host output). */
cg_status cg_matvec(cg_ctx *ctx, const double *x, double *y);
cg_status cg_write_params(const cg_setup *s, /* params.txt + force.txt */ const char *outdir, const double *x,
                          int ncols, int64_t nrows, const cg_fit_result *res,
                          const double *y, const double *b);

/* ------------------------------------------------------ force evaluation */
/* Force/energy/stress of one configuration with coefficients x (the fitted
 * vector in A column order).  F: [natoms*3] kcal/mol/A, E kcal/mol,
 * virial9: sum over interactions of -f r (kcal/mol), no 1/V. */
cg_status cg_force(cg_ctx *ctx, const double *x, const cg_frame *f,
                   double *F, double *E, double *virial9);

/* The same for a batch of independent configurations in one set of launches
 * (each configuration alone, as ZCalc does per MD step; src/functions.C:815-891).
 * F: [3 * sum natoms] in frame order, E: [nframes], virial9: [9 * nframes].
 * The coefficient tensors built from x stay on the device until x changes. */
cg_status cg_force_batch(cg_ctx *ctx, const double *x, const cg_frame *frames, int nframes,
                         double *F, double *E, double *virial9);

/* cg_force without any device or pinned-memory allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference: it is
 * an addition for those callers.  Every
 * device buffer of the call (coefficient tensors, frame arrays, neighbour and
 * cluster lists, sort scratch, task lists, results) is carved from one
 * fixed array of CG_NOALLOC_SCRATCH_MIB MiB per GPU that the library holds
 * until the process ends; the call uses exact list sizes (host reads of the
 * sizes), no graphs, and returns when the results are in F, E and virial9.
 * A configuration whose buffers do not fit is refused with CG_ERR_MEMORY and
 * a message naming the size it needs; the need grows with atoms x density^3
 * (measured: 4096-atom C/H/N/O configurations at 0.15-0.2 atoms/A^3 with
 * orders 12/6/3 fit, 6000 atoms at 0.22 atoms/A^3 do not).  Calls on one GPU
 * share the array:
 * calls that overlap in time on the same GPU (two threads, two contexts)
 * race and must be serialised by the caller; the array keeps the last
 * call's contents.  Same arguments and results as cg_force. */
#define CG_NOALLOC_SCRATCH_MIB 256
cg_status cg_force_noalloc(cg_ctx *ctx, const double *x, const cg_frame *f, double *F, double *E,
                           double *virial9);

/* Force calls made on ctx, how many of them were repeated with exact list
 * sizes, and how many were served by replaying a captured HIP graph.  After
 * its first call a context sizes the neighbour and cluster lists of a force
 * call from the largest per-atom counts it has seen (plus a margin) instead
 * of reading them back; a call whose lists did not fit is repeated once with
 * exact sizes.  A call of the same shape (frame sizes, cell lists, list
 * capacities) as the previous call is captured into a graph and later calls
 * of that shape replay it.  No counterpart in the reference (its lists are
 * host vectors). */
cg_status cg_force_stats(cg_ctx *ctx, int64_t *calls, int64_t *repeats, int64_t *graph_replays);

/* 2-body penalty of the force evaluation, A_p (s_min + d_p - r)^3 for
 * r < s_min + d_p (chimes_md defaults d_p = 0.01, A_p = 1e4;
 * src/chimes_md.C:1747-1748, overridden by PAIR CHEBYSHEV PENALTY lines). */
cg_status cg_set_penalty(cg_ctx *ctx, double dist, double scale);

/* Device-to-device copy of `bytes` bytes on the context stream (one kernel
 * that reads and writes exactly `bytes`): the reference of known traffic
 * used to calibrate HBM byte counters.  No reference counterpart. */
cg_status cg_copy_probe(cg_ctx *ctx, size_t bytes);

/* ---------------------------------------------------- force-evaluation server
 * No reference counterpart (the reference evaluates one configuration per MD
 * step through ZCalc, src/functions.C:815-891).  Configurations are submitted
 * from any thread; worker threads (one device context and HIP stream each)
 * take every queued request that fits a batch (max_batch_atoms atoms, and the
 * estimated device bytes of all batches in flight within hbm_budget_bytes;
 * 0 = 4 GiB) and evaluate it with cg_force_batch, so batches overlap on the
 * device and new requests are admitted as soon as a worker and budget free
 * up.  A request larger than the budget runs alone.  x is copied. */
typedef struct cg_server cg_server;
cg_status cg_serve_create(const cg_model *m, int device, const double *x, int nworkers,
                          size_t hbm_budget_bytes, int max_batch_atoms, cg_server **out);
/* Copies the frame's positions, types and cell; returns a ticket. */
cg_status cg_serve_submit(cg_server *s, const cg_frame *f, int64_t *ticket);
/* done = 1 when the result is ready (wait != 0 blocks until then); the
 * outputs are written and the ticket released on the call that sees done.
 * latency_ms: submit to completion, measured in the library.
 * A batch whose evaluation fails is evaluated again request by request, so a
 * request reports an error only when it fails on its own. */
cg_status cg_serve_poll(cg_server *s, int64_t ticket, int wait, int *done, double *F, double *E,
                        double *virial9, double *latency_ms);
/* hold != 0: workers start no new batch (requests keep queueing); hold = 0
 * releases them, and the queued requests are then admitted together under the
 * usual batch and budget limits.  Lets a caller group submissions. */
cg_status cg_serve_hold(cg_server *s, int hold);
cg_status cg_serve_stats(cg_server *s, int64_t *batches, int64_t *requests);
/* Server counters, the first n of: batches, requests, batches re-run request
 * by request after a failure, calls replayed from a captured graph, calls
 * captured into a new graph, calls repeated with exact list sizes, and
 * device or pinned buffer moves (each an allocation) made by the workers. */
cg_status cg_serve_counters(cg_server *s, int64_t *out, int n);
/* Finishes every queued request, then stops the workers. */
void cg_serve_destroy(cg_server *s);

#ifdef __cplusplus
}
#endif
#endif
