/*
 * oracle.c -- CPU FP64 restatement of the ChIMES-lsq design-matrix and
 * force-evaluation arithmetic.  TEST INFRASTRUCTURE ONLY: it is the checker
 * for the HIP path (tests/, __graft_entry__.smoke(), bench.py's cpu_baseline
 * leg).  Nothing in the shipped library links or calls it.
 *
 * It follows the reference loop for loop (paths relative to the ChIMES-lsq tree):
 *   cell geometry      BOX::UPDATE_CELL/UPDATE_INVER_CELL/LAMMPSIFY/UPDATE_EXTENT/
 *                      UPDATE_VOLUME           src/ClassDefs.C:1740-2021,2194-2211
 *   wrap + ghosts      BOX::WRAP_ATOM, LAYER_ATOM, FRAME::build_layers
 *                                              src/ClassDefs.C:2053-2163,2685-2752
 *   pair vector        BOX::GET_DISTANCE (no MIC with ghosts) src/ClassDefs.C:2321-2358
 *   neighbour lists    NEIGHBORS::DO_UPDATE_BIG / DO_UPDATE_SMALL (same sets; the
 *                      bins only accelerate)   src/ClassDefs.C:137-393
 *   transform/polys    Cheby::transform, set_polys, set_polys_out_of_range,
 *                      cheby_var_deriv        src/Cheby.C:58-165,207-443
 *   cutoff             FCUT::get_fcut, PROCEED src/Fcut.C:17-72,125-140
 *   assembly           Cheby::Deriv_2B/3B/4B  src/Cheby.C:709-1729
 *   row layout         A_MAT::PRINT_FRAME      src/A_Matrix.C:210-567
 *   Ewald columns      ZCalc_Ewald_Deriv, optimal_ewald_params, generate_trig,
 *                      add_sines              src/Ewald.C:428-825
 *   force evaluation   Cheby::Force_all/Force_3B/Force_4B src/Cheby.C:1731-2675
 *
 * Parity pins for each function are listed in tests/test_oracle.py.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../include/chimes_gpu.h"

#define PI_REF 3.14159265359 /* src/functions.h:47 */
#define PI_M 3.14159265358979323846 /* M_PI of <cmath> */

typedef struct {
  double H[9], Hi[9];           /* HMAT, INVR_HMAT (row-major as reference) */
  double ax, ay, az, bx, by, bz, cx, cy, cz;
  double la, lb, lc, alpha, beta, gamma;
  double lx, ly, lz, xy, xz, yz;
  double xlo, ylo, zlo, xhi, yhi, zhi, ex, ey, ez, vol;
  int ortho;
} box_t;

static double vmag(double x, double y, double z) { return sqrt(x * x + y * y + z * z); }
static double vangle(double x1, double y1, double z1, double x2, double y2, double z2) {
  return acos((x1 * x2 + y1 * y2 + z1 * z2) / (vmag(x1, y1, z1) * vmag(x2, y2, z2)));
}

static int box_setup(const double *cell, box_t *b) {
  b->ax = cell[0]; b->ay = cell[1]; b->az = cell[2];
  b->bx = cell[3]; b->by = cell[4]; b->bz = cell[5];
  b->cx = cell[6]; b->cy = cell[7]; b->cz = cell[8];
  double *H = b->H;
  H[0] = b->ax; H[1] = b->bx; H[2] = b->cx;
  H[3] = b->ay; H[4] = b->by; H[5] = b->cy;
  H[6] = b->az; H[7] = b->bz; H[8] = b->cz;
  b->ortho = (fabs(b->ay) + fabs(b->az) + fabs(b->bx) + fabs(b->bz) + fabs(b->cx) + fabs(b->cy) > 1.0e-10) ? 0 : 1;
  double det = H[0] * (H[4] * H[8] - H[5] * H[7]) - H[1] * (H[3] * H[8] - H[5] * H[6]) +
               H[2] * (H[3] * H[7] - H[4] * H[6]);
  if (det == 0) return -1;
  double T[9];
  T[0] = (H[4] * H[8] - H[5] * H[7]);
  T[1] = -1 * (H[3] * H[8] - H[5] * H[6]);
  T[2] = (H[3] * H[7] - H[4] * H[6]);
  T[3] = -1 * (H[1] * H[8] - H[2] * H[7]);
  T[4] = (H[0] * H[8] - H[2] * H[6]);
  T[5] = -1 * (H[0] * H[7] - H[1] * H[6]);
  T[6] = (H[1] * H[5] - H[2] * H[4]);
  T[7] = -1 * (H[0] * H[5] - H[2] * H[3]);
  T[8] = (H[0] * H[4] - H[1] * H[3]);
  double *I = b->Hi;
  I[0] = T[0]; I[1] = T[3]; I[2] = T[6];
  I[3] = T[1]; I[4] = T[4]; I[5] = T[7];
  I[6] = T[2]; I[7] = T[5]; I[8] = T[8];
  for (int i = 0; i < 9; ++i) I[i] /= det;
  b->la = vmag(b->ax, b->ay, b->az);
  b->lb = vmag(b->bx, b->by, b->bz);
  b->lc = vmag(b->cx, b->cy, b->cz);
  b->alpha = vangle(b->bx, b->by, b->bz, b->cx, b->cy, b->cz);
  b->beta = vangle(b->cx, b->cy, b->cz, b->ax, b->ay, b->az);
  b->gamma = vangle(b->ax, b->ay, b->az, b->bx, b->by, b->bz);
  /* LAMMPSIFY */
  b->lx = b->la;
  b->xy = b->lb * cos(b->gamma); if (fabs(b->xy) < 1e-12) b->xy = 0.0;
  b->xz = b->lc * cos(b->beta); if (fabs(b->xz) < 1e-12) b->xz = 0.0;
  b->ly = sqrt(b->lb * b->lb - b->xy * b->xy);
  b->yz = (b->lb * b->lc * cos(b->alpha) - b->xy * b->xz) / b->ly; if (fabs(b->yz) < 1e-12) b->yz = 0.0;
  b->lz = sqrt(b->lc * b->lc - b->xz * b->xz - b->yz * b->yz);
  /* UPDATE_EXTENT */
  b->xlo = b->ylo = b->zlo = 0.0;
  if (b->xy < b->xlo) b->xlo = b->xy;
  if (b->xz < b->xlo) b->xlo = b->xz;
  if (b->xy + b->xz < b->xlo) b->xlo = b->xy + b->xz;
  if (b->yz < b->ylo) b->ylo = b->yz;
  double t = 0.0;
  b->xhi = b->ax;
  if (b->xy > t) t = b->xy;
  if (b->xz > t) t = b->xz;
  if (b->xy + b->xz > t) t = b->xy + b->xz;
  b->xhi += t;
  b->yhi = b->by; if (b->yz > 0.0) b->yhi += b->yz;
  b->zhi = b->cz;
  b->ex = b->xhi - b->xlo; b->ey = b->yhi - b->ylo; b->ez = b->zhi - b->zlo;
  /* UPDATE_VOLUME */
  if (b->ortho) b->vol = b->ax * b->by * b->cz;
  else {
    double v = 1;
    v += 2 * cos(b->alpha) * cos(b->beta) * cos(b->gamma);
    v -= cos(b->alpha) * cos(b->alpha);
    v -= cos(b->beta) * cos(b->beta);
    v -= cos(b->gamma) * cos(b->gamma);
    b->vol = b->la * b->lb * b->lc * sqrt(v);
  }
  return 0;
}

static void mulv(const double *M, const double *x, double *y) {
  y[0] = M[0] * x[0] + M[1] * x[1] + M[2] * x[2];
  y[1] = M[3] * x[0] + M[4] * x[1] + M[5] * x[2];
  y[2] = M[6] * x[0] + M[7] * x[1] + M[8] * x[2];
}

static void wrap_atom(const box_t *b, const double *u, double *w) {
  if (b->ortho) {
    double ix = floor(u[0] / b->lx), iy = floor(u[1] / b->ly), iz = floor(u[2] / b->lz);
    w[0] = u[0] - ix * b->lx; w[1] = u[1] - iy * b->ly; w[2] = u[2] - iz * b->lz;
  } else {
    double f[3];
    mulv(b->Hi, u, f);
    f[0] -= floor(f[0]); f[1] -= floor(f[1]); f[2] -= floor(f[2]);
    mulv(b->H, f, w);
  }
}

static void layer_atom(const box_t *b, const double *r, const int *n, double *o) {
  if (b->ortho) {
    o[0] = r[0] + n[0] * b->ax; o[1] = r[1] + n[1] * b->by; o[2] = r[2] + n[2] * b->cz;
  } else {
    double f[3];
    mulv(b->Hi, r, f);
    f[0] += n[0]; f[1] += n[1]; f[2] += n[2];
    mulv(b->H, f, o);
  }
}

static void get_dist_vec(const box_t *b, const double *p1, const double *p2, double *rab) {
  double s1[3], s2[3], d[3];
  mulv(b->Hi, p1, s1);
  mulv(b->Hi, p2, s2);
  d[0] = s2[0] - s1[0]; d[1] = s2[1] - s1[1]; d[2] = s2[2] - s1[2];
  mulv(b->H, d, rab);
}

/* ---------------------------------------------------- polynomials, cutoff */
static void transform(const cg_model *m, double r, double xdiff, double xavg, double lambda, double *x, double *e) {
  switch (m->transform) {
    case CG_MORSE: *e = exp(-r / lambda); *x = (*e - xavg) / xdiff; break;
    case CG_INVRSE_R: *e = 0.0; *x = (1.0 / r - xavg) / xdiff; break;
    default: *e = 0.0; *x = (r - xavg) / xdiff; break;
  }
}

static double dxdr_of(const cg_model *m, double xdiff, double r, double lambda, double e) {
  switch (m->transform) {
    case CG_MORSE: return (-e / lambda) / xdiff;
    case CG_INVRSE_R: return -1.0 / (r * r * xdiff);
    default: return 1.0 / xdiff;
  }
}

static void polys_in_range(const cg_model *m, double *Tn, double *Tnd, double r, double xdiff, double xavg,
                           double lambda, int snum) {
  double x, e;
  transform(m, r, xdiff, xavg, lambda, &x, &e);
  Tn[0] = 1.0; Tn[1] = x;
  Tnd[0] = 1.0; Tnd[1] = 2.0 * x;
  for (int i = 2; i <= snum; ++i) {
    Tn[i] = 2.0 * x * Tn[i - 1] - Tn[i - 2];
    Tnd[i] = 2.0 * x * Tnd[i - 1] - Tnd[i - 2];
  }
  double dxdr = m->deriv_const * dxdr_of(m, xdiff, r, lambda, e);
  for (int i = snum; i >= 1; --i) Tnd[i] = i * dxdr * Tnd[i - 1];
  Tnd[0] = 0.0;
}

/* Cheby::set_polys / set_polys_out_of_range */
static void set_polys(const cg_model *m, int ptype, double *Tn, double *Tnd, double r, double xdiff, double xavg,
                      int snum, double smin) {
  double lambda = m->pair_lambda[ptype];
  if (r >= smin) { polys_in_range(m, Tn, Tnd, r, xdiff, xavg, lambda, snum); return; }
  if (m->fix_type == CG_FIX_ZERO_DERIV) {
    double x, e;
    transform(m, smin, xdiff, xavg, lambda, &x, &e);
    Tn[0] = 1.0; Tn[1] = x;
    Tnd[0] = 0.0; Tnd[1] = 0.0;
    for (int i = 2; i <= snum; ++i) { Tn[i] = 2.0 * x * Tn[i - 1] - Tn[i - 2]; Tnd[i] = 0.0; }
  } else if (m->fix_type == CG_FIX_CONSTANT_DERIV) {
    polys_in_range(m, Tn, Tnd, smin, xdiff, xavg, lambda, snum);
    for (int i = 0; i <= snum; ++i) Tn[i] += Tnd[i] * (r - smin);
  } else {
    polys_in_range(m, Tn, Tnd, smin, xdiff, xavg, lambda, snum);
    double d = m->smooth_dist;
    double f = exp((r - smin) / d);
    for (int i = 0; i <= snum; ++i) {
      Tn[i] += d * (f - 1.0) * Tnd[i];
      Tnd[i] *= f;
    }
  }
}

static void get_fcut(const cg_model *m, double *fc, double *fcd, double r, double rmax) {
  if (m->fcut_type == CG_FCUT_CUBIC) {
    double f0 = 1.0 - r / rmax;
    *fc = pow(f0, 3);
    *fcd = pow(f0, 2);
    *fcd *= -1.0 * 3 / rmax;
  } else {
    double th = rmax - m->fcut_offset * rmax;
    if (r < th) { *fc = 1.0; *fcd = 0.0; }
    else if (r > rmax) { *fc = 0.0; *fcd = 0.0; }
    else {
      double f0 = (r - th) / (rmax - th) * PI_REF + PI_REF / 2.0;
      double f0d = PI_REF / (rmax - th);
      *fc = 0.5 + 0.5 * sin(f0);
      *fcd = 0.5 * cos(f0) * f0d;
    }
  }
}

static int proceed(double r, double rmax) { return r < rmax; }

/* ----------------------------------------------------------- frame setup */
typedef struct {
  box_t box;
  int n, nall;
  double *all;     /* [nall*3] ALL_COORDS */
  int *parent;     /* [nall] */
  int *type;       /* [nall] */
  int small;
  double ps[5];    /* PERM_SCALE */
  int *l2, *l3, *l4; /* flattened lists */
  int *o2, *o3, *o4; /* offsets [n+1] */
} sys_t;

static void sys_free(sys_t *s) {
  free(s->all); free(s->parent); free(s->type);
  free(s->l2); free(s->l3); free(s->l4); free(s->o2); free(s->o3); free(s->o4);
}

static double dist(const sys_t *s, int a1, int a2, double *rab) {
  get_dist_vec(&s->box, s->all + 3 * a1, s->all + 3 * a2, rab);
  return sqrt(rab[0] * rab[0] + rab[1] * rab[1] + rab[2] * rab[2]);
}

/* Largest list cutoff of the model (NEIGHBORS::MAX_CUTOFF over the families). */
static double model_max_cutoff(const cg_model *m) {
  double mx = m->max_cutoff_2b;
  if (m->c3.nbody && m->c3.max_cutoff > mx) mx = m->c3.max_cutoff;
  if (m->c4.nbody && m->c4.max_cutoff > mx) mx = m->c4.max_cutoff;
  return mx;
}

/* Ghost layers a force evaluation needs so that every image within the list
 * cutoff exists: ceil(cutoff / smallest perpendicular cell width), at least
 * one (chimes_md run with an NLAYERS large enough for its cell). */
static int force_layers(const cg_model *m, const double *cell) {
  const double *a = cell, *b = cell + 3, *c = cell + 6;
  double bc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
  double ca[3] = {c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0]};
  double ab[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  double det = fabs(a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2]);
  double w = det / vmag(bc[0], bc[1], bc[2]);
  double w2 = det / vmag(ca[0], ca[1], ca[2]), w3 = det / vmag(ab[0], ab[1], ab[2]);
  if (w2 < w) w = w2;
  if (w3 < w) w = w3;
  int L = (int)ceil(model_max_cutoff(m) / w);
  return L < 1 ? 1 : L;
}

/* small_state: in/out sticky SMALL flag (NEIGHBORS::UPDATE_WITH_BIG).
 * layers: ghost layers (0: the model's NLAYERS). */
static int sys_build(const cg_model *m, const cg_frame *f, int *small_state, sys_t *s, int layers) {
  memset(s, 0, sizeof(*s));
  if (box_setup(f->cell, &s->box)) return -1;
  int L = layers > 0 ? layers : m->nlayers;
  int ncell = (2 * L + 1) * (2 * L + 1) * (2 * L + 1);
  s->n = f->natoms;
  s->nall = f->natoms * ncell;
  s->all = (double *)malloc(sizeof(double) * 3 * s->nall);
  s->parent = (int *)malloc(sizeof(int) * s->nall);
  s->type = (int *)malloc(sizeof(int) * s->nall);
  for (int a = 0; a < s->n; ++a) {
    wrap_atom(&s->box, f->xyz + 3 * a, s->all + 3 * a);
    s->parent[a] = a;
    s->type[a] = f->type[a];
  }
  int idx = s->n;
  for (int n1 = -L; n1 <= L; ++n1)
    for (int n2 = -L; n2 <= L; ++n2)
      for (int n3 = -L; n3 <= L; ++n3) {
        if (n1 == 0 && n2 == 0 && n3 == 0) continue;
        int nn[3] = {n1, n2, n3};
        for (int a = 0; a < s->n; ++a) {
          layer_atom(&s->box, s->all + 3 * a, nn, s->all + 3 * idx);
          s->parent[idx] = a;
          s->type[idx] = f->type[a];
          ++idx;
        }
      }
  double mc2 = m->max_cutoff_2b, mc3 = mc2, mc4 = mc2;
  if (m->c3.nbody) mc3 = m->c3.max_cutoff;
  if (m->c4.nbody) mc4 = m->c4.max_cutoff;
  double mx = mc2;
  if (mc3 > mx) mx = mc3;
  if (mc4 > mx) mx = mc4;
  if (m->small_mode || s->box.ex <= mx || s->box.ey <= mx || s->box.ez <= mx) *small_state = 1;
  s->small = *small_state;
  s->ps[0] = s->ps[1] = 1.0;
  for (int j = 2; j < 5; ++j) s->ps[j] = s->small ? s->ps[j - 1] / j : 1.0;
  if (!s->small) for (int j = 0; j < 5; ++j) s->ps[j] = 1.0;
  const double pad = 0.3;
  /* lists in ghost-index order */
  int cap = 64;
  s->l2 = (int *)malloc(sizeof(int) * cap); s->l3 = (int *)malloc(sizeof(int) * cap); s->l4 = (int *)malloc(sizeof(int) * cap);
  int c2 = 0, c3 = 0, c4 = 0, k2 = cap, k3 = cap, k4 = cap;
  s->o2 = (int *)malloc(sizeof(int) * (s->n + 1));
  s->o3 = (int *)malloc(sizeof(int) * (s->n + 1));
  s->o4 = (int *)malloc(sizeof(int) * (s->n + 1));
  for (int a1 = 0; a1 < s->n; ++a1) {
    s->o2[a1] = c2; s->o3[a1] = c3; s->o4[a1] = c4;
    for (int a2 = 0; a2 < s->nall; ++a2) {
      if (a2 == a1) continue;
      if (!s->small && a1 > s->parent[a2]) continue;
      double rab[3];
      double r = dist(s, a1, a2, rab);
#define PUSH(l, c, k) do { if (c == k) { k *= 2; l = (int *)realloc(l, sizeof(int) * k); } l[c++] = a2; } while (0)
      if (r < mc2 + pad) PUSH(s->l2, c2, k2);
      if (r < mc3 + pad) PUSH(s->l3, c3, k3);
      if (r < mc4 + pad) PUSH(s->l4, c4, k4);
#undef PUSH
    }
  }
  s->o2[s->n] = c2; s->o3[s->n] = c3; s->o4[s->n] = c4;
  return 0;
}

/* ------------------------------------------------------------- assembly */
/* Accumulators for one frame: FORCES[N][P] xyz, STRESSES[P] (6), ENERGIES[P]. */
typedef struct {
  int P;
  double *F;   /* [N*P*3] */
  double *S;   /* [P*6]  xx yy zz xy xz yz */
  double *E;   /* [P] */
} acc_t;

static void add_pair_terms(const cg_model *m, acc_t *A, int a, int b, int col, double g, const double *rab, double r) {
  for (int c = 0; c < 3; ++c) {
    double d = g * rab[c] / r;
    A->F[((size_t)a * A->P + col) * 3 + c] += d;
    A->F[((size_t)b * A->P + col) * 3 + c] -= d;
  }
  if (m->fit_stress == CG_STRESS_DIAG) {
    A->S[col * 6 + 0] -= g * rab[0] * rab[0] / r;
    A->S[col * 6 + 1] -= g * rab[1] * rab[1] / r;
    A->S[col * 6 + 2] -= g * rab[2] * rab[2] / r;
  } else if (m->fit_stress == CG_STRESS_ALL) {
    A->S[col * 6 + 0] -= g * rab[0] * rab[0] / r;
    A->S[col * 6 + 3] -= g * rab[0] * rab[1] / r;
    A->S[col * 6 + 4] -= g * rab[0] * rab[2] / r;
    A->S[col * 6 + 1] -= g * rab[1] * rab[1] / r;
    A->S[col * 6 + 5] -= g * rab[1] * rab[2] / r;
    A->S[col * 6 + 2] -= g * rab[2] * rab[2] / r;
  }
}

static int pair_index(const cg_model *m, const sys_t *s, int a1, int a2) {
  return m->int_pair_map[s->type[s->parent[a1]] * m->ntypes + s->type[s->parent[a2]]];
}

static void deriv_2b(const cg_model *m, const sys_t *s, acc_t *A, int fit_stress, int fit_ener, int md) {
  int snum = m->order2b;
  double Tn[64], Tnd[64];
  cg_model mm = *m;
  mm.fit_stress = fit_stress;
  for (int a1 = 0; a1 < s->n; ++a1)
    for (int k = s->o2[a1]; k < s->o2[a1 + 1]; ++k) {
      int a2 = s->l2[k];
      int p = pair_index(m, s, a1, a2);
      int vstart = p * snum;
      double rab[3];
      double r = dist(s, a1, a2, rab);
      /* least squares: S_MINIM < r < S_MAXIM (Deriv_2B); force evaluation: r < S_MAXIM (Force_all) */
      if (!(r < m->pair_smax[p] && (md || r > m->pair_smin[p]))) continue;
      set_polys(m, p, Tn, Tnd, r, m->pair_xdiff[p], m->pair_xavg[p], snum, m->pair_smin[p]);
      double fc, fcd;
      get_fcut(m, &fc, &fcd, r, m->pair_smax[p]);
      int b = s->parent[a2];
      for (int i = 0; i < snum; ++i) {
        double g = s->ps[2] * (fc * Tnd[i + 1] + fcd * Tn[i + 1]);
        add_pair_terms(&mm, A, a1, b, vstart + i, g, rab, r);
        if (fit_ener) A->E[vstart + i] += s->ps[2] * fc * Tn[i + 1];
      }
    }
}

static int make_id(const int *t, int n) {
  int sum = 0, w = 1;
  for (int i = 0; i < n; ++i) { sum += w * (t[i] + 1); w *= CG_MAX_ATOM_TYPES; }
  return sum;
}

/* Deriv_3B (nb=3) and Deriv_4B (nb=4) share one enumeration. */
static void deriv_nb(const cg_model *m, const cg_cluster_set *cs, const sys_t *s, acc_t *A, int fit_stress,
                     int fit_ener) {
  const int nb = cs->nbody, np = cs->npairs;
  const int *lst = nb == 3 ? s->l3 : s->l4;
  const int *off = nb == 3 ? s->o3 : s->o4;
  const double perm_scale = s->ps[nb];
  double Tn[6][32], Tnd[6][32];
  int at[4];
  cg_model mm = *m;
  mm.fit_stress = fit_stress;
  for (int a1 = 0; a1 < s->n; ++a1) {
    int b = off[a1], e = off[a1 + 1];
    for (int i2 = b; i2 < e; ++i2)
      for (int i3 = b; i3 < e; ++i3) {
        int a2 = lst[i2], a3 = lst[i3];
        if (a3 == a2) continue;
        if (perm_scale == 1.0 && s->parent[a2] > s->parent[a3]) continue;
        for (int i4 = (nb == 4 ? b : 0); i4 < (nb == 4 ? e : 1); ++i4) {
          int a4 = nb == 4 ? lst[i4] : -1;
          if (nb == 4) {
            if (a2 == a4 || a3 == a4) continue;
            if (perm_scale == 1.0 && s->parent[a3] > s->parent[a4]) continue;
          }
          int atoms[4] = {a1, a2, a3, a4};
          for (int q = 0; q < nb; ++q) at[q] = s->type[atoms[q]];
          int id = make_id(at, nb);
          int ct = cs->id_map[id];
          if (ct < 0) continue;
          const int *pmap = cs->id_pair_map + (size_t)id * np;
          double rab[6][3], r[6];
          int pa[6], pb[6];
          int q = 0;
          for (int i = 0; i < nb; ++i)
            for (int j = i + 1; j < nb; ++j) {
              r[q] = dist(s, atoms[i], atoms[j], rab[q]);
              pa[q] = i; pb[q] = j;
              ++q;
            }
          int ok = 1;
          double smin[6], smax[6], xavg[6], xdiff[6];
          for (int f = 0; f < np; ++f) {
            smax[f] = cs->s_max[ct * np + pmap[f]];
            smin[f] = cs->s_min[ct * np + pmap[f]];
            xavg[f] = cs->x_avg[ct * np + pmap[f]];
            xdiff[f] = cs->x_diff[ct * np + pmap[f]];
            if (!proceed(r[f], smax[f])) { ok = 0; break; }
          }
          if (!ok) continue;
          double fc[6], fcd[6];
          for (int f = 0; f < np; ++f) {
            int pt = pair_index(m, s, atoms[pa[f]], atoms[pb[f]]);
            set_polys(m, pt, Tn[f], Tnd[f], r[f], xdiff[f], xavg[f], cs->order, smin[f]);
          }
          for (int f = 0; f < np; ++f) get_fcut(m, &fc[f], &fcd[f], r[f], smax[f]);
          int vstart = cs->col_offset[ct];
          int parent[4];
          for (int i = 0; i < nb; ++i) parent[i] = i == 0 ? a1 : s->parent[atoms[i]];
          for (int i = cs->perm_offset[ct]; i < cs->perm_offset[ct + 1]; ++i) {
            int col = vstart + cs->perm_col[i];
            int pw[6];
            for (int f = 0; f < np; ++f) pw[f] = cs->perm_powers[(size_t)i * np + pmap[f]];
            double deriv[6], val[6];
            for (int f = 0; f < np; ++f) {
              deriv[f] = perm_scale * (fc[f] * Tnd[f][pw[f]] + fcd[f] * Tn[f][pw[f]]);
              val[f] = fc[f] * Tn[f][pw[f]];
            }
            for (int f = 0; f < np; ++f) {
              double g = deriv[f];
              for (int h = 0; h < np; ++h) if (h != f) g *= val[h];
              add_pair_terms(&mm, A, parent[pa[f]], parent[pb[f]], col, g, rab[f], r[f]);
            }
            if (fit_ener) {
              double en = perm_scale;
              for (int f = 0; f < np; ++f) en *= val[f];
              A->E[col] += en;
            }
          }
        }
      }
  }
}

/* ---------------------------------------------------------------- Ewald */
/* optimal_ewald_params (src/Ewald.C:428-475). */
static void ewald_opt(double accuracy, int nat, const box_t *b, double *alpha, double *rc, int *kc) {
  double effort_ratio = 5.5, balance_factor = 1.1, accuracy_factor = 0.8;
  double min_boxdim = b->ax;
  if (b->by < min_boxdim) min_boxdim = b->by;
  if (b->cz < min_boxdim) min_boxdim = b->cz;
  double V = b->ax * b->by * b->cz;
  double p = -log(accuracy) * accuracy_factor;
  *alpha = sqrt(PI_M) * pow(effort_ratio * nat / (V * V), 1.0 / 6.0);
  *rc = balance_factor * sqrt(p) / *alpha;
  if (*rc > 0.5 * min_boxdim) {
    *rc = 0.5 * min_boxdim;
    *alpha = sqrt(p) / *rc;
  }
  double rkc = 2.0 * *alpha * sqrt(p) / balance_factor;
  double lavg = pow(V, 1.0 / 3.0);
  *kc = (int)ceil(rkc * lavg / (2.0 * PI_M));
}

/* generate_trig (src/Ewald.C:797-825): sin/cos of k * 2 pi r / L for
 * k = 0..kmax by angle addition. */
static void gen_trig(const double *rvec, const box_t *b, int kmax, double (*S)[3], double (*C)[3]) {
  const double L[3] = {b->ax, b->by, b->cz};
  for (int d = 0; d < 3; ++d) {
    S[0][d] = 0.0;
    C[0][d] = 1.0;
    S[1][d] = sin(2.0 * PI_M * rvec[d] / L[d]);
    C[1][d] = cos(2.0 * PI_M * rvec[d] / L[d]);
  }
  for (int k = 2; k <= kmax; ++k)
    for (int d = 0; d < 3; ++d) {
      S[k][d] = S[k - 1][d] * C[1][d] + C[k - 1][d] * S[1][d];
      C[k][d] = C[k - 1][d] * C[1][d] - S[k - 1][d] * S[1][d];
    }
}

/* add_sines (src/Ewald.C:752-795): sin(kx tx + ky ty + kz tz). */
static double add_sines(int kx, int ky, int kz, double (*S)[3], double (*C)[3]) {
  int k[3] = {kx, ky, kz};
  double sk[3], ck[3];
  for (int d = 0; d < 3; ++d) {
    if (k[d] >= 0) { sk[d] = S[k[d]][d]; ck[d] = C[k[d]][d]; }
    else { sk[d] = -S[-k[d]][d]; ck[d] = C[-k[d]][d]; }
  }
  double sinyz = sk[1] * ck[2] + ck[1] * sk[2];
  double cosyz = ck[1] * ck[2] - sk[1] * sk[2];
  return sk[0] * cosyz + ck[0] * sinyz;
}

/* ZCalc_Ewald_Deriv (src/Ewald.C:477-750): per pair type p and atom a the
 * Ewald force per unit charge product, CH[(p * n + a) * 3 + c].  Real space
 * over the Ewald list (half list, or the unordered one with PERM_SCALE[2] in
 * small cells; ghost coordinates), k space over atom pairs a2 < a1 of the
 * unwrapped coordinates with the reference's k-vector enumeration. */
static int ewald_deriv(const cg_model *m, const cg_frame *f, const sys_t *s, double *CH) {
  const box_t *b = &s->box;
  if (!b->ortho) return -2;
  const int n = s->n, np = m->npairs;
  const double PI = 3.14159265359; /* src/Ewald.C:489 */
  double alpha, r_cut;
  int kmax;
  ewald_opt(1.0e-6, n, b, &alpha, &r_cut, &kmax);
  const double alphasq = alpha * alpha, Volume = b->vol;
  memset(CH, 0, sizeof(double) * (size_t)np * n * 3);
  /* k vectors (src/Ewald.C:554-598) */
  int cap = (kmax + 1) * (kmax + 1) * (kmax + 1), totk = 0;
  int *KV = (int *)malloc(sizeof(int) * 3 * cap);
  double *Kfac = (double *)malloc(sizeof(double) * cap);
  for (int kx = 0; kx <= kmax; kx++)
    for (int ky = 0; ky <= kmax; ky++)
      for (int kz = 0; kz <= kmax; kz++) {
        int ksq = kx * kx + ky * ky + kz * kz;
        if (ksq != 0 && ksq < 50) {
          double RK[3] = {(2.0 * PI / b->ax) * kx, (2.0 * PI / b->by) * ky, (2.0 * PI / b->cz) * kz};
          double rksq = RK[0] * RK[0] + RK[1] * RK[1] + RK[2] * RK[2];
          KV[3 * totk] = kx; KV[3 * totk + 1] = ky; KV[3 * totk + 2] = kz;
          Kfac[totk] = exp(-rksq / (4.0 * alphasq)) / rksq;
          totk++;
        }
      }
  const double perm_scale = s->ps[2];
  const double pad = 0.3; /* RCUT_PADDING */
  /* real space (src/Ewald.C:608-651) */
  for (int a1 = 0; a1 < n; a1++)
    for (int a2 = 0; a2 < s->nall; a2++) {
      if (a2 == a1) continue;
      if (!s->small && a1 > s->parent[a2]) continue;
      double rvec[3];
      double rlen = dist(s, a1, a2, rvec);
      if (!(rlen < r_cut + pad)) continue; /* LIST_EWALD */
      int fa2 = s->parent[a2];
      int ip = m->int_pair_map[s->type[a1] * m->ntypes + s->type[fa2]];
      if (rlen < r_cut) {
        double tempd = ((-2.0 / sqrt(PI)) * exp(-1 * alpha * alpha * rlen * rlen) * rlen * alpha - erfc(alpha * rlen)) /
                       (rlen * rlen);
        tempd *= -1.0;
        tempd *= 2.0;
        tempd *= 0.5 / rlen;
        tempd *= perm_scale;
        for (int c = 0; c < 3; ++c) {
          CH[((size_t)ip * n + a1) * 3 + c] += tempd * rvec[c];
          CH[((size_t)ip * n + fa2) * 3 + c] -= tempd * rvec[c];
        }
      }
    }
  /* k space (src/Ewald.C:653-738) */
  double (*S)[3] = (double (*)[3])malloc(sizeof(double) * 3 * (kmax + 1));
  double (*C)[3] = (double (*)[3])malloc(sizeof(double) * 3 * (kmax + 1));
  const double pre = 0.5 * (4 * PI / Volume);
  for (int a1 = 0; a1 < n; a1++)
    for (int a2 = 0; a2 < a1; a2++) {
      int ip = m->int_pair_map[f->type[a1] * m->ntypes + f->type[a2]];
      double rvec[3];
      for (int c = 0; c < 3; ++c) rvec[c] = f->xyz[3 * a2 + c] - f->xyz[3 * a1 + c];
      gen_trig(rvec, b, kmax, S, C);
      for (int ik = 0; ik < totk; ik++) {
        int kx = KV[3 * ik], ky = KV[3 * ik + 1], kz = KV[3 * ik + 2];
        double RK[3] = {(2.0 * PI / b->ax) * kx, (2.0 * PI / b->by) * ky, (2.0 * PI / b->cz) * kz};
        double t1 = add_sines(kx, ky, kz, S, C) * 2.0 * Kfac[ik] * pre, t2 = 0.0, t3 = 0.0, t4 = 0.0;
        if (ky > 0) {
          t2 = add_sines(kx, -ky, kz, S, C) * 2.0 * Kfac[ik] * pre;
          if (kz > 0) t3 = add_sines(kx, -ky, -kz, S, C) * 2.0 * Kfac[ik] * pre;
        }
        if (kz > 0) t4 = add_sines(kx, ky, -kz, S, C) * 2.0 * Kfac[ik] * pre;
        if (kx > 0) { t1 *= 2.0; t2 *= 2.0; t3 *= 2.0; t4 *= 2.0; }
        double v[3] = {(t1 + t2 + t3 + t4) * RK[0], (t1 - t2 - t3 + t4) * RK[1], (t1 + t2 - t3 - t4) * RK[2]};
        for (int c = 0; c < 3; ++c) {
          CH[((size_t)ip * n + a1) * 3 + c] += v[c];
          CH[((size_t)ip * n + a2) * 3 + c] -= v[c];
        }
      }
    }
  for (size_t i = 0; i < (size_t)np * n * 3; ++i) CH[i] *= -1.0;
  free(S); free(C); free(KV); free(Kfac);
  return 0;
}

/* chimes_md's Ewald forces, energy and virial (ZCalc_Ewald, Ewald_K_Space_New,
 * src/Ewald.C:28-426) of a model fitted with charges: per-type charges from
 * the charge products in x (src/chimes_md.C:2044-2078), the mean charge taken
 * off every atom (check_charges, src/functions.C:996-1050).  Adds to F, *e and
 * sv (xx yy zz xy xz yz).  Returns -3 for inconsistent pair charges, -4 for
 * too large a total charge, -2 for a non-orthorhombic cell. */
static int ewald_md(const cg_model *m, const cg_frame *f, const sys_t *s, const double *x, double *F, double *e,
                    double *sv) {
  const double ke = 332.0637157615209; /* src/functions.h:40 */
  const box_t *b = &s->box;
  if (!b->ortho) return -2;
  const int n = s->n, nt = m->ntypes, P = m->ncols_short;
  double qt[CG_MAX_ATOM_TYPES];
  for (int j = 0; j < nt; ++j) {
    double Q = x[P + m->int_pair_map[j * nt + j]] / ke;
    qt[j] = sqrt(fabs(Q)) * (m->charge_sign ? m->charge_sign[j] : 1);
  }
  for (int i = 0; i < m->npairs; ++i)
    if (fabs(qt[m->pair_t1[i]] * qt[m->pair_t2[i]] - x[P + i] / ke) > 1.0e-4) return -3;
  double *q = (double *)malloc(sizeof(double) * (n + 1));
  double total = 0.0;
  for (int a = 0; a < n; ++a) { q[a] = qt[f->type[a]]; total += q[a]; }
  if (total > 1.0e-3 * n) { free(q); return -4; }
  for (int a = 0; a < n; ++a) q[a] -= total / n;
  double alpha, r_cut;
  int k_cut;
  ewald_opt(1.0e-6, n, b, &alpha, &r_cut, &k_cut);
  const double alphasq = alpha * alpha, SQRT_PI = sqrt(PI_M);
  const double perm_scale = s->ps[2], pad = 0.3;
  double *T = (double *)calloc((size_t)3 * n + 3, sizeof(double));
  double U = 0.0, W[6] = {0, 0, 0, 0, 0, 0};
  /* real space over the Ewald list (src/Ewald.C:345-403) */
  for (int a1 = 0; a1 < n; a1++)
    for (int a2 = 0; a2 < s->nall; a2++) {
      if (a2 == a1) continue;
      if (!s->small && a1 > s->parent[a2]) continue;
      double R[3];
      double rlen = dist(s, a1, a2, R);
      if (!(rlen < r_cut + pad)) continue;
      int fa2 = s->parent[a2];
      if (rlen < r_cut) {
        double erfc_val = erfc(alpha * rlen);
        double tempy = erfc_val / rlen;
        double tempd = ((-2.0 / SQRT_PI) * exp(-1.0 * alphasq * rlen * rlen) * rlen * alpha - erfc_val) /
                       (rlen * rlen * rlen);
        double tempx = -tempd * q[a1] * q[fa2];
        tempx *= ke;
        tempx *= perm_scale;
        tempy *= perm_scale;
        for (int c = 0; c < 3; ++c) {
          T[3 * a1 + c] += R[c] * tempx;
          T[3 * fa2 + c] -= R[c] * tempx;
        }
        W[0] += tempx * R[0] * R[0]; W[1] += tempx * R[1] * R[1]; W[2] += tempx * R[2] * R[2];
        W[3] += tempx * R[0] * R[1]; W[4] += tempx * R[0] * R[2]; W[5] += tempx * R[1] * R[2];
        U += ke * q[a1] * q[fa2] * tempy;
      }
    }
  /* k space (src/Ewald.C:28-240) */
  const double Volume = b->ax * b->by * b->cz;
  double tempk = 0.0;
  double *sn = (double *)malloc(sizeof(double) * (n + 1)), *cs = (double *)malloc(sizeof(double) * (n + 1));
  for (int kx = -k_cut; kx <= k_cut; kx++)
    for (int ky = -k_cut; ky <= k_cut; ky++)
      for (int kz = -k_cut; kz <= k_cut; kz++) {
        int ksq = kx * kx + ky * ky + kz * kz;
        if (ksq == 0 || ksq >= k_cut * k_cut) continue;
        double K[3] = {(2.0 * PI_M / b->ax) * kx, (2.0 * PI_M / b->by) * ky, (2.0 * PI_M / b->cz) * kz};
        double rksq = K[0] * K[0] + K[1] * K[1] + K[2] * K[2];
        double kfac = exp(-rksq / (4.0 * alphasq)) / rksq;
        double tc = 0.0, ts = 0.0;
        for (int a = 0; a < n; a++) {
          double kdotr = f->xyz[3 * a] * K[0] + f->xyz[3 * a + 1] * K[1] + f->xyz[3 * a + 2] * K[2];
          sn[a] = sin(kdotr);
          cs[a] = cos(kdotr);
          tc += q[a] * cs[a];
          ts += q[a] * sn[a];
        }
        double kweight = kfac * (tc * tc + ts * ts);
        tempk += kweight;
        double pre = (2 * PI_M * ke / Volume) * kweight;
        W[0] += pre * (1.0 - 2.0 * K[0] * K[0] / rksq - 0.5 * K[0] * K[0] / alphasq);
        W[1] += pre * (1.0 - 2.0 * K[1] * K[1] / rksq - 0.5 * K[1] * K[1] / alphasq);
        W[2] += pre * (1.0 - 2.0 * K[2] * K[2] / rksq - 0.5 * K[2] * K[2] / alphasq);
        W[3] += pre * (-2.0 * K[0] * K[1] / rksq - 0.5 * K[0] * K[1] / alphasq);
        W[4] += pre * (-2.0 * K[0] * K[2] / rksq - 0.5 * K[0] * K[2] / alphasq);
        W[5] += pre * (-2.0 * K[1] * K[2] / rksq - 0.5 * K[1] * K[2] / alphasq);
        for (int a = 0; a < n; a++) {
          double tempd = -4.0 * PI_M * kfac * ke * q[a] * (tc * sn[a] - ts * cs[a]) / Volume;
          for (int c = 0; c < 3; ++c) T[3 * a + c] += K[c] * tempd;
        }
      }
  U += ke * (2 * PI_M / Volume) * tempk;
  for (int a = 0; a < n; a++) U -= ke * alpha / sqrt(PI_M) * q[a] * q[a];
  for (int i = 0; i < 3 * n; ++i) F[i] -= T[i];
  *e += U;
  for (int c = 0; c < 6; ++c) sv[c] += W[c];
  free(q); free(T); free(sn); free(cs);
  return 0;
}

/* Ewald parameters of a frame (for tests). */
int orc_ewald_params(int natoms, const double *cell, double *alpha, double *rc, int *kmax) {
  box_t b;
  if (box_setup(cell, &b)) return -1;
  ewald_opt(1.0e-6, natoms, &b, alpha, rc, kmax);
  return 0;
}

/* Rows of one frame in reference order; returns number of rows written.
 * out: [rows x ncols] row-major.  small_state is the sticky SMALL flag. */
int orc_assemble_frame(const cg_model *m, const cg_frame *f, int *small_state, double *out) {
  sys_t s;
  if (sys_build(m, f, small_state, &s, 0)) return -1;
  const int P = m->ncols_short, nc = m->ncols;
  int fit_stress = (m->fit_stress != CG_STRESS_NONE && (m->nstress < 0 || f->index < m->nstress)) ? m->fit_stress : 0;
  int fit_ener = m->fit_energy && (m->nener < 0 || f->index < m->nener);
  acc_t A;
  A.P = P;
  A.F = (double *)calloc((size_t)s.n * P * 3, sizeof(double));
  A.S = (double *)calloc((size_t)P * 6, sizeof(double));
  A.E = (double *)calloc((size_t)P, sizeof(double));
  double inv_vol = 1.0 / s.box.vol;
  deriv_2b(m, &s, &A, fit_stress, fit_ener, 0);
  for (int i = 0; i < m->n2b_cols; ++i) for (int c = 0; c < 6; ++c) A.S[i * 6 + c] *= inv_vol;
  if (m->c3.nbody) {
    deriv_nb(m, &m->c3, &s, &A, fit_stress, fit_ener);
    for (int i = 0; i < m->c3.ncols; ++i) for (int c = 0; c < 6; ++c) A.S[(m->n2b_cols + i) * 6 + c] *= inv_vol;
  }
  if (m->c4.nbody) {
    deriv_nb(m, &m->c4, &s, &A, fit_stress, fit_ener);
    int b = m->n2b_cols + m->c3.ncols;
    for (int i = 0; i < m->c4.ncols; ++i) for (int c = 0; c < 6; ++c) A.S[(b + i) * 6 + c] *= inv_vol;
  }
  int counts[CG_MAX_ATOM_TYPES] = {0};
  for (int a = 0; a < s.n; ++a) counts[f->type[a]]++;
  /* FITCOUL: Ewald columns after the short-range ones, force rows only */
  const int nq = m->n_charge_cols;
  double *CH = NULL;
  if (nq > 0) {
    CH = (double *)malloc(sizeof(double) * (size_t)m->npairs * s.n * 3 + 8);
    if (ewald_deriv(m, f, &s, CH)) {
      free(CH); free(A.F); free(A.S); free(A.E); sys_free(&s);
      return -2;
    }
  }
  int row = 0;
  for (int a = 0; a < s.n; ++a)
    for (int c = 0; c < 3; ++c) {
      double *o = out + (size_t)row * nc;
      for (int p = 0; p < P; ++p) o[p] = A.F[((size_t)a * P + p) * 3 + c];
      for (int q = 0; q < nq; ++q) o[P + q] = CH[((size_t)q * s.n + a) * 3 + c];
      for (int t = 0; t < m->n_offset_cols; ++t) o[P + nq + t] = 0.0;
      ++row;
    }
  free(CH);
  if (fit_stress) {
    static const int diag[3] = {0, 1, 2};
    static const int all[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
    int nr = fit_stress == CG_STRESS_ALL ? 9 : 3;
    const int *sel = fit_stress == CG_STRESS_ALL ? all : diag;
    for (int k = 0; k < nr; ++k) {
      double *o = out + (size_t)row * nc;
      for (int p = 0; p < P; ++p) o[p] = A.S[p * 6 + sel[k]];
      for (int q = 0; q < nq; ++q) o[P + q] = 0.0;
      for (int t = 0; t < m->n_offset_cols; ++t) o[P + nq + t] = 0.0;
      ++row;
    }
  }
  if (fit_ener) {
    for (int k = 0; k < 3; ++k) {
      double *o = out + (size_t)row * nc;
      for (int p = 0; p < P; ++p) o[p] = A.E[p];
      for (int q = 0; q < nq; ++q) o[P + q] = 0.0;
      for (int t = 0; t < m->n_offset_cols; ++t) o[P + nq + t] = counts[t];
      ++row;
    }
  }
  free(A.F); free(A.S); free(A.E);
  sys_free(&s);
  return row;
}

/* Number of (2-body pairs, 3-body, 4-body clusters) the frame contributes;
 * used by tests to cross-check enumeration counts. */
int orc_count_clusters(const cg_model *m, const cg_frame *f, int *small_state, long long *counts3) {
  sys_t s;
  if (sys_build(m, f, small_state, &s, 0)) return -1;
  counts3[0] = counts3[1] = counts3[2] = 0;
  for (int a1 = 0; a1 < s.n; ++a1)
    for (int k = s.o2[a1]; k < s.o2[a1 + 1]; ++k) {
      int a2 = s.l2[k];
      int p = pair_index(m, &s, a1, a2);
      double rab[3];
      double r = dist(&s, a1, a2, rab);
      if (r > m->pair_smin[p] && r < m->pair_smax[p]) counts3[0]++;
    }
  for (int fam = 0; fam < 2; ++fam) {
    const cg_cluster_set *cs = fam == 0 ? &m->c3 : &m->c4;
    if (!cs->nbody) continue;
    int nb = cs->nbody, np = cs->npairs;
    const int *lst = nb == 3 ? s.l3 : s.l4;
    const int *off = nb == 3 ? s.o3 : s.o4;
    double ps = s.ps[nb];
    for (int a1 = 0; a1 < s.n; ++a1) {
      int b = off[a1], e = off[a1 + 1];
      for (int i2 = b; i2 < e; ++i2)
        for (int i3 = b; i3 < e; ++i3) {
          int a2 = lst[i2], a3 = lst[i3];
          if (a3 == a2) continue;
          if (ps == 1.0 && s.parent[a2] > s.parent[a3]) continue;
          for (int i4 = (nb == 4 ? b : 0); i4 < (nb == 4 ? e : 1); ++i4) {
            int a4 = nb == 4 ? lst[i4] : -1;
            if (nb == 4 && (a2 == a4 || a3 == a4)) continue;
            if (nb == 4 && ps == 1.0 && s.parent[a3] > s.parent[a4]) continue;
            int atoms[4] = {a1, a2, a3, a4}, at[4];
            for (int q = 0; q < nb; ++q) at[q] = s.type[atoms[q]];
            int id = make_id(at, nb);
            int ct = cs->id_map[id];
            if (ct < 0) continue;
            const int *pmap = cs->id_pair_map + (size_t)id * np;
            int q = 0, ok = 1;
            for (int i = 0; i < nb && ok; ++i)
              for (int j = i + 1; j < nb; ++j) {
                double rab[3];
                double r = dist(&s, atoms[i], atoms[j], rab);
                if (!proceed(r, cs->s_max[ct * np + pmap[q]])) { ok = 0; break; }
                ++q;
              }
            if (ok) counts3[1 + fam]++;
          }
        }
    }
  }
  sys_free(&s);
  return 0;
}

/* Primitive pieces exposed for known-answer tests. */
void orc_polys(const cg_model *m, int ptype, double r, double xdiff, double xavg, int snum, double smin,
               double *Tn, double *Tnd) {
  set_polys(m, ptype, Tn, Tnd, r, xdiff, xavg, snum, smin);
}
void orc_fcut(const cg_model *m, double r, double rmax, double *fc, double *fcd) { get_fcut(m, fc, fcd, r, rmax); }
double orc_volume(const double *cell) {
  box_t b;
  if (box_setup(cell, &b)) return 0;
  return b.vol;
}

/* Force evaluation of one configuration with coefficients x (A column order):
 * Cheby::Force_all (2-body incl. r < S_MINIM and the penalty
 * A_p (s_min + d_p - r)^3), Force_3B, Force_4B.  F [3N] kcal/mol/A, E kcal/mol,
 * vir9 = -sum g r_a r_b / r (no 1/V).  penalty_scale/dist: chimes_md defaults
 * 1e4 / 0.01 (src/chimes_md.C:1747-1748). */
int orc_force(const cg_model *m, const cg_frame *f, const double *x, double penalty_scale, double penalty_dist,
              double *F, double *E, double *vir9) {
  int small = 0;
  sys_t s;
  if (sys_build(m, f, &small, &s, force_layers(m, f->cell))) return -1;
  const int P = m->ncols_short;
  acc_t A;
  A.P = P;
  A.F = (double *)calloc((size_t)s.n * P * 3, sizeof(double));
  A.S = (double *)calloc((size_t)P * 6, sizeof(double));
  A.E = (double *)calloc((size_t)P, sizeof(double));
  deriv_2b(m, &s, &A, CG_STRESS_ALL, 1, 1);
  if (m->c3.nbody) deriv_nb(m, &m->c3, &s, &A, CG_STRESS_ALL, 1);
  if (m->c4.nbody) deriv_nb(m, &m->c4, &s, &A, CG_STRESS_ALL, 1);
  for (int a = 0; a < s.n; ++a)
    for (int c = 0; c < 3; ++c) {
      double v = 0;
      for (int p = 0; p < P; ++p) v += A.F[((size_t)a * P + p) * 3 + c] * x[p];
      F[3 * a + c] = v;
    }
  double e = 0, sv[6] = {0, 0, 0, 0, 0, 0};
  for (int p = 0; p < P; ++p) {
    e += A.E[p] * x[p];
    for (int c = 0; c < 6; ++c) sv[c] += A.S[p * 6 + c] * x[p];
  }
  /* penalty (not linear in x) */
  for (int a1 = 0; a1 < s.n; ++a1)
    for (int k = s.o2[a1]; k < s.o2[a1 + 1]; ++k) {
      int a2 = s.l2[k];
      int p = pair_index(m, &s, a1, a2);
      double rab[3];
      double r = dist(&s, a1, a2, rab);
      if (!(r < m->pair_smax[p])) continue;
      if (r - penalty_dist < m->pair_smin[p]) {
        double rp = m->pair_smin[p] + penalty_dist - r;
        double coeff = -3.0 * rp * rp * penalty_scale;
        int b = s.parent[a2];
        for (int c = 0; c < 3; ++c) {
          F[3 * a1 + c] += coeff * rab[c] / r;
          F[3 * b + c] -= coeff * rab[c] / r;
        }
        e += rp * rp * rp * penalty_scale;
        sv[0] -= coeff * rab[0] * rab[0] / r; sv[1] -= coeff * rab[1] * rab[1] / r; sv[2] -= coeff * rab[2] * rab[2] / r;
        sv[3] -= coeff * rab[0] * rab[1] / r; sv[4] -= coeff * rab[0] * rab[2] / r; sv[5] -= coeff * rab[1] * rab[2] / r;
      }
    }
  /* per-atom energy offsets (ZCalc with INCLUDE_ATOM_OFFSETS, its default:
   * src/functions.C:843-852, src/input.C:2097); they are the energy-offset
   * columns of x, after the short-range ones */
  if (m->n_offset_cols > 0)
    for (int a = 0; a < f->natoms; ++a) e += x[m->ncols_short + m->n_charge_cols + f->type[a]];
  /* Ewald terms of a model fitted with charges */
  if (m->n_charge_cols > 0) {
    int rc = ewald_md(m, f, &s, x, F, &e, sv);
    if (rc) { free(A.F); free(A.S); free(A.E); sys_free(&s); return rc; }
  }
  *E = e;
  /* xx xy xz / yx yy yz / zx zy zz */
  vir9[0] = sv[0]; vir9[1] = sv[3]; vir9[2] = sv[4];
  vir9[3] = sv[3]; vir9[4] = sv[1]; vir9[5] = sv[5];
  vir9[6] = sv[4]; vir9[7] = sv[5]; vir9[8] = sv[2];
  free(A.F); free(A.S); free(A.E);
  sys_free(&s);
  return 0;
}

/* Assemble n frames with nthreads OpenMP threads (one frame per thread at a
 * time, the analogue of the reference's MPI frame split) and discard the rows;
 * used to time the CPU baseline.  Returns the number of rows produced. */
long long orc_assemble_many(const cg_model *m, const cg_frame *frames, int n, int nthreads) {
  long long rows = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(+ : rows)
  for (int i = 0; i < n; ++i) {
    int small = 0;
    const cg_frame *f = &frames[i];
    long long r = 3LL * f->natoms + 9 + 3;
    double *out = (double *)malloc(sizeof(double) * (size_t)r * m->ncols);
    int got = orc_assemble_frame(m, f, &small, out);
    free(out);
    rows += got > 0 ? got : 0;
  }
  return rows;
}
