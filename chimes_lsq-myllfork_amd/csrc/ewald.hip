// Ewald charge columns of the design matrix for FITCOUL fits
// (ZCalc_Ewald_Deriv, src/Ewald.C:477-750).
//
// For atom i and pair type p = (type i, type t) the reference sums, per unit
// charge product, the Ewald force from every atom j of type t:
//
//   A[3i+c][P+p] = sum_j  g(r_ij) (r_i - r_j)_c                     r_ij < r_c
//                + sum_j  sum_{k != 0}  (4 pi / V) Kfac(k) k_c sin(k.(r_i - r_j))
//
//   g(r) = (erfc(a r) + 2 a r exp(-a^2 r^2) / sqrt(pi)) / r^3,
//   Kfac(k) = exp(-|k|^2 / (4 a^2)) / |k|^2,
//
// with k over the integer vectors of |k_x|, |k_y|, |k_z| <= kmax and
// 0 < |k|^2 < 50 (in units of 2 pi / L per axis).  On the GPU:
//
// * real space: the cutoff r_c is at most half the shortest cell length
//   (src/Ewald.C:455-459), so at most one image of j lies inside it and the
//   minimum image is that one; the reference's ghost half list (or the
//   unordered list with PERM_SCALE 1/2 in small cells) sums the same set.
// * k space: k and -k give the same term, so half of the vectors are summed
//   twice; sin(k.(r_i - r_j)) = sin(k.r_i) cos(k.r_j) - cos(k.r_i) sin(k.r_j),
//   so the sum over j of one type is two per-type structure factors,
//   C_t(k) = sum_j cos(k.r_j) and S_t(k) = sum_j sin(k.r_j): O(N nk) work in
//   place of the reference's O(N^2 nk) pair loop.  (The j = i term cancels.)
//
// One kernel builds the structure factors (one thread per frame k-vector),
// the second gives one wavefront to each atom: lanes split the k-vectors, then
// the atoms of the frame, and a shuffle reduction sums the 3 x ntypes values.
// Parameters follow optimal_ewald_params (src/Ewald.C:428-475) per frame.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "device.hpp"
#include "setup.hpp"

namespace cg {
namespace {

constexpr int kEwMaxT = 9;  // NATMTYP <= 9
// The reference's k-vector, volume and real-space prefactors use this value of
// pi (src/Ewald.C:489); its phases use M_PI (src/Ewald.C:805-811).
constexpr double kRefPi = 3.14159265359;

struct EwFrame {
  int natoms, atom_off;
  long long row_off;
  int k0, nk;
  double alpha, rc, two_a_sqrtpi;
  double L[3], invL[3];
};

struct EwK {
  int k[3];
  int frame;
  double coef[3];  // 2 (4 pi / V) Kfac(k) k_c: the k and -k terms together
};

__global__ void k_ewald_sf(const EwFrame *__restrict__ fr, const EwK *__restrict__ ks, int nk_total,
                           const double *__restrict__ wrap, const int *__restrict__ type, int ntypes,
                           double2 *__restrict__ sf) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nk_total) return;
  const EwK K = ks[g];
  const EwFrame F = fr[K.frame];
  const double qx = K.k[0] * F.invL[0], qy = K.k[1] * F.invL[1], qz = K.k[2] * F.invL[2];
  double C[kEwMaxT], S[kEwMaxT];
#pragma unroll
  for (int t = 0; t < kEwMaxT; ++t) C[t] = S[t] = 0.0;
  for (int a = F.atom_off; a < F.atom_off + F.natoms; ++a) {
    double u = qx * wrap[3 * a] + qy * wrap[3 * a + 1] + qz * wrap[3 * a + 2];
    u -= rint(u);
    double s, c;
    sincospi(2.0 * u, &s, &c);
    const int ty = type[a];
#pragma unroll
    for (int t = 0; t < kEwMaxT; ++t) {
      const bool m = t == ty;
      C[t] += m ? c : 0.0;
      S[t] += m ? s : 0.0;
    }
  }
  for (int t = 0; t < ntypes; ++t) sf[(size_t)g * kEwMaxT + t] = make_double2(C[t], S[t]);
}

template <class T>
__global__ void __launch_bounds__(256) k_ewald_cols(const EwFrame *__restrict__ fr, const EwK *__restrict__ ks,
                                                    const double2 *__restrict__ sf, const int *__restrict__ atom_frame,
                                                    const double *__restrict__ wrap, const int *__restrict__ type,
                                                    int natoms_total, int ntypes, int npairs,
                                                    const int *__restrict__ pair_map, int col0, T *__restrict__ A,
                                                    long long ld) {
  const int lane = threadIdx.x & 63;
  const int gi = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (gi >= natoms_total) return;  // whole wavefront
  const EwFrame F = fr[atom_frame[gi]];
  const int ti = type[gi];
  const double xi = wrap[3 * gi], yi = wrap[3 * gi + 1], zi = wrap[3 * gi + 2];
  double acc[3][kEwMaxT];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int t = 0; t < kEwMaxT; ++t) acc[c][t] = 0.0;
  // k space
  for (int k = F.k0 + lane; k < F.k0 + F.nk; k += 64) {
    const EwK K = ks[k];
    double u = K.k[0] * F.invL[0] * xi + K.k[1] * F.invL[1] * yi + K.k[2] * F.invL[2] * zi;
    u -= rint(u);
    double s, c;
    sincospi(2.0 * u, &s, &c);
#pragma unroll
    for (int t = 0; t < kEwMaxT; ++t) {
      if (t < ntypes) {
        const double2 q = sf[(size_t)k * kEwMaxT + t];
        const double v = s * q.x - c * q.y;
        acc[0][t] += K.coef[0] * v;
        acc[1][t] += K.coef[1] * v;
        acc[2][t] += K.coef[2] * v;
      }
    }
  }
  // real space, minimum image
  const double rc = F.rc, al = F.alpha;
  for (int j = F.atom_off + lane; j < F.atom_off + F.natoms; j += 64) {
    if (j == gi) continue;
    double d[3] = {wrap[3 * j] - xi, wrap[3 * j + 1] - yi, wrap[3 * j + 2] - zi};
#pragma unroll
    for (int c = 0; c < 3; ++c) d[c] -= F.L[c] * rint(d[c] * F.invL[c]);
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double r = sqrt(r2);
    if (!(r < rc)) continue;
    const double ar = al * r;
    const double g = (erfc(ar) + F.two_a_sqrtpi * r * exp(-ar * ar)) / (r2 * r);
    const int tj = type[j];
#pragma unroll
    for (int t = 0; t < kEwMaxT; ++t) {
      const double m = t == tj ? g : 0.0;
      acc[0][t] -= m * d[0];
      acc[1][t] -= m * d[1];
      acc[2][t] -= m * d[2];
    }
  }
  // wavefront sums
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int t = 0; t < kEwMaxT; ++t) {
      if (t < ntypes) {
        double v = acc[c][t];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        acc[c][t] = v;
      }
    }
  // every pair column of the atom's 3 rows; a pair without the atom's type
  // gets -0.0, as the reference's final negation of its zero entries prints
  for (int idx = lane; idx < 3 * npairs; idx += 64) {
    const int c = idx / npairs, p = idx % npairs;
    double v = -0.0;
#pragma unroll
    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
      for (int t = 0; t < kEwMaxT; ++t)
        v = (cc == c && t < ntypes && pair_map[ti * ntypes + t] == p) ? acc[cc][t] : v;
    const long long row = F.row_off + 3LL * (gi - F.atom_off) + c;
    A[row * ld + col0 + p] = (T)v;
  }
}

}  // namespace

// optimal_ewald_params (src/Ewald.C:428-475) with accuracy 1e-6 (EWALD_ACCURACY,
// src/Ewald.C:6): Fincham's balance of real- and k-space work, cutoff capped
// at half the shortest cell length.
void ewald_params(int natoms, const double L[3], double *alpha, double *rc, int *kmax) {
  const double accuracy = 1.0e-6, effort_ratio = 5.5, balance = 1.1, acc_factor = 0.8;
  const double min_l = std::min(L[0], std::min(L[1], L[2]));
  const double V = L[0] * L[1] * L[2];
  const double p = -std::log(accuracy) * acc_factor;
  double a = std::sqrt(M_PI) * std::pow(effort_ratio * natoms / (V * V), 1.0 / 6.0);
  double r = balance * std::sqrt(p) / a;
  if (r > 0.5 * min_l) {
    r = 0.5 * min_l;
    a = std::sqrt(p) / r;
  }
  const double rkc = 2.0 * a * std::sqrt(p) / balance;
  *kmax = (int)std::ceil(rkc * std::pow(V, 1.0 / 3.0) / (2.0 * M_PI));
  *alpha = a;
  *rc = r;
}

void ewald_columns(Ctx &c, const cg_frame *frames, const std::vector<DevFrame> &hf, const int *d_af,
                   const int *d_type, const double *d_wrap, int natoms_total, void *A, long long ld, int prec) {
  const cg_model &m = c.m;
  const int nframes = (int)hf.size(), nt = m.ntypes;
  if (nt > kEwMaxT) throw Error(CG_ERR_UNSUPPORTED, "Ewald columns support at most 9 atom types");
  std::vector<EwFrame> ef(nframes);
  std::vector<EwK> ks;
  for (int f = 0; f < nframes; ++f) {
    const DevFrame &d = hf[f];
    if (!d.ortho) throw Error(CG_ERR_INPUT, "Ewald evaluator does not support a non-orthorhombic simulation box");
    EwFrame &e = ef[f];
    e.natoms = d.natoms;
    e.atom_off = d.atom_off;
    e.row_off = d.row_off;
    const double L[3] = {frames[f].cell[0], frames[f].cell[4], frames[f].cell[8]};
    int kmax = 0;
    ewald_params(d.natoms, L, &e.alpha, &e.rc, &kmax);
    e.two_a_sqrtpi = 2.0 * e.alpha / std::sqrt(kRefPi);
    for (int q = 0; q < 3; ++q) { e.L[q] = L[q]; e.invL[q] = 1.0 / L[q]; }
    const double V = L[0] * L[1] * L[2];
    const double a2 = e.alpha * e.alpha;
    e.k0 = (int)ks.size();
    if (d.natoms > 0)
      for (int kx = 0; kx <= kmax; ++kx)
        for (int ky = -kmax; ky <= kmax; ++ky)
          for (int kz = -kmax; kz <= kmax; ++kz) {
            const int ksq = kx * kx + ky * ky + kz * kz;
            if (ksq == 0 || ksq >= 50) continue;
            if (kx == 0 && (ky < 0 || (ky == 0 && kz < 0))) continue;  // one of k, -k
            EwK K;
            K.k[0] = kx; K.k[1] = ky; K.k[2] = kz;
            K.frame = f;
            double R[3];
            for (int q = 0; q < 3; ++q) R[q] = (2.0 * kRefPi / L[q]) * K.k[q];
            const double rksq = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
            const double kfac = std::exp(-rksq / (4.0 * a2)) / rksq;
            for (int q = 0; q < 3; ++q) K.coef[q] = 2.0 * kfac * (4.0 * kRefPi / V) * R[q];
            ks.push_back(K);
          }
    e.nk = (int)ks.size() - e.k0;
  }
  const int nk_total = (int)ks.size();
  // one upload: frames | k-vectors; the structure factors follow them
  auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t o_k = align(ef.size() * sizeof(EwFrame));
  const size_t o_sf = o_k + align(std::max<size_t>(ks.size(), 1) * sizeof(EwK));
  const size_t bytes = o_sf + (size_t)std::max(nk_total, 1) * kEwMaxT * sizeof(double2);
  char *dev = static_cast<char *>(c.b_ewald.get(bytes));
  if (!dev) throw Error(CG_ERR_MEMORY, "Ewald tables do not fit in device memory");
  c.h_ewald.assign(o_sf, 0);
  std::memcpy(c.h_ewald.data(), ef.data(), ef.size() * sizeof(EwFrame));
  if (!ks.empty()) std::memcpy(c.h_ewald.data() + o_k, ks.data(), ks.size() * sizeof(EwK));
  CG_HIP(hipMemcpyAsync(dev, c.h_ewald.data(), o_sf, hipMemcpyHostToDevice, c.stream));
  const EwFrame *d_ef = reinterpret_cast<const EwFrame *>(dev);
  const EwK *d_ks = reinterpret_cast<const EwK *>(dev + o_k);
  const int *d_pm = c.d_int_pair_map;
  double2 *d_sf = reinterpret_cast<double2 *>(dev + o_sf);
  if (nk_total > 0)
    hipLaunchKernelGGL(k_ewald_sf, dim3((nk_total + 255) / 256), dim3(256), 0, c.stream, d_ef, d_ks, nk_total, d_wrap,
                       d_type, nt, d_sf);
  if (natoms_total > 0) {
    const dim3 grid((natoms_total + 3) / 4), block(256);
    if (prec == CG_FP32)
      hipLaunchKernelGGL(k_ewald_cols<float>, grid, block, 0, c.stream, d_ef, d_ks, d_sf, d_af, d_wrap, d_type,
                         natoms_total, nt, m.npairs, d_pm, m.ncols_short, (float *)A, ld);
    else
      hipLaunchKernelGGL(k_ewald_cols<double>, grid, block, 0, c.stream, d_ef, d_ks, d_sf, d_af, d_wrap, d_type,
                         natoms_total, nt, m.npairs, d_pm, m.ncols_short, (double *)A, ld);
  }
}

// ------------------------------------------------------------------ MD side
// Ewald forces, energy and virial of configurations with charges (ZCalc_Ewald
// and Ewald_K_Space_New, src/Ewald.C:28-426), added to a force call's results.
// Differences from the fit's columns kept as the reference has them: the
// k-set is |n|^2 < kmax^2 over the cube |n_i| <= kmax, every prefactor uses
// the exact pi, and the atom charges are the per-type charges shifted by the
// configuration's mean charge (check_charges, src/functions.C:996-1050).
//   F_a  = sum_k 4 pi ke q_a Kfac(k) (C sin(k.r_a) - S cos(k.r_a)) k / V
//        + sum_b ke q_a q_b g(r_ab) (r_a - r_b)                      r_ab < r_c
//   E    = sum_pairs ke q_a q_b erfc(a r)/r + 2 pi ke / V sum_k Kfac (C^2 + S^2)
//          - ke a / sqrt(pi) sum_a q_a^2
//   W_ab = sum_pairs ke q q g r_a r_b + 2 pi ke / V sum_k Kfac (C^2 + S^2) B_ab(k),
//   B_ab = delta_ab - 2 k_a k_b / k^2 - k_a k_b / (2 a^2)      (Heyes 1994)
// with C = sum_a q_a cos(k.r_a), S = sum_a q_a sin(k.r_a); half of the k-set
// is summed twice.
namespace {

constexpr double kKe = 332.0637157615209;  // kcal/mol A per e^2 (src/functions.h:40)

struct EmFrame {
  int natoms, atom_off, k0, nk;
  double alpha, rc, two_a_sqrtpi, shift, V;
  double L[3], invL[3];
};

struct EmK {
  int k[3];
  int frame;
  double R[3];   // 2 pi n / L
  double kf2;    // 2 Kfac(k): the k and -k terms together
  double B[6];   // Heyes B tensor: xx yy zz xy xz yz
};

__global__ void k_ewmd_sf(const EmFrame *__restrict__ fr, const EmK *__restrict__ ks, int nk_total,
                          const double *__restrict__ wrap, const int *__restrict__ type, const double *__restrict__ qt,
                          double2 *__restrict__ sf, double *__restrict__ EV) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nk_total) return;
  const EmK K = ks[g];
  const EmFrame F = fr[K.frame];
  const double qx = K.k[0] * F.invL[0], qy = K.k[1] * F.invL[1], qz = K.k[2] * F.invL[2];
  double C = 0.0, S = 0.0;
  for (int a = F.atom_off; a < F.atom_off + F.natoms; ++a) {
    double u = qx * wrap[3 * a] + qy * wrap[3 * a + 1] + qz * wrap[3 * a + 2];
    u -= rint(u);
    double sn, cs;
    sincospi(2.0 * u, &sn, &cs);
    const double q = qt[type[a]] - F.shift;
    C += q * cs;
    S += q * sn;
  }
  sf[g] = make_double2(C, S);
  const double kw = (2.0 * M_PI * kKe / F.V) * K.kf2 * (C * C + S * S);
  double *ev = EV + 10 * (size_t)K.frame;
  atomicAdd(ev, kw);
  const int map9[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
  for (int q = 0; q < 9; ++q) atomicAdd(ev + 1 + q, kw * K.B[map9[q]]);
}

__global__ void __launch_bounds__(256) k_ewmd_force(const EmFrame *__restrict__ fr, const EmK *__restrict__ ks,
                                                    const double2 *__restrict__ sf, const int *__restrict__ atom_frame,
                                                    const double *__restrict__ wrap, const int *__restrict__ type,
                                                    const double *__restrict__ qt, int natoms_total,
                                                    double *__restrict__ Fout, double *__restrict__ EV) {
  const int lane = threadIdx.x & 63;
  const int gi = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (gi >= natoms_total) return;  // whole wavefront
  const int f = atom_frame[gi];
  const EmFrame F = fr[f];
  const double qa = qt[type[gi]] - F.shift;
  const double xi = wrap[3 * gi], yi = wrap[3 * gi + 1], zi = wrap[3 * gi + 2];
  double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0, w[6] = {0, 0, 0, 0, 0, 0};
  const double kpre = 4.0 * M_PI * kKe * qa / F.V;
  for (int k = F.k0 + lane; k < F.k0 + F.nk; k += 64) {
    const EmK K = ks[k];
    double u = K.k[0] * F.invL[0] * xi + K.k[1] * F.invL[1] * yi + K.k[2] * F.invL[2] * zi;
    u -= rint(u);
    double sn, cs;
    sincospi(2.0 * u, &sn, &cs);
    const double2 q = sf[k];
    const double t = kpre * K.kf2 * (q.x * sn - q.y * cs);
    fx += t * K.R[0];
    fy += t * K.R[1];
    fz += t * K.R[2];
  }
  const double rc = F.rc, al = F.alpha;
  for (int j = F.atom_off + lane; j < F.atom_off + F.natoms; j += 64) {
    if (j == gi) continue;
    double d[3] = {wrap[3 * j] - xi, wrap[3 * j + 1] - yi, wrap[3 * j + 2] - zi};
#pragma unroll
    for (int c = 0; c < 3; ++c) d[c] -= F.L[c] * rint(d[c] * F.invL[c]);
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double r = sqrt(r2);
    if (!(r < rc)) continue;
    const double ar = al * r;
    const double ec = erfc(ar);
    const double qq = kKe * qa * (qt[type[j]] - F.shift);
    const double g = qq * (ec + F.two_a_sqrtpi * r * exp(-ar * ar)) / (r2 * r);
    fx -= g * d[0];
    fy -= g * d[1];
    fz -= g * d[2];
    // each pair is seen from both atoms: half of its energy and virial here
    e += 0.5 * qq * ec / r;
    w[0] += 0.5 * g * d[0] * d[0];
    w[1] += 0.5 * g * d[1] * d[1];
    w[2] += 0.5 * g * d[2] * d[2];
    w[3] += 0.5 * g * d[0] * d[1];
    w[4] += 0.5 * g * d[0] * d[2];
    w[5] += 0.5 * g * d[1] * d[2];
  }
  double v[10] = {fx, fy, fz, e, w[0], w[1], w[2], w[3], w[4], w[5]};
#pragma unroll
  for (int q = 0; q < 10; ++q)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off, 64);
  if (lane == 0) {
    Fout[3 * (size_t)gi] += v[0];
    Fout[3 * (size_t)gi + 1] += v[1];
    Fout[3 * (size_t)gi + 2] += v[2];
    double *ev = EV + 10 * (size_t)f;
    atomicAdd(ev, v[3] - kKe * al / sqrt(M_PI) * qa * qa);  // with the self term
    const int map9[9] = {4, 7, 8, 7, 5, 9, 8, 9, 6};
    for (int q = 0; q < 9; ++q) atomicAdd(ev + 1 + q, v[map9[q]]);
  }
}

}  // namespace

// Charges per atom type from the fitted charge products (chimes_md.C:2044-2078):
// q_t = sign_t sqrt(|Q_tt| / ke); every pair product must agree within 1e-4.
void ewald_type_charges(const cg_model &m, const double *x, std::vector<double> &qt) {
  const int nt = m.ntypes;
  qt.assign(nt, 0.0);
  for (int t = 0; t < nt; ++t) {
    const int p = m.int_pair_map[t * nt + t];
    const int sg = m.charge_sign ? m.charge_sign[t] : 1;
    qt[t] = std::sqrt(std::fabs(x[m.ncols_short + p] / kKe)) * sg;
  }
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < nt; ++j) {
      const int p = m.int_pair_map[i * nt + j];
      if (std::fabs(qt[i] * qt[j] - x[m.ncols_short + p] / kKe) > 1.0e-4) {
        std::string a(m.type_names + 8 * m.pair_t1[p]), b(m.type_names + 8 * m.pair_t2[p]);
        throw Error(CG_ERR_INPUT, "Error: Inconsistent pair charges for pair: " + a + b);
      }
    }
}

void ewald_md(Ctx &c, const cg_frame *frames, const std::vector<DevFrame> &hf, const int *d_af, const int *d_type,
              const double *d_wrap, int natoms_total, double *d_F, double *d_EV) {
  const int nframes = (int)hf.size();
  const std::vector<double> &qt = c.q_type;
  std::vector<EmFrame> ef(nframes);
  std::vector<EmK> ks;
  for (int f = 0; f < nframes; ++f) {
    const DevFrame &d = hf[f];
    if (!d.ortho) throw Error(CG_ERR_INPUT, "Ewald evaluator does not support a non-orthorhombic simulation box");
    EmFrame &e = ef[f];
    e.natoms = d.natoms;
    e.atom_off = d.atom_off;
    const double L[3] = {frames[f].cell[0], frames[f].cell[4], frames[f].cell[8]};
    // neutrality (check_charges): the mean charge is taken off every atom
    double total = 0.0;
    for (int a = 0; a < d.natoms; ++a) total += qt[frames[f].type[a]];
    if (total > 1.0e-3 * d.natoms)
      throw Error(CG_ERR_INPUT, "System total charge is too large: " + std::to_string(total));
    e.shift = d.natoms > 0 ? total / d.natoms : 0.0;
    int kmax = 0;
    if (d.natoms > 0) ewald_params(d.natoms, L, &e.alpha, &e.rc, &kmax);
    e.two_a_sqrtpi = 2.0 * e.alpha / std::sqrt(M_PI);
    e.V = L[0] * L[1] * L[2];
    for (int q = 0; q < 3; ++q) { e.L[q] = L[q]; e.invL[q] = 1.0 / L[q]; }
    const double a2 = e.alpha * e.alpha;
    e.k0 = (int)ks.size();
    if (d.natoms > 0)
      for (int kx = 0; kx <= kmax; ++kx)
        for (int ky = -kmax; ky <= kmax; ++ky)
          for (int kz = -kmax; kz <= kmax; ++kz) {
            const int ksq = kx * kx + ky * ky + kz * kz;
            if (ksq == 0 || ksq >= kmax * kmax) continue;
            if (kx == 0 && (ky < 0 || (ky == 0 && kz < 0))) continue;  // one of k, -k
            EmK K;
            K.k[0] = kx; K.k[1] = ky; K.k[2] = kz;
            K.frame = f;
            for (int q = 0; q < 3; ++q) K.R[q] = (2.0 * M_PI / L[q]) * K.k[q];
            const double rksq = K.R[0] * K.R[0] + K.R[1] * K.R[1] + K.R[2] * K.R[2];
            K.kf2 = 2.0 * std::exp(-rksq / (4.0 * a2)) / rksq;
            const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
            for (int q = 0; q < 6; ++q)
              K.B[q] = (q < 3 ? 1.0 : 0.0) - 2.0 * K.R[ia[q]] * K.R[ib[q]] / rksq - 0.5 * K.R[ia[q]] * K.R[ib[q]] / a2;
            ks.push_back(K);
          }
    e.nk = (int)ks.size() - e.k0;
  }
  const int nk_total = (int)ks.size();
  auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t o_k = align(ef.size() * sizeof(EmFrame));
  const size_t o_q = o_k + align(std::max<size_t>(ks.size(), 1) * sizeof(EmK));
  const size_t o_sf = o_q + align(qt.size() * sizeof(double));
  const size_t bytes = o_sf + (size_t)std::max(nk_total, 1) * sizeof(double2);
  char *dev = static_cast<char *>(c.b_ewald.get(bytes));
  if (!dev) throw Error(CG_ERR_MEMORY, "Ewald tables do not fit in device memory");
  c.h_ewald.assign(o_sf, 0);
  std::memcpy(c.h_ewald.data(), ef.data(), ef.size() * sizeof(EmFrame));
  if (!ks.empty()) std::memcpy(c.h_ewald.data() + o_k, ks.data(), ks.size() * sizeof(EmK));
  std::memcpy(c.h_ewald.data() + o_q, qt.data(), qt.size() * sizeof(double));
  CG_HIP(hipMemcpyAsync(dev, c.h_ewald.data(), o_sf, hipMemcpyHostToDevice, c.stream));
  const EmFrame *d_ef = reinterpret_cast<const EmFrame *>(dev);
  const EmK *d_ks = reinterpret_cast<const EmK *>(dev + o_k);
  const double *d_qt = reinterpret_cast<const double *>(dev + o_q);
  double2 *d_sf = reinterpret_cast<double2 *>(dev + o_sf);
  if (nk_total > 0)
    hipLaunchKernelGGL(k_ewmd_sf, dim3((nk_total + 255) / 256), dim3(256), 0, c.stream, d_ef, d_ks, nk_total, d_wrap,
                       d_type, d_qt, d_sf, d_EV);
  if (natoms_total > 0)
    hipLaunchKernelGGL(k_ewmd_force, dim3((natoms_total + 3) / 4), dim3(256), 0, c.stream, d_ef, d_ks, d_sf, d_af,
                       d_wrap, d_type, d_qt, natoms_total, d_F, d_EV);
  CG_HIP(hipGetLastError());
}

}  // namespace cg

extern "C" cg_status cg_ewald_params(int natoms, const double *cell, double *alpha, double *rc, int *kmax) {
  if (!cell || !alpha || !rc || !kmax || natoms < 1) return CG_ERR_ARG;
  const double L[3] = {cell[0], cell[4], cell[8]};
  if (!(L[0] > 0 && L[1] > 0 && L[2] > 0)) return CG_ERR_ARG;
  cg::ewald_params(natoms, L, alpha, rc, kmax);
  return CG_OK;
}
