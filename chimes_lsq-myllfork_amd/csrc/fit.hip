// Least-squares fit of the assembled design matrix on gfx950.
//
// Replaces the SVD path of chimes_lsq.py (src/chimes_lsq.py:169-214,284-306):
//   x = V diag(1/sigma_i for sigma_i > eps*sigma_max) U^T b.
// Here the same truncated solution is obtained from the Gram matrix
// G = A^T A (eigen-decomposition G = V diag(lambda) V^T, lambda = sigma^2,
// kept while lambda_i > eps^2 lambda_max) followed by FP64 iterative
// refinement on the residual r = b - A x, which converges to the truncated
// least-squares solution in the kept subspace.
//
//   CG_GRAM_FP64   : G on FP64 MFMA (v_mfma_f64_16x16x4f64) from an FP64 or
//                    FP32 A (FP32 entries widened on load: exact products).
//   CG_GRAM_BF16X3 : FP32 A; G as split-BF16x3 (hi*hi + hi*lo + lo*hi) on
//                    v_mfma_f32_16x16x32_bf16 with FP32 accumulation per
//                    K-slab and FP64 accumulation across slabs.
// Residual and A^T r are always FP64.
//
//   CG_FIT_QR      : no Gram matrix; LQ factorization of A and SVD of its
//                    triangular factor (fit_qr below).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "device.hpp"
#include "setup.hpp"

namespace cg {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int GT = 64;      // Gram output tile (GT x GT per workgroup)
constexpr int GK = 32;      // rows staged per LDS step (FP64 path)
constexpr int GK32 = 32;    // rows staged per LDS step (BF16 path)

// Upper-triangular tile pair index -> (ti, tj) with ti <= tj.
__device__ __forceinline__ void tile_of(int p, int nt, int &ti, int &tj) {
  int i = 0;
  while (p >= nt - i) { p -= nt - i; ++i; }
  ti = i;
  tj = i + p;
}

// Partial Gram of rows [k0, k1) for one 64x64 tile (FP64 MFMA).
// v_mfma_f64_16x16x4f64 layouts: operand A lane l supplies A[l%16][l/16],
// operand B supplies B[l/16][l%16]; the result register v of lane l holds
// D[4*v + l/16][l%16] (one row per register, rows interleaved across the
// four 16-lane groups -- unlike the FP32-output 16x16 instructions, whose
// register v holds row 4*(l/16) + v).
template <class T>
__global__ __launch_bounds__(256) void k_gram_f64(const T *__restrict__ A, long long ld, long long m, int n,
                                                  long long kchunk, int nt, double *__restrict__ part) {
  __shared__ double sI[GK][GT + 1];
  __shared__ double sJ[GK][GT + 1];
  int ti, tj;
  tile_of(blockIdx.x, nt, ti, tj);
  const int i0 = ti * GT, j0 = tj * GT;
  const long long k0 = (long long)blockIdx.y * kchunk;
  const long long k1 = min(m, k0 + kchunk);
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int wr = (wave / 2) * 32, wc = (wave % 2) * 32;  // wave's 32x32 sub-tile
  d4 acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = (d4)0.0;
  // each thread stages PER elements of each panel per slab of GK rows; the
  // next slab is loaded into registers while the current one is multiplied
  // (an MFMA slab takes longer than the loads' latency)
  // (FP32 entries are widened when they go to LDS, so that the loads'
  // results are not waited for at the load)
  constexpr int PER = GK * GT / 256;
  T ri[PER], rj[PER];
  auto load = [&](long long k) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = threadIdx.x + q * 256;
      const int r = e / GT, c = e % GT;
      const long long row = k + r;
      ri[q] = (T)0;
      rj[q] = (T)0;
      if (row < k1) {
        if (i0 + c < n) ri[q] = A[row * ld + i0 + c];
        if (j0 + c < n) rj[q] = A[row * ld + j0 + c];
      }
    }
  };
  if (k0 < k1) load(k0);
  for (long long k = k0; k < k1; k += GK) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = threadIdx.x + q * 256;
      sI[e / GT][e % GT] = (double)ri[q];
      sJ[e / GT][e % GT] = (double)rj[q];
    }
    __syncthreads();
    if (k + GK < k1) load(k + GK);
#pragma unroll
    for (int kk = 0; kk < GK; kk += 4) {
      const int kr = kk + lane / 16;
      double av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) av[a] = sI[kr][wr + a * 16 + lane % 16];
#pragma unroll
      for (int b = 0; b < 2; ++b) bv[b] = sJ[kr][wc + b * 16 + lane % 16];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
  double *P = part + (size_t)blockIdx.y * n * n;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int gi = i0 + wr + a * 16 + 4 * v + lane / 16;
        int gj = j0 + wc + b * 16 + lane % 16;
        if (gi < n && gj < n) P[(size_t)gi * n + gj] = acc[a][b][v];
      }
}

// Split-BF16x3 partial Gram (A in FP32).  Operand A of
// v_mfma_f32_16x16x32_bf16: lane l supplies A[l%16][8*(l/16) .. +7]; operand B
// supplies B[8*(l/16) .. +7][l%16]; result register v of lane l holds
// D[4*(l/16) + v][l%16].
__global__ __launch_bounds__(256) void k_gram_bf16x3(const float *__restrict__ A, long long ld, long long m, int n,
                                                     long long kchunk, int nt, double *__restrict__ part) {
  __shared__ float sI[GK32][GT + 1];
  __shared__ float sJ[GK32][GT + 1];
  int ti, tj;
  tile_of(blockIdx.x, nt, ti, tj);
  const int i0 = ti * GT, j0 = tj * GT;
  const long long k0 = (long long)blockIdx.y * kchunk;
  const long long k1 = min(m, k0 + kchunk);
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int wr = (wave / 2) * 32, wc = (wave % 2) * 32;
  f4 acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = (f4)0.0f;
  for (long long k = k0; k < k1; k += GK32) {
    for (int e = threadIdx.x; e < GK32 * GT; e += 256) {
      int r = e / GT, c = e % GT;
      long long row = k + r;
      float vi = 0.f, vj = 0.f;
      if (row < k1) {
        if (i0 + c < n) vi = A[row * ld + i0 + c];
        if (j0 + c < n) vj = A[row * ld + j0 + c];
      }
      sI[r][c] = vi;
      sJ[r][c] = vj;
    }
    __syncthreads();
    const int kb = 8 * (lane / 16);
    bf8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        float v = sI[kb + q][wr + a * 16 + lane % 16];
        __bf16 h = (__bf16)v;
        ah[a][q] = h;
        al[a][q] = (__bf16)(v - (float)h);
      }
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        float v = sJ[kb + q][wc + b * 16 + lane % 16];
        __bf16 h = (__bf16)v;
        bh[b][q] = h;
        bl[b][q] = (__bf16)(v - (float)h);
      }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
      }
    __syncthreads();
  }
  double *P = part + (size_t)blockIdx.y * n * n;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int gi = i0 + wr + a * 16 + 4 * (lane / 16) + v;
        int gj = j0 + wc + b * 16 + lane % 16;
        if (gi < n && gj < n) P[(size_t)gi * n + gj] = (double)acc[a][b][v];
      }
}

// Sum the split-K partials of the upper triangle and mirror to a full
// symmetric matrix.
__global__ void k_gram_reduce(const double *part, int S, int n, double *G) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)n * n) return;
  int i = (int)(idx / n), j = (int)(idx % n);
  int a = min(i, j), b = max(i, j);
  int ta = a / GT, tb = b / GT;
  // partial of tile (ta, tb) holds row index from the I panel (ta) and column from J (tb)
  (void)ta; (void)tb;
  double s = 0.0;
  for (int q = 0; q < S; ++q) s += part[(size_t)q * n * n + (size_t)a * n + b];
  G[(size_t)i * n + j] = s;
}

// y = b - A x   (FP64 accumulation; one wave per row)
template <class T>
__global__ __launch_bounds__(256) void k_residual(const T *__restrict__ A, long long ld, long long m, int n,
                                                  const double *__restrict__ x, const double *__restrict__ b,
                                                  double *__restrict__ r) {
  long long row = (long long)blockIdx.x * 4 + threadIdx.x / 64;
  int lane = threadIdx.x % 64;
  if (row >= m) return;
  const T *a = A + row * ld;
  // four independent partial sums keep four row loads in flight per lane
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int j = lane;
  for (; j + 192 < n; j += 256) {
    s0 += (double)a[j] * x[j];
    s1 += (double)a[j + 64] * x[j + 64];
    s2 += (double)a[j + 128] * x[j + 128];
    s3 += (double)a[j + 192] * x[j + 192];
  }
  for (; j < n; j += 64) s0 += (double)a[j] * x[j];
  double s = (s0 + s1) + (s2 + s3);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) r[row] = (b ? b[row] : 0.0) - s;
}

// A^T v over a slab of rows; one thread per column, the slab's sum stored at
// part[column * nslab + slab] (summed over the slabs by k_slab_sum).
template <class T>
__global__ __launch_bounds__(256) void k_atv(const T *__restrict__ A, long long ld, long long m, int n,
                                             const double *__restrict__ v, long long slab, double *__restrict__ part) {
  int j = blockIdx.x * 256 + threadIdx.x;
  long long r0 = (long long)blockIdx.y * slab, r1 = min(m, r0 + slab);
  if (j >= n) return;
  // four rows per step with independent partial sums (four loads in flight)
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  long long r = r0;
  for (; r + 3 < r1; r += 4) {
    s0 += (double)A[r * ld + j] * v[r];
    s1 += (double)A[(r + 1) * ld + j] * v[r + 1];
    s2 += (double)A[(r + 2) * ld + j] * v[r + 2];
    s3 += (double)A[(r + 3) * ld + j] * v[r + 3];
  }
  for (; r < r1; ++r) s0 += (double)A[r * ld + j] * v[r];
  part[(size_t)j * gridDim.y + blockIdx.y] = (s0 + s1) + (s2 + s3);
}

// g[j] = sum over the slabs of part[j * nslab + slab]: one workgroup per
// column, each thread a fixed strided subset, then a fixed tree in LDS, so
// A^T v comes out the same bits on every call.
__global__ __launch_bounds__(256) void k_slab_sum(const double *__restrict__ part, int nslab, double *__restrict__ g) {
  __shared__ double red[256];
  const double *p = part + (size_t)blockIdx.x * nslab;
  double s = 0.0;
  for (int k = threadIdx.x; k < nslab; k += 256) s += p[k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) g[blockIdx.x] = red[0];
}

__global__ void k_scale(double *v, const double *w, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] *= w[i];
}
// Kept rows/columns of the Gram (column-major) and kept entries of a vector
// (fits over a column subset: hierarchical exclusions).
__global__ void k_gather_sym(const double *G, int n, const int *keep, int nk, double *Gk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i < nk && j < nk) Gk[(size_t)j * nk + i] = G[(size_t)keep[j] * n + keep[i]];
}
__global__ void k_gather(const double *v, const int *keep, int nk, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nk) out[i] = v[keep[i]];
}
__global__ void k_zero(double *out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = 0.0;
}
__global__ void k_scatter_vals(const double *v, const int *keep, int nk, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nk) out[keep[i]] = v[i];
}

// FP64 copy of the fitted columns of A: row-major m x ne, i.e. the
// column-major ne x m matrix A^T that the LQ factorization takes.
template <class T>
__global__ void k_copy_cols(const T *__restrict__ A, long long ld, long long m, int ne, const int *__restrict__ keep,
                            const double *__restrict__ w, double *__restrict__ W) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= m * ne) return;
  const long long r = idx / ne;
  const int j = (int)(idx - r * ne);
  const double v = (double)A[r * ld + (keep ? keep[j] : j)];
  W[idx] = w ? w[r] * v : v;
}
__global__ void k_mul(double *v, const double *w, long long m) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < m) v[i] *= w[i];
}
// Lower triangle of the leading ne x ne block of a column-major matrix.
__global__ void k_lower(const double *W, int ne, double *L) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i < ne) L[(size_t)j * ne + i] = i >= j ? W[(size_t)j * ne + i] : 0.0;
}

__global__ void k_axpy(double *x, const double *d, int n, double alpha) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += alpha * d[i];
}

struct FitState {
  rocblas_handle h = nullptr;
};

static FitState &fit_state(Ctx &c) {
  static thread_local FitState fs;
  if (!fs.h) {
    if (rocblas_create_handle(&fs.h) != rocblas_status_success) throw Error(CG_ERR_DEVICE, "rocblas_create_handle");
    // no atomics in rocBLAS (and in the rocSOLVER routines that call it): a
    // fit of the same A gives the same bits on every call
    if (rocblas_set_atomics_mode(fs.h, rocblas_atomics_not_allowed) != rocblas_status_success)
      throw Error(CG_ERR_DEVICE, "rocblas_set_atomics_mode");
  }
  rocblas_set_stream(fs.h, c.stream);
  return fs;
}

template <class T>
static void atv(Ctx &c, const T *A, long long m, int n, const double *v, double *g) {
  if (m <= 0 || n <= 0) {
    if (n > 0) CG_HIP(hipMemsetAsync(g, 0, n * sizeof(double), c.stream));
    return;
  }
  // slabs small enough for several workgroups per CU (at most 4096 of them,
  // which bounds the partial-sum buffer); their partial sums are added in slab
  // order (no atomics: the same bits on every call)
  long long slab = std::max(256LL, (m + 4095) / 4096);
  long long nslab = (m + slab - 1) / slab;
  double *part = c.b_atvp.as<double>((size_t)nslab * n);
  if (!part) throw Error(CG_ERR_MEMORY, "A^T v partial sums");
  hipLaunchKernelGGL(k_atv<T>, dim3((n + 255) / 256, (unsigned)nslab), dim3(256), 0, c.stream, A, (long long)c.A_ld,
                     m, n, v, slab, part);
  hipLaunchKernelGGL(k_slab_sum, dim3((unsigned)n), dim3(256), 0, c.stream, part, (int)nslab, g);
}

template <class T>
static void residual(Ctx &c, const T *A, long long m, int n, const double *x, const double *b, double *r) {
  hipLaunchKernelGGL(k_residual<T>, dim3((unsigned)((m + 3) / 4)), dim3(256), 0, c.stream, A, (long long)c.A_ld, m, n,
                     x, b, r);
}

// G = A^T A into Gd [n*n] (row-major, symmetric); part: S partial slabs.
static void gram(Ctx &c, int mode, double *Gd, double *part, long long S, long long kchunk) {
  const long long m = c.A_rows;
  const int n = c.A_ld;
  const int nt = (n + GT - 1) / GT;
  const int npair = nt * (nt + 1) / 2;
  if (mode == CG_GRAM_BF16X3) {
    if (c.A_prec != CG_FP32) throw Error(CG_ERR_ARG, "the split-BF16x3 Gram needs an FP32 A");
    hipLaunchKernelGGL(k_gram_bf16x3, dim3(npair, (unsigned)S), dim3(256), 0, c.stream, (const float *)c.A,
                       (long long)n, m, n, kchunk, nt, part);
  } else if (c.A_prec == CG_FP64) {
    hipLaunchKernelGGL(k_gram_f64<double>, dim3(npair, (unsigned)S), dim3(256), 0, c.stream, (const double *)c.A,
                       (long long)n, m, n, kchunk, nt, part);
  } else {
    hipLaunchKernelGGL(k_gram_f64<float>, dim3(npair, (unsigned)S), dim3(256), 0, c.stream, (const float *)c.A,
                       (long long)n, m, n, kchunk, nt, part);
  }
  CG_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_gram_reduce, dim3((unsigned)(((long long)n * n + 255) / 256)), dim3(256), 0, c.stream, part,
                     (int)S, n, Gd);
  CG_HIP(hipGetLastError());
}

// Split-K plan of the Gram launch: S row slabs of kchunk rows (multiple of
// the staging depth), enough workgroups to fill the chip.
static void gram_plan(long long m, int n, long long &S, long long &kchunk) {
  const int nt = (n + GT - 1) / GT;
  const int npair = nt * (nt + 1) / 2;
  long long target = std::max<long long>(1, (long long)(1024 + npair - 1) / npair);
  S = std::min<long long>(target, std::max<long long>(1, m / 256));
  S = std::min<long long>(S, 64);
  kchunk = (m + S - 1) / S;
  kchunk = (kchunk + GK32 - 1) / GK32 * GK32;
  S = std::max<long long>(1, (m + kchunk - 1) / kchunk);
}

static void check_mode(int mode) {
  if (mode != CG_GRAM_FP64 && mode != CG_GRAM_BF16X3 && mode != CG_FIT_QR) throw Error(CG_ERR_ARG, "unknown fit mode");
}

// CG_FIT_QR: the truncated SVD solution of chimes_lsq.py from an SVD of A
// itself, to working precision (no Gram matrix, whose eigenvectors resolve
// singular directions near the cut eps * s_max only to about
// 1e-16 s_max^2 / (s_i^2 - s_j^2)).  The row-major A is the column-major
// ne x m matrix A^T = L Q (LQ factorization, rocsolver_dgelqf), so
// A = Q^T L^T; with the SVD L = U S V^T (rocsolver_dgesvd of the ne x ne
// factor) A = (Q^T V) S U^T and
//   x = sum_{s_i > eps s_max} U[:, i] (V^T (Q b))_i / s_i,
// Q b applied from the stored reflectors (rocsolver_dormlq).
//
// Row weights (chimes_lsq.py --weights, src/chimes_lsq.py:75-150) scale the
// rows of A and b for the solve; the RMS, BIC and A x use the unweighted
// rows, as the script does.  ridge >= 0 gives sklearn's Ridge without
// intercept (src/chimes_lsq.py:216-225) from the same SVD:
//   x = sum_i s_i / (s_i^2 + ridge) (V^T (Q b))_i U[:, i], every column kept.
static void fit_qr(Ctx &c, const double *b_host, double eps, double ridge, double *x_host, cg_fit_result *res) {
  const long long m = c.A_rows;
  const int n = c.A_ld;
  FitState &fs = fit_state(c);
  const bool masked = !c.fit_keep.empty() && (int)c.fit_keep.size() < n;
  const int ne = masked ? (int)c.fit_keep.size() : n;
  const size_t nn = (size_t)ne * ne;
  double *W = c.b_qr.as<double>((size_t)m * ne + 3 * (size_t)m + 3 * nn + 6 * (size_t)ne + (size_t)n + 64);
  if (!W) throw Error(CG_ERR_MEMORY, "QR fit buffers");
  double *d_b = W + (size_t)m * ne, *d_q = d_b + m, *d_r = d_q + m;
  double *L = d_r + m, *U = L + nn, *VT = U + nn;
  double *S = VT + nn, *E = S + ne, *tau = E + ne, *t = tau + ne, *winv = t + ne, *xk = winv + ne, *d_x = xk + ne;
  rocblas_int *d_info = reinterpret_cast<rocblas_int *>(d_x + n);
  double *d_z = d_x + n + 8;
  int *d_keep = nullptr;
  if (masked) {
    d_keep = c.b_fitkeep.as<int>(ne);
    if (!d_keep) throw Error(CG_ERR_MEMORY, "fit column list");
    CG_HIP(hipMemcpyAsync(d_keep, c.fit_keep.data(), ne * sizeof(int), hipMemcpyHostToDevice, c.stream));
  }
  hipEvent_t e0, e1, e2;
  CG_HIP(hipEventCreate(&e0)); CG_HIP(hipEventCreate(&e1)); CG_HIP(hipEventCreate(&e2));
  const bool weighted = !c.fit_w.empty() && ridge < 0.0;
  double *d_w = nullptr;
  if (!c.fit_w.empty() && (long long)c.fit_w.size() != m)
    throw Error(CG_ERR_ARG, "Wrong number of lines in WEIGHTS file");
  if (weighted) {
    d_w = c.b_fitw.as<double>(m);
    if (!d_w) throw Error(CG_ERR_MEMORY, "fit weights");
    CG_HIP(hipMemcpyAsync(d_w, c.fit_w.data(), m * sizeof(double), hipMemcpyHostToDevice, c.stream));
  }
  CG_HIP(hipMemcpyAsync(d_b, b_host, m * sizeof(double), hipMemcpyHostToDevice, c.stream));
  CG_HIP(hipMemcpyAsync(d_q, d_b, m * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  if (weighted) hipLaunchKernelGGL(k_mul, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c.stream, d_q, d_w, m);
  CG_HIP(hipEventRecord(e0, c.stream));
  const long long tot = m * ne;
  const unsigned grid = (unsigned)((tot + 255) / 256);
  if (c.A_prec == CG_FP64)
    hipLaunchKernelGGL(k_copy_cols<double>, dim3(grid), dim3(256), 0, c.stream, (const double *)c.A, (long long)n, m, ne,
                       (const int *)d_keep, d_w, W);
  else
    hipLaunchKernelGGL(k_copy_cols<float>, dim3(grid), dim3(256), 0, c.stream, (const float *)c.A, (long long)n, m, ne,
                       (const int *)d_keep, d_w, W);
  CG_HIP(hipGetLastError());
  if (rocsolver_dgelqf(fs.h, ne, (rocblas_int)m, W, ne, tau) != rocblas_status_success)
    throw Error(CG_ERR_NUMERIC, "rocsolver_dgelqf failed");
  if (rocsolver_dormlq(fs.h, rocblas_side_left, rocblas_operation_none, (rocblas_int)m, 1, ne, W, ne, tau, d_q,
                       (rocblas_int)m) != rocblas_status_success)
    throw Error(CG_ERR_NUMERIC, "rocsolver_dormlq failed");
  hipLaunchKernelGGL(k_lower, dim3((ne + 255) / 256, ne), dim3(256), 0, c.stream, W, ne, L);
  CG_HIP(hipEventRecord(e1, c.stream));
  if (rocsolver_dgesvd(fs.h, rocblas_svect_all, rocblas_svect_all, ne, ne, L, ne, S, U, ne, VT, ne, E,
                       rocblas_outofplace, d_info) != rocblas_status_success)
    throw Error(CG_ERR_NUMERIC, "rocsolver_dgesvd failed");
  std::vector<double> sv(ne);
  int info = 0;
  CG_HIP(hipMemcpyAsync(sv.data(), S, ne * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  CG_HIP(hipMemcpyAsync(&info, d_info, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  CG_HIP(hipStreamSynchronize(c.stream));
  if (info != 0) throw Error(CG_ERR_NUMERIC, "SVD did not converge (info=" + std::to_string(info) + ")");
  double smax = 0;
  for (double v : sv) smax = std::max(smax, std::fabs(v));
  const double thr = eps * smax;
  std::vector<double> wi(ne, 0.0);
  int nvars = 0;
  if (ridge >= 0.0) {
    for (int i = 0; i < ne; ++i) wi[i] = sv[i] / (sv[i] * sv[i] + ridge);
    nvars = ne;
  } else {
    for (int i = 0; i < ne; ++i)
      if (std::fabs(sv[i]) > thr) { wi[i] = 1.0 / sv[i]; ++nvars; }
  }
  CG_HIP(hipMemcpyAsync(winv, wi.data(), ne * sizeof(double), hipMemcpyHostToDevice, c.stream));
  // t = S+ V^T (Q b)[0:ne];  x = U t
  const double one = 1.0, zero = 0.0;
  if (rocblas_dgemv(fs.h, rocblas_operation_none, ne, ne, &one, VT, ne, d_q, 1, &zero, t, 1) != rocblas_status_success)
    throw Error(CG_ERR_DEVICE, "rocblas_dgemv");
  hipLaunchKernelGGL(k_scale, dim3((ne + 255) / 256), dim3(256), 0, c.stream, t, winv, ne);
  if (rocblas_dgemv(fs.h, rocblas_operation_none, ne, ne, &one, U, ne, t, 1, &zero, xk, 1) != rocblas_status_success)
    throw Error(CG_ERR_DEVICE, "rocblas_dgemv");
  if (masked) {
    hipLaunchKernelGGL(k_zero, dim3((n + 255) / 256), dim3(256), 0, c.stream, d_x, n);
    hipLaunchKernelGGL(k_scatter_vals, dim3((ne + 255) / 256), dim3(256), 0, c.stream, xk, (const int *)d_keep, ne, d_x);
  } else {
    CG_HIP(hipMemcpyAsync(d_x, xk, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  }
  CG_HIP(hipEventRecord(e2, c.stream));
  // residual over every row (the reference's RMS and BIC use b - A x)
  if (c.A_prec == CG_FP64) residual(c, (const double *)c.A, m, n, d_x, d_b, d_r);
  else residual(c, (const float *)c.A, m, n, d_x, d_b, d_r);
  double Z = 0;
  if (rocblas_set_pointer_mode(fs.h, rocblas_pointer_mode_device) != rocblas_status_success ||
      rocblas_ddot(fs.h, (rocblas_int)m, d_r, 1, d_r, 1, d_z) != rocblas_status_success)
    throw Error(CG_ERR_DEVICE, "rocblas_ddot");
  rocblas_set_pointer_mode(fs.h, rocblas_pointer_mode_host);
  CG_HIP(hipMemcpyAsync(&Z, d_z, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  CG_HIP(hipMemcpyAsync(x_host, d_x, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  CG_HIP(hipStreamSynchronize(c.stream));
  double mx = 0;
  for (int i = 0; i < n; ++i) mx = std::max(mx, std::fabs(x_host[i]));
  float tq = 0, ts = 0;
  CG_HIP(hipEventElapsedTime(&tq, e0, e1));
  CG_HIP(hipEventElapsedTime(&ts, e1, e2));
  for (auto e : {e0, e1, e2}) (void)hipEventDestroy(e);
  res->eps = eps;
  res->rms = std::sqrt(Z / (double)m);
  res->max_abs_x = mx;
  res->nvars = nvars;
  res->eps_abs = thr;
  res->bic = (double)m * std::log(Z / (double)m) + (double)nvars * std::log((double)m);
  res->ms_gram = tq;
  res->ms_solve = ts;
  res->ms_refine = 0.f;
  res->iters = 0;
  res->algorithm = ridge >= 0.0 ? CG_ALG_RIDGE : CG_ALG_SVD;
  res->alpha = ridge >= 0.0 ? ridge : 0.0;
}

void gram_host(Ctx &c, int mode, const double *b_host, double *G_host, double *Atb_host) {
  if (!c.A) throw Error(CG_ERR_ARG, "nothing assembled");
  check_mode(mode);
  if (mode == CG_FIT_QR) throw Error(CG_ERR_ARG, "cg_gram takes a Gram mode");
  const long long m = c.A_rows;
  const int n = c.A_ld;
  long long S, kchunk;
  gram_plan(m, n, S, kchunk);
  double *G = c.b_x.as<double>((size_t)n * n + (size_t)S * n * n + (size_t)m + (size_t)n + 64);
  if (!G) throw Error(CG_ERR_MEMORY, "Gram buffers");
  double *part = G + (size_t)n * n;
  double *d_b = part + (size_t)S * n * n;
  double *d_c = d_b + m;
  gram(c, mode, G, part, S, kchunk);
  if (G_host) CG_HIP(hipMemcpyAsync(G_host, G, (size_t)n * n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  if (Atb_host) {
    if (!b_host) throw Error(CG_ERR_ARG, "A^T b needs b");
    CG_HIP(hipMemcpyAsync(d_b, b_host, m * sizeof(double), hipMemcpyHostToDevice, c.stream));
    if (c.A_prec == CG_FP64) atv(c, (const double *)c.A, m, n, d_b, d_c);
    else atv(c, (const float *)c.A, m, n, d_b, d_c);
    CG_HIP(hipMemcpyAsync(Atb_host, d_c, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  }
  CG_HIP(hipStreamSynchronize(c.stream));
}

void fit(Ctx &c, const double *b_host, double eps, int refine, int mode, double *x_host, cg_fit_result *res) {
  if (!c.A) throw Error(CG_ERR_ARG, "nothing assembled");
  check_mode(mode);
  if (!(eps > 0.0)) throw Error(CG_ERR_ARG, "eps must be positive");
  const long long m = c.A_rows;
  const int n = c.A_ld;
  const int nfit = !c.fit_keep.empty() && (int)c.fit_keep.size() < n ? (int)c.fit_keep.size() : n;
  if (mode == CG_FIT_QR) {
    if (m >= nfit) return fit_qr(c, b_host, eps, -1.0, x_host, res);
    mode = CG_GRAM_FP64;  // fewer rows than columns: the Gram path
  }
  if (!c.fit_w.empty()) throw Error(CG_ERR_UNSUPPORTED, "row weights need the CG_FIT_QR fit");
  FitState &fs = fit_state(c);
  hipEvent_t e0, e1, e2, e3;
  CG_HIP(hipEventCreate(&e0)); CG_HIP(hipEventCreate(&e1)); CG_HIP(hipEventCreate(&e2)); CG_HIP(hipEventCreate(&e3));
  // buffers: G [n*n], part [S*n*n], b [m], r [m], x, c, t, dx, w, W, E [n], info
  long long S, kchunk;
  gram_plan(m, n, S, kchunk);
  double *G = c.b_x.as<double>((size_t)n * n + (size_t)S * n * n + 2 * (size_t)m + 10 * (size_t)n + 128);
  if (!G) throw Error(CG_ERR_MEMORY, "fit buffers");
  double *part = G + (size_t)n * n;
  double *d_b = part + (size_t)S * n * n;
  double *d_r = d_b + m;
  double *d_x = d_r + m;
  double *d_c = d_x + n, *d_t = d_c + n, *d_dx = d_t + n, *d_w = d_dx + n, *d_W = d_w + n, *d_E = d_W + n;
  CG_HIP(hipMemcpyAsync(d_b, b_host, m * sizeof(double), hipMemcpyHostToDevice, c.stream));
  CG_HIP(hipEventRecord(e0, c.stream));
  gram(c, mode, G, part, S, kchunk);
  // a fit over a column subset works on the kept rows/columns of G (gathered
  // into the Gram's split-K scratch, free after the reduction) and keeps the
  // other coefficients at zero
  const bool masked = !c.fit_keep.empty() && (int)c.fit_keep.size() < n;
  const int ne = masked ? (int)c.fit_keep.size() : n;
  double *Ge = G;
  double *d_g = d_E + n + 16, *d_g2 = d_g + n;
  int *d_keep = nullptr;
  if (masked) {
    d_keep = c.b_fitkeep.as<int>(ne);
    if (!d_keep) throw Error(CG_ERR_MEMORY, "fit column list");
    CG_HIP(hipMemcpyAsync(d_keep, c.fit_keep.data(), ne * sizeof(int), hipMemcpyHostToDevice, c.stream));
    Ge = part;
    hipLaunchKernelGGL(k_gather_sym, dim3((ne + 255) / 256, ne), dim3(256), 0, c.stream, G, n, (const int *)d_keep,
                       ne, Ge);
  }
  CG_HIP(hipEventRecord(e1, c.stream));
  // eigen-decomposition G = V diag(W) V^T (V column-major in G; G is symmetric)
  rocblas_int *d_info = reinterpret_cast<rocblas_int *>(d_E + n);
  if (rocsolver_dsyevd(fs.h, rocblas_evect_original, rocblas_fill_upper, ne, Ge, ne, d_W, d_E, d_info) !=
      rocblas_status_success)
    throw Error(CG_ERR_NUMERIC, "rocsolver_dsyevd failed");
  std::vector<double> W(ne);
  int info = 0;
  CG_HIP(hipMemcpyAsync(W.data(), d_W, ne * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  CG_HIP(hipMemcpyAsync(&info, d_info, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  CG_HIP(hipStreamSynchronize(c.stream));
  if (info != 0) throw Error(CG_ERR_NUMERIC, "eigensolver did not converge (info=" + std::to_string(info) + ")");
  double lmax = 0;
  for (double v : W) lmax = std::max(lmax, std::fabs(v));
  const double thr = eps * eps * lmax;
  std::vector<double> winv(ne, 0.0);
  int nvars = 0;
  for (int i = 0; i < ne; ++i)
    if (W[i] > thr) { winv[i] = 1.0 / W[i]; ++nvars; }
  CG_HIP(hipMemcpyAsync(d_w, winv.data(), ne * sizeof(double), hipMemcpyHostToDevice, c.stream));
  // P v = V W+ V^T v (on the kept entries; zero elsewhere)
  auto project = [&](const double *in, double *out) {
    const double one = 1.0, zero = 0.0;
    const double *src = in;
    double *dst = out;
    if (masked) {
      hipLaunchKernelGGL(k_gather, dim3((ne + 255) / 256), dim3(256), 0, c.stream, in, (const int *)d_keep, ne, d_g);
      src = d_g;
      dst = d_g2;
    }
    if (rocblas_dgemv(fs.h, rocblas_operation_transpose, ne, ne, &one, Ge, ne, src, 1, &zero, d_t, 1) !=
        rocblas_status_success)
      throw Error(CG_ERR_DEVICE, "rocblas_dgemv");
    hipLaunchKernelGGL(k_scale, dim3((ne + 255) / 256), dim3(256), 0, c.stream, d_t, d_w, ne);
    if (rocblas_dgemv(fs.h, rocblas_operation_none, ne, ne, &one, Ge, ne, d_t, 1, &zero, dst, 1) !=
        rocblas_status_success)
      throw Error(CG_ERR_DEVICE, "rocblas_dgemv");
    if (masked) {
      hipLaunchKernelGGL(k_zero, dim3((n + 255) / 256), dim3(256), 0, c.stream, out, n);
      hipLaunchKernelGGL(k_scatter_vals, dim3((ne + 255) / 256), dim3(256), 0, c.stream, d_g2, (const int *)d_keep,
                         ne, out);
    }
  };
  auto resid = [&](const double *x) {
    if (c.A_prec == CG_FP64) residual(c, (const double *)c.A, m, n, x, d_b, d_r);
    else residual(c, (const float *)c.A, m, n, x, d_b, d_r);
  };
  auto grad = [&]() {
    if (c.A_prec == CG_FP64) atv(c, (const double *)c.A, m, n, d_r, d_c);
    else atv(c, (const float *)c.A, m, n, d_r, d_c);
  };
  // |r|^2 on the device (one value back per iteration)
  double *d_z = d_E + n + 2;
  auto rnorm = [&]() {
    double z = 0;
    if (rocblas_set_pointer_mode(fs.h, rocblas_pointer_mode_device) != rocblas_status_success ||
        rocblas_ddot(fs.h, (rocblas_int)m, d_r, 1, d_r, 1, d_z) != rocblas_status_success)
      throw Error(CG_ERR_DEVICE, "rocblas_ddot");
    rocblas_set_pointer_mode(fs.h, rocblas_pointer_mode_host);
    CG_HIP(hipMemcpyAsync(&z, d_z, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    CG_HIP(hipStreamSynchronize(c.stream));
    return z;
  };
  if (c.A_prec == CG_FP64) atv(c, (const double *)c.A, m, n, d_b, d_c);
  else atv(c, (const float *)c.A, m, n, d_b, d_c);
  project(d_c, d_x);
  CG_HIP(hipEventRecord(e2, c.stream));
  // refinement: x += P A^T (b - A x), stopped when |r| no longer decreases
  // (kept x is the best iterate; with an inexact G the update can grow)
  resid(d_x);
  double Z = rnorm();
  int iters = 0;
  for (int it = 0; it < refine; ++it) {
    grad();
    project(d_c, d_dx);
    hipLaunchKernelGGL(k_axpy, dim3((n + 255) / 256), dim3(256), 0, c.stream, d_x, d_dx, n, 1.0);
    resid(d_x);
    const double Zn = rnorm();
    if (!(Zn < Z)) {  // undo the step and stop
      hipLaunchKernelGGL(k_axpy, dim3((n + 255) / 256), dim3(256), 0, c.stream, d_x, d_dx, n, -1.0);
      resid(d_x);
      break;
    }
    Z = Zn;
    ++iters;
  }
  CG_HIP(hipEventRecord(e3, c.stream));
  CG_HIP(hipMemcpyAsync(x_host, d_x, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  CG_HIP(hipStreamSynchronize(c.stream));
  double mx = 0;
  for (int i = 0; i < n; ++i) mx = std::max(mx, std::fabs(x_host[i]));
  float tg = 0, ts = 0, tr = 0;
  CG_HIP(hipEventElapsedTime(&tg, e0, e1));
  CG_HIP(hipEventElapsedTime(&ts, e1, e2));
  CG_HIP(hipEventElapsedTime(&tr, e2, e3));
  for (auto e : {e0, e1, e2, e3}) (void)hipEventDestroy(e);
  res->eps = eps;
  res->rms = std::sqrt(Z / (double)m);
  res->max_abs_x = mx;
  res->nvars = nvars;
  res->eps_abs = eps * std::sqrt(lmax);
  res->bic = (double)m * std::log(Z / (double)m) + (double)nvars * std::log((double)m);
  res->ms_gram = tg;
  res->ms_solve = ts;
  res->ms_refine = tr;
  res->iters = iters;
  res->algorithm = CG_ALG_SVD;
  res->alpha = 0.0;
}

void fit_ridge(Ctx &c, const double *b_host, double alpha, double *x_host, cg_fit_result *res) {
  if (!c.A) throw Error(CG_ERR_ARG, "nothing assembled");
  if (!(alpha >= 0.0)) throw Error(CG_ERR_ARG, "ridge alpha must be >= 0");
  const int n = c.A_ld;
  const int nfit = !c.fit_keep.empty() && (int)c.fit_keep.size() < n ? (int)c.fit_keep.size() : n;
  if (c.A_rows < nfit) throw Error(CG_ERR_UNSUPPORTED, "ridge fit needs at least as many rows as columns");
  fit_qr(c, b_host, 1.0, alpha, x_host, res);
  res->eps = 0.0;
  res->eps_abs = 0.0;
}

// G = A^T A (FP64 Gram) and A^T b of the last assembled A, left on the device
// in the fit workspace (for the LassoLars solver).
void gram_and_atb(Ctx &c, const double *b_host, double **G_out, double **Atb_out) {
  if (!c.A) throw Error(CG_ERR_ARG, "nothing assembled");
  const long long m = c.A_rows;
  const int n = c.A_ld;
  long long S, kchunk;
  gram_plan(m, n, S, kchunk);
  double *G = c.b_x.as<double>((size_t)n * n + (size_t)S * n * n + (size_t)m + (size_t)n + 64);
  if (!G) throw Error(CG_ERR_MEMORY, "Gram buffers");
  double *part = G + (size_t)n * n;
  double *d_b = part + (size_t)S * n * n;
  double *d_c = d_b + m;
  gram(c, CG_GRAM_FP64, G, part, S, kchunk);
  CG_HIP(hipMemcpyAsync(d_b, b_host, m * sizeof(double), hipMemcpyHostToDevice, c.stream));
  if (c.A_prec == CG_FP64) atv(c, (const double *)c.A, m, n, d_b, d_c);
  else atv(c, (const float *)c.A, m, n, d_b, d_c);
  *G_out = G;
  *Atb_out = d_c;
}

void matvec(Ctx &c, const double *x_host, double *y_host) {
  const long long m = c.A_rows;
  const int n = c.A_ld;
  double *d = c.b_tmp.as<double>(m + n);
  if (!d) throw Error(CG_ERR_MEMORY, "matvec");
  double *dx = d + m;
  CG_HIP(hipMemcpyAsync(dx, x_host, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  if (c.A_prec == CG_FP64) residual(c, (const double *)c.A, m, n, dx, nullptr, d);
  else residual(c, (const float *)c.A, m, n, dx, nullptr, d);
  CG_HIP(hipMemcpyAsync(y_host, d, m * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  CG_HIP(hipStreamSynchronize(c.stream));
  for (long long i = 0; i < m; ++i) y_host[i] = -y_host[i];
}

}  // namespace cg


extern "C" {

cg_status cg_gram(cg_ctx *ctx, int gram_mode, const double *b_host, double *G, double *Atb) {
  if (!ctx || (!G && !Atb)) { cg::set_error("null argument"); return CG_ERR_ARG; }
  try {
    CG_HIP(hipSetDevice(ctx->c.device));
    cg::gram_host(ctx->c, gram_mode, b_host, G, Atb);
    return CG_OK;
  } catch (const cg::Error &e) {
    cg::set_error(e.what());
    return e.code;
  }
}

cg_status cg_fit(cg_ctx *ctx, const double *b_host, double eps, int refine_iters, int gram_mode, double *x,
                 cg_fit_result *res) {
  if (!ctx || !b_host || !x || !res) { cg::set_error("null argument"); return CG_ERR_ARG; }
  try {
    CG_HIP(hipSetDevice(ctx->c.device));
    cg::fit(ctx->c, b_host, eps, refine_iters, gram_mode, x, res);
    return CG_OK;
  } catch (const cg::Error &e) {
    cg::set_error(e.what());
    return e.code;
  }
}

cg_status cg_fit_ridge(cg_ctx *ctx, const double *b_host, double alpha, double *x, cg_fit_result *res) {
  if (!ctx || !b_host || !x || !res) { cg::set_error("null argument"); return CG_ERR_ARG; }
  try {
    CG_HIP(hipSetDevice(ctx->c.device));
    cg::fit_ridge(ctx->c, b_host, alpha, x, res);
    return CG_OK;
  } catch (const cg::Error &e) {
    cg::set_error(e.what());
    return e.code;
  }
}

cg_status cg_set_fit_weights(cg_ctx *ctx, const double *w, int64_t nrows) {
  if (!ctx || nrows < 0 || (nrows > 0 && !w)) { cg::set_error("cg_set_fit_weights: bad argument"); return CG_ERR_ARG; }
  if (w) ctx->c.fit_w.assign(w, w + nrows);
  else ctx->c.fit_w.clear();
  return CG_OK;
}

cg_status cg_set_fit_columns(cg_ctx *ctx, const int *keep, int nkeep) {
  if (!ctx || nkeep < 0 || (nkeep > 0 && !keep)) { cg::set_error("cg_set_fit_columns: bad argument"); return CG_ERR_ARG; }
  const int n = ctx->c.m.ncols;
  for (int i = 0; i < nkeep; ++i)
    if (keep[i] < 0 || keep[i] >= n || (i > 0 && keep[i] <= keep[i - 1])) {
      cg::set_error("cg_set_fit_columns: columns must be increasing and below ncols");
      return CG_ERR_ARG;
    }
  ctx->c.fit_keep.assign(keep, keep + nkeep);
  return CG_OK;
}

cg_status cg_matvec(cg_ctx *ctx, const double *x, double *y) {
  if (!ctx || !x || !y || !ctx->c.A) { cg::set_error("nothing assembled"); return CG_ERR_ARG; }
  try {
    CG_HIP(hipSetDevice(ctx->c.device));
    cg::matvec(ctx->c, x, y);
    return CG_OK;
  } catch (const cg::Error &e) {
    cg::set_error(e.what());
    return e.code;
  }
}

}  // extern "C"
