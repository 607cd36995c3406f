// LASSO by least-angle regression on the GPU: chimes_lsq.py --algorithm=lassolars
// (src/chimes_lsq.py:248-264), i.e. scikit-learn's LassoLars(alpha,
// fit_intercept=False, fit_path=False, max_iter), whose solver works on the
// Gram matrix G = A^T A and Xy = A^T b when A has more rows than columns
// (scikit-learn 1.7.2, linear_model/_least_angle.py, _lars_path_solver).
//
// The path is sequential -- one variable enters or leaves per step -- and each
// step is a few O(n k) products and O(k^2) triangular solves on the active
// set (k variables), so the whole loop runs in ONE persistent workgroup of
// 1024 threads: G, Xy and the Cholesky factor of the active Gram stay in
// global memory (L2-resident), the per-feature state in LDS, and no step
// goes back to the host.  The steps restate the published algorithm:
//
//   C = max |Cov| over inactive positions (first on ties), alpha = C / m;
//   stop at alpha <= alpha_min (+ float32 eps), interpolating the last step
//   to alpha_min;
//   add: Cholesky row L[k, :k] = G[j, active] solved against L, diagonal
//   max(sqrt(|G_jj - |w|^2|), eps) (a pivot below 1e-7 drops the candidate);
//   ls = (L L^T)^-1 sign, AA = 1/sqrt(ls . sign), ls *= AA;
//   corr = G[active, inactive]^T ls rounded to 15 decimals;
//   gamma = min(min_pos((C - Cov)/(AA - corr)), min_pos((C + Cov)/(AA + corr)), C/AA);
//   a coefficient crossing zero first (min_pos(-coef/ls) < gamma) shortens the
//   step to it and is dropped (Givens update of L), re-entering the inactive
//   list at its front with Cov = Xy_j - G_j . coef.
// Positions follow the reference solver's bookkeeping (selected variable
// swapped to the front of the inactive block), so ties resolve alike.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <string>
#include <vector>

#include "device.hpp"
#include "setup.hpp"

namespace cg {
namespace {

constexpr int LARS_BS = 1024;
constexpr int LARS_NMAX = 2048;  // columns (LDS state: 6 doubles + 1 int per column)

struct LarsOut {
  int n_iter, n_active, status;  // status: 0 ok, 1 alpha increased (early stop), 2 max_iter, 3 guard
  int pad;
  double alpha;
};

// block-wide reductions over the 16 waves; every thread returns the result
__device__ double block_sum(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < LARS_BS / 64; ++i) s += red[i];
  return s;
}

__device__ double block_min(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double s = red[0];
  for (int i = 1; i < LARS_BS / 64; ++i) s = fmin(s, red[i]);
  return s;
}

// (largest |v|, smallest position among equals)
__device__ void block_argmax(double v, int pos, double *red, int *redi, double &best, int &bpos) {
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off, 64);
    const int op = __shfl_xor(pos, off, 64);
    if (ov > v || (ov == v && op < pos)) { v = ov; pos = op; }
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) { red[w] = v; redi[w] = pos; }
  __syncthreads();
  best = red[0];
  bpos = redi[0];
  for (int i = 1; i < LARS_BS / 64; ++i)
    if (red[i] > best || (red[i] == best && redi[i] < bpos)) { best = red[i]; bpos = redi[i]; }
}

// Lower-triangular solve L x = r in place (r in LDS, L row-major with stride
// ld), column by column as reference BLAS dtrsv does; dshift adds to the
// diagonal (0 unless the system has to be regularised).
__device__ void trsv_lower(const double *L, int ld, int k, double *r, double dshift) {
  for (int j = 0; j < k; ++j) {
    __syncthreads();
    const double xj = r[j] / (L[(size_t)j * ld + j] + dshift);
    __syncthreads();
    if (threadIdx.x == 0) r[j] = xj;
    for (int i = j + 1 + threadIdx.x; i < k; i += LARS_BS) r[i] -= L[(size_t)i * ld + j] * xj;
  }
  __syncthreads();
}

// L^T x = r in place, column by column from the last row up.
__device__ void trsv_upper_t(const double *L, int ld, int k, double *r, double dshift) {
  for (int j = k - 1; j >= 0; --j) {
    __syncthreads();
    const double xj = r[j] / (L[(size_t)j * ld + j] + dshift);
    __syncthreads();
    if (threadIdx.x == 0) r[j] = xj;
    for (int i = threadIdx.x; i < j; i += LARS_BS) r[i] -= L[(size_t)j * ld + i] * xj;
  }
  __syncthreads();
}

// reference BLAS drotg: r = sign * sqrt(a^2 + b^2) (sign of the larger), c = a/r, s = b/r
__device__ void rotg(double &a, double &b, double &c, double &s) {
  const double roe = fabs(a) > fabs(b) ? a : b;
  const double scale = fabs(a) + fabs(b);
  if (scale == 0.0) { c = 1.0; s = 0.0; a = 0.0; b = 0.0; return; }
  double r = scale * sqrt((a / scale) * (a / scale) + (b / scale) * (b / scale));
  r = copysign(1.0, roe) * r;
  c = a / r;
  s = b / r;
  a = r;
}

__global__ void __launch_bounds__(LARS_BS) k_lars(const double *__restrict__ G, const double *__restrict__ Xy, int n,
                                                  double n_samples, double alpha_min, int max_iter,
                                                  double *__restrict__ L, double *__restrict__ coef_out,
                                                  LarsOut *__restrict__ out) {
  __shared__ double covp[LARS_NMAX];  // Cov by position (inactive block [k, n))
  __shared__ double coef[LARS_NMAX], prev[LARS_NMAX];  // by feature
  __shared__ double corr[LARS_NMAX];  // equiangular correlations by position
  __shared__ double ls[LARS_NMAX];    // least-squares direction / solve scratch (active positions)
  __shared__ double sgn[LARS_NMAX];   // sign_active by active position
  __shared__ int ind[LARS_NMAX];      // position -> feature
  __shared__ double red[LARS_BS / 64];
  __shared__ int redi[LARS_BS / 64];
  const int tid = threadIdx.x;
  const double eps = 2.220446049250313e-16;    // LassoLars default eps (float64 epsilon)
  const double tol = 1.1920928955078125e-07;   // float32 epsilon (equality tolerance)
  const double tiny32 = 1.1754943508222875e-38;
  for (int i = tid; i < n; i += LARS_BS) {
    covp[i] = Xy[i];
    coef[i] = prev[i] = 0.0;
    ind[i] = i;
  }
  __syncthreads();
  int k = 0, n_iter = 0, status = 0;
  bool drop = false;
  double alpha = 0.0, prev_alpha = 0.0;
  // every pass either ends the path, takes a step (n_iter + 1) or zeroes one
  // degenerate candidate's correlation; the bound keeps the loop finite
  const long long guard_max = (long long)max_iter + 4LL * n + 16;
  long long guard = 0;
  while (true) {
    if (++guard > guard_max) { status = 3; break; }
    // ---- largest correlation among the inactive positions
    double C = 0.0;
    int cpos = -1;
    {
      double v = -1.0;
      int p = n;
      for (int i = k + tid; i < n; i += LARS_BS) {
        const double a = fabs(covp[i]);
        if (a > v) { v = a; p = i; }
      }
      double best;
      int bp;
      block_argmax(v, p, red, redi, best, bp);
      if (k < n) { C = best; cpos = bp; }
    }
    const double Csigned = cpos >= 0 ? covp[cpos] : 0.0;
    __syncthreads();
    alpha = C / n_samples;
    if (alpha <= alpha_min + tol) {
      if (fabs(alpha - alpha_min) > tol) {
        if (n_iter > 0) {
          const double ss = (prev_alpha - alpha_min) / (prev_alpha - alpha);
          for (int i = tid; i < n; i += LARS_BS) coef[i] = prev[i] + ss * (coef[i] - prev[i]);
        }
        alpha = alpha_min;
      }
      break;
    }
    if (n_iter >= max_iter || k >= n) { status = n_iter >= max_iter ? 2 : 0; break; }
    if (!drop) {
      // ---- the selected position moves to the front of the inactive block
      const int m = k, p = cpos;
      const double sg = Csigned > 0 ? 1.0 : (Csigned < 0 ? -1.0 : 0.0);
      __syncthreads();
      if (tid == 0) {
        const double t = covp[m]; covp[m] = covp[p]; covp[p] = t;
        const int ti = ind[m]; ind[m] = ind[p]; ind[p] = ti;
        sgn[k] = sg;
      }
      __syncthreads();
      const int j = ind[k];
      // Cholesky row: L[k, :k] = G[j, active] solved against L[:k, :k]
      for (int a = tid; a < k; a += LARS_BS) ls[a] = G[(size_t)j * n + ind[a]];
      trsv_lower(L, n, k, ls, 0.0);
      double v = 0.0;
      for (int a = tid; a < k; a += LARS_BS) v += ls[a] * ls[a];
      v = block_sum(v, red);
      const double c = G[(size_t)j * n + j];
      const double diag = fmax(sqrt(fabs(c - v)), eps);
      if (diag < 1e-7) {
        // degenerate candidate: its correlation is zeroed, the position
        // swap undone on Cov only (as the reference solver does)
        __syncthreads();
        if (tid == 0) {
          covp[m] = 0.0;
          const double t = covp[m]; covp[m] = covp[p]; covp[p] = t;
        }
        __syncthreads();
        continue;
      }
      for (int a = tid; a < k; a += LARS_BS) L[(size_t)k * n + a] = ls[a];
      if (tid == 0) L[(size_t)k * n + k] = diag;
      k += 1;
      __syncthreads();
    }
    if (n_iter > 0 && prev_alpha < alpha) { status = 1; break; }  // alpha no longer controlled
    // ---- least-squares direction on the active set
    double dshift = 0.0, AA;
    {
      for (int i = 0;; ++i) {
        if (i > 200) { AA = 1.0; break; }  // not reached for a finite Gram
        for (int a = tid; a < k; a += LARS_BS) ls[a] = sgn[a];
        trsv_lower(L, n, k, ls, dshift);
        trsv_upper_t(L, n, k, ls, dshift);
        double sm = 0.0;
        for (int a = tid; a < k; a += LARS_BS) sm += ls[a] * sgn[a];
        sm = block_sum(sm, red);
        if (i > 0) sm = fmax(sm, eps);
        AA = 1.0 / sqrt(sm);
        if (isfinite(AA)) break;
        dshift += ldexp(1.0, i) * eps;  // regularise an ill-conditioned factor
      }
      for (int a = tid; a < k; a += LARS_BS) ls[a] *= AA;
      __syncthreads();
    }
    // ---- equiangular correlations of the inactive positions (rounded to 15 decimals)
    for (int q = k + tid; q < n; q += LARS_BS) {
      const size_t row = (size_t)ind[q] * n;
      double s = 0.0;
      for (int a = 0; a < k; ++a) s += G[row + ind[a]] * ls[a];
      corr[q] = rint(s * 1e15) / 1e15;
    }
    __syncthreads();
    double g = DBL_MAX;
    for (int q = k + tid; q < n; q += LARS_BS) {
      const double g1 = (C - covp[q]) / (AA - corr[q] + tiny32);
      const double g2 = (C + covp[q]) / (AA + corr[q] + tiny32);
      if (g1 > 0.0 && g1 < g) g = g1;
      if (g2 > 0.0 && g2 < g) g = g2;
    }
    double gamma = fmin(block_min(g, red), C / AA);
    // ---- first coefficient to cross zero
    double zp = DBL_MAX;
    for (int a = tid; a < k; a += LARS_BS) {
      const double z = -coef[ind[a]] / (ls[a] + tiny32);
      if (z > 0.0 && z < zp) zp = z;
    }
    const double z_pos = block_min(zp, red);
    drop = z_pos < gamma;
    if (drop) gamma = z_pos;
    n_iter += 1;
    prev_alpha = alpha;
    for (int i = tid; i < n; i += LARS_BS) {
      prev[i] = coef[i];
      coef[i] = 0.0;
    }
    __syncthreads();
    for (int a = tid; a < k; a += LARS_BS) coef[ind[a]] = prev[ind[a]] + gamma * ls[a];
    for (int q = k + tid; q < n; q += LARS_BS) covp[q] -= gamma * corr[q];
    __syncthreads();
    if (drop) {
      // positions where z equals z_pos, in decreasing order (one in practice)
      for (int a = k - 1; a >= 0; --a) {
        const double z = -prev[ind[a]] / (ls[a] + tiny32);
        if (z != z_pos) continue;
        // Cholesky downdate: remove row/column a (sklearn cholesky_delete)
        const int kk = k;
        for (int i = a; i < kk - 1; ++i) {
          __syncthreads();
          for (int c2 = tid; c2 < i + 2; c2 += LARS_BS) {
            const double t = L[(size_t)(i + 1) * n + c2];
            ls[c2] = t;  // staged row
          }
          __syncthreads();
          for (int c2 = tid; c2 < i + 2; c2 += LARS_BS) L[(size_t)i * n + c2] = ls[c2];
        }
        __syncthreads();
        for (int i = a; i < kk - 1; ++i) {
          __shared__ double cs_[2];
          if (tid == 0) {
            double x0 = L[(size_t)i * n + i], x1 = L[(size_t)i * n + i + 1], c = 0.0, s = 0.0;
            rotg(x0, x1, c, s);
            if (x0 < 0.0) { x0 = fabs(x0); c = -c; s = -s; }
            L[(size_t)i * n + i] = x0;
            L[(size_t)i * n + i + 1] = 0.0;
            cs_[0] = c;
            cs_[1] = s;
          }
          __syncthreads();
          const double c = cs_[0], s = cs_[1];
          for (int r = i + 1 + tid; r < kk - 1; r += LARS_BS) {
            const double x = L[(size_t)r * n + i], y = L[(size_t)r * n + i + 1];
            L[(size_t)r * n + i] = c * x + s * y;
            L[(size_t)r * n + i + 1] = c * y - s * x;
          }
          __syncthreads();
        }
        // the dropped variable moves to the end of the active block
        const int feat = ind[a];
        __syncthreads();
        if (tid == 0) {
          for (int i = a; i < kk - 1; ++i) {
            ind[i] = ind[i + 1];
            sgn[i] = sgn[i + 1];
          }
          ind[kk - 1] = feat;
          sgn[kk - 1] = 0.0;
        }
        __syncthreads();
        k -= 1;
        // its correlation from the Gram: Xy_j - G_j . coef
        double t = 0.0;
        for (int i = tid; i < n; i += LARS_BS) t += G[(size_t)feat * n + i] * coef[i];
        t = block_sum(t, red);
        if (tid == 0) covp[k] = Xy[feat] - t;
        __syncthreads();
        break;
      }
    }
  }
  __syncthreads();
  for (int i = tid; i < n; i += LARS_BS) coef_out[i] = coef[i];
  if (tid == 0) {
    out->n_iter = n_iter;
    out->n_active = k;
    out->status = status;
    out->alpha = alpha;
  }
}

}  // namespace

// G = A^T A and A^T b of the last assembled A (FP64, fit.hip).
void gram_and_atb(Ctx &c, const double *b_host, double **G, double **Atb);

void fit_lassolars(Ctx &c, const double *b_host, double alpha, int max_iter, double *x_host, cg_fit_result *res) {
  if (!c.A) throw Error(CG_ERR_ARG, "nothing assembled");
  const long long m = c.A_rows;
  const int n = c.A_ld;
  if (n > LARS_NMAX) throw Error(CG_ERR_UNSUPPORTED, "LassoLars supports at most 2048 columns");
  if (m <= n) throw Error(CG_ERR_UNSUPPORTED, "LassoLars needs more rows than columns (Gram form)");
  if (!c.fit_keep.empty() || !c.fit_w.empty())
    throw Error(CG_ERR_UNSUPPORTED, "LassoLars with column subsets or row weights is not supported");
  if (!(alpha > 0.0)) throw Error(CG_ERR_ARG, "LASSO alpha must be positive");
  double *G = nullptr, *Xy = nullptr;
  hipEvent_t e0, e1, e2;
  CG_HIP(hipEventCreate(&e0)); CG_HIP(hipEventCreate(&e1)); CG_HIP(hipEventCreate(&e2));
  CG_HIP(hipEventRecord(e0, c.stream));
  gram_and_atb(c, b_host, &G, &Xy);
  CG_HIP(hipEventRecord(e1, c.stream));
  char *w = static_cast<char *>(c.b_lars.get((size_t)n * n * sizeof(double) + 2 * (size_t)n * sizeof(double) + 256));
  if (!w) throw Error(CG_ERR_MEMORY, "LassoLars buffers");
  double *L = reinterpret_cast<double *>(w);
  double *d_x = L + (size_t)n * n;
  LarsOut *d_out = reinterpret_cast<LarsOut *>(d_x + n);
  hipLaunchKernelGGL(k_lars, dim3(1), dim3(LARS_BS), 0, c.stream, (const double *)G, (const double *)Xy, n, (double)m,
                     alpha, max_iter, L, d_x, d_out);
  CG_HIP(hipGetLastError());
  CG_HIP(hipEventRecord(e2, c.stream));
  LarsOut out{};
  CG_HIP(hipMemcpyAsync(x_host, d_x, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  CG_HIP(hipMemcpyAsync(&out, d_out, sizeof(LarsOut), hipMemcpyDeviceToHost, c.stream));
  CG_HIP(hipStreamSynchronize(c.stream));
  if (out.status == 3) throw Error(CG_ERR_NUMERIC, "LassoLars path did not terminate");
  float tg = 0, ts = 0;
  CG_HIP(hipEventElapsedTime(&tg, e0, e1));
  CG_HIP(hipEventElapsedTime(&ts, e1, e2));
  for (auto e : {e0, e1, e2}) (void)hipEventDestroy(e);
  // residual and summary as chimes_lsq.py prints them (count_nonzero_vars: |x| > 1e-5)
  std::vector<double> y(m);
  matvec(c, x_host, y.data());
  std::vector<double> b(b_host, b_host + m);
  double Z = 0.0, mx = 0.0;
  for (long long i = 0; i < m; ++i) Z += (y[i] - b[i]) * (y[i] - b[i]);
  int nv = 0;
  for (int i = 0; i < n; ++i) {
    mx = std::max(mx, std::fabs(x_host[i]));
    if (std::fabs(x_host[i]) > 1.0e-5) ++nv;
  }
  res->eps = 0.0;
  res->eps_abs = 0.0;
  res->rms = std::sqrt(Z / (double)m);
  res->max_abs_x = mx;
  res->nvars = nv;
  res->bic = (double)m * std::log(Z / (double)m) + (double)nv * std::log((double)m);
  res->ms_gram = tg;
  res->ms_solve = ts;
  res->ms_refine = 0.f;
  res->iters = out.n_iter;
  res->algorithm = CG_ALG_LASSOLARS;
  res->alpha = alpha;
}

}  // namespace cg

extern "C" cg_status cg_fit_lassolars(cg_ctx *ctx, const double *b_host, double alpha, int max_iter, double *x,
                                      cg_fit_result *res) {
  if (!ctx || !b_host || !x || !res || max_iter < 1) {
    cg::set_error("cg_fit_lassolars: bad argument");
    return CG_ERR_ARG;
  }
  try {
    CG_HIP(hipSetDevice(ctx->c.device));
    cg::fit_lassolars(ctx->c, b_host, alpha, max_iter, x, res);
    return CG_OK;
  } catch (const cg::Error &e) {
    cg::set_error(e.what());
    return e.code;
  }
}
