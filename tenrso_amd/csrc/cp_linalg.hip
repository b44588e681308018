// cp_linalg.hip -- FP64 normal-equation kernels of the ALS step on gfx950.
//
//   Gram-Hadamard  Gamma_n = hadamard_{k != n} A_k^T A_k      (cp/helpers.rs:409-452)
//   solve          A_n[i,:] = M[i,:] Gamma_n^{-1} for every row (cp/helpers.rs:455-495)
//   fit            1 - sqrt(max(0, |X|^2 + |Xhat|^2 - 2<X,Xhat>)) / |X|   (cp/helpers.rs:498-579)
//
// All of it runs in FP64 on the device; the factors stay FP32.  Every reduction
// has a fixed order (per-block partials summed in block order), so results are
// run-to-run deterministic.  The R x R inverse is an in-LDS Gauss-Jordan
// elimination without pivoting, valid for the symmetric positive definite
// Gamma_n; a non-positive pivot triggers the reference's ridge retry
// (Gamma += T::epsilon() * 10R * I with T = f32, cp/helpers.rs:477-490).
#include <cfloat>

#include "cp_linalg.hpp"

namespace tenrso {

static inline unsigned cdiv(i64 a, i64 b) { return (unsigned)((a + b - 1) / b); }

// wave-wide sum (fixed butterfly order, every lane gets the total)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// block-wide sum of one double per thread (256 threads), fixed order
__device__ double block_sum_256(double v, double* red) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  const double r = (red[0] + red[1]) + (red[2] + red[3]);
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------- ||X||^2, min X
__global__ __launch_bounds__(256) void norm_min_partial(const float* __restrict__ X, i64 n, double* __restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  float mn = FLT_MAX;
  const i64 stride = (i64)gridDim.x * 256;
  for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
    const float x = X[i];
    s += (double)x * (double)x;
    mn = fminf(mn, x);
  }
  const double tot = block_sum_256(s, red);
  red[threadIdx.x] = (double)mn;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x] = tot;
    part[gridDim.x + blockIdx.x] = red[0];
  }
}

__global__ __launch_bounds__(256) void norm_min_final(const double* __restrict__ part, int nblk, double* __restrict__ scal) {
  __shared__ double red[256];
  double s = 0.0, mn = DBL_MAX;
  for (int b = threadIdx.x; b < nblk; b += 256) {
    s += part[b];
    mn = fmin(mn, part[nblk + b]);
  }
  const double tot = block_sum_256(s, red);
  red[threadIdx.x] = mn;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    scal[SC_NORM_X] = tot;
    scal[SC_MIN_X] = red[0];
  }
}

int tensor_norm_min(tenrso_ctx* ctx, const float* X, i64 n, double* tmp, double* scal) {
  const int nblk = (int)std::min<i64>(1024, std::max<i64>(1, cdiv(n, 256 * 16)));
  hipLaunchKernelGGL(norm_min_partial, dim3(nblk), dim3(256), 0, ctx->stream, X, n, tmp);
  hipLaunchKernelGGL(norm_min_final, dim3(1), dim3(256), 0, ctx->stream, tmp, nblk, scal);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

// ---------------------------------------------------------------- Gram A^T A
// grid: R workgroups of 1024 threads; workgroup r computes Gram row r.  Thread
// (ch, c) sums A[i][r] A[i][c] in FP64 over row chunk ch (16 chunks), with all of
// its loads independent; the 16 chunk sums are added in fixed order.
__global__ __launch_bounds__(1024) void gram_rows_kernel(const float* __restrict__ A, i64 rows, int R,
                                                         double* __restrict__ G) {
  __shared__ double part[16 * SOLVE_RMAX];
  const int c = threadIdx.x & 63, ch = threadIdx.x >> 6;
  const int r = blockIdx.x;
  double acc0 = 0.0, acc1 = 0.0;
  if (c < R) {
    const i64 i0 = (i64)ch * rows / 16, i1 = (i64)(ch + 1) * rows / 16;
    i64 i = i0;
    for (; i + 2 <= i1; i += 2) {
      acc0 += (double)A[i * R + r] * (double)A[i * R + c];
      acc1 += (double)A[(i + 1) * R + r] * (double)A[(i + 1) * R + c];
    }
    if (i < i1) acc0 += (double)A[i * R + r] * (double)A[i * R + c];
  }
  part[ch * SOLVE_RMAX + c] = acc0 + acc1;
  __syncthreads();
  if (ch == 0 && c < R) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) s += part[k * SOLVE_RMAX + c];
    G[(i64)r * R + c] = s;
  }
}

int gram_f64(tenrso_ctx* ctx, const float* A, i64 rows, i64 R, double* G, hipStream_t stream) {
  if (R > SOLVE_RMAX) {
    set_error("rank %lld exceeds the on-device Gram bound %d", (long long)R, SOLVE_RMAX);
    return TENRSO_ERR_UNSUPPORTED;
  }
  if (R == 0) return TENRSO_OK;
  hipLaunchKernelGGL(gram_rows_kernel, dim3((unsigned)R), dim3(1024), 0, stream ? stream : ctx->stream, A, rows,
                     (int)R, G);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

// ---------------------------------------------------------------- normal equations
// One workgroup: Gamma = hadamard_{k != skip} G_k (ascending k, from ones) + reg,
// padded to 64 x 64 with the identity, inverted in place by Gauss-Jordan.
// Thread t owns column c = t % 64 and rows rg + 4m (m < 16) in registers; each
// pivot step publishes the pivot row and column through double-buffered LDS,
// so one barrier per step suffices.  The output Winv is 64 x 64 (stride 64);
// entries outside R x R are the identity and never used.
constexpr int NE = 64;

struct NormalEqArgs {
  const double* G;
  int nd, skip;
  int R;
  double l2;
  int tik_order;
  double tik_lambda;
  double* Winv;
  double* scal;
};

__device__ __forceinline__ double gamma_entry(const NormalEqArgs& a, int i, int j, double ridge) {
  const int R = a.R;
  if (i >= R || j >= R) return i == j ? 1.0 : 0.0;
  double v = 1.0;
  for (int k = 0; k < a.nd; ++k)
    if (k != a.skip) v = v * a.G[(i64)k * R * R + i * R + j];
  if (a.tik_order > 0) {
    // lambda * D^T D for first / second differences (cp/helpers.rs:630-702)
    const double lam = a.tik_lambda;
    const int n = R;
    if (a.tik_order == 1) {
      if (n >= 2) {
        if (i == j) v += lam * ((i == 0 || i == n - 1) ? 1.0 : 2.0);
        else if (j == i + 1 || i == j + 1) v -= lam;
      }
    } else {
      if (n >= 3) {
        if (i == j) v += lam * ((i == 0 || i == n - 1) ? 1.0 : ((i == 1 || i == n - 2) ? 5.0 : 6.0));
        else if (j == i + 1 || i == j + 1) {
          const int lo = i < j ? i : j;
          v += lam * ((lo == 0 || lo == n - 2) ? -2.0 : -4.0);
        } else if (j == i + 2 || i == j + 2) {
          v += lam;
        }
      } else if (i == j) {
        v += lam;
      }
    }
  } else if (a.l2 > 0.0 && i == j) {
    v += a.l2;
  }
  if (i == j) v += ridge;
  return v;
}

// FP64 reciprocal: hardware estimate refined by two Newton steps.
__device__ __forceinline__ double recip(double p) {
  double r = __builtin_amdgcn_rcp(p);
  r = fma(r, fma(-p, r, 1.0), r);
  r = fma(r, fma(-p, r, 1.0), r);
  return r;
}

// In-place Gauss-Jordan inverse of the 64 x 64 matrix held as av[m] = A[rg+4m][c]
// (classic in-place form: after step j, column j holds column j of the inverse).
// Step j:
//   * the wave owning row j (rg == j % 4) reads the pivot p = A[j][j] from lane j
//     (readlane), publishes 1/p and the scaled pivot row prow[c] = A[j][c] / p;
//   * the four threads with c == j publish the pivot column pcol[i] = A[i][j],
//     except pcol[j] = p - 1, and reset their own entries to the identity
//     column (1 on row j, 0 elsewhere);
//   * one barrier; then every entry does a single FMA
//       A[i][c] -= pcol[i] * prow[c]
//     which yields A[i][c] - A[i][j] A[j][c] / p off the pivot row and column,
//     A[j][c] / p on the pivot row, -A[i][j] / p on the pivot column and 1/p on
//     the pivot.  A bad pivot (<= 0 or not finite) clears `ok` (no branch).
__device__ bool gj_invert(double (&av)[16], int c, int rg, double* prow, double* pcol, double* pinv) {
  bool ok = true;
  for (int j = 0; j < NE; ++j) {
    const int b = j & 1;
    double p = 0.0;
    if (rg == (j & 3)) {
      // pv = av[j / 4] without a dynamic register index
      double pv = 0.0;
#pragma unroll
      for (int m = 0; m < 16; ++m) pv = (m == (j >> 2)) ? av[m] : pv;
      const int lo = __builtin_amdgcn_readlane(__double2loint(pv), j);
      const int hi = __builtin_amdgcn_readlane(__double2hiint(pv), j);
      p = __hiloint2double(hi, lo);
      const double ip = recip(p);
      prow[b * NE + c] = (c == j) ? ip : pv * ip;
      if (c == 0) pinv[b] = (p > 0.0 && isfinite(p)) ? ip : __longlong_as_double(0x7ff8000000000000LL);
    }
    if (c == j) {
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        const bool diag = (rg + 4 * m) == j;
        pcol[b * NE + rg + 4 * m] = diag ? av[m] - 1.0 : av[m];
        av[m] = diag ? 1.0 : 0.0;
      }
    }
    __syncthreads();
    ok = ok && !isnan(pinv[b]);
    const double ajc = prow[b * NE + c];
#pragma unroll
    for (int m = 0; m < 16; ++m) av[m] = fma(-pcol[b * NE + rg + 4 * m], ajc, av[m]);
  }
  __syncthreads();
  return ok;
}

__global__ __launch_bounds__(256) void normal_eq_kernel(NormalEqArgs a) {
  __shared__ double prow[2 * NE], pcol[2 * NE], pinv[2];
  const int c = threadIdx.x & 63, rg = threadIdx.x >> 6;
  double av[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) av[m] = gamma_entry(a, rg + 4 * m, c, 0.0);
  int path = 0;
  bool ok = gj_invert(av, c, rg, prow, pcol, pinv);
  if (!ok) {
    path = 1;
#pragma unroll
    for (int m = 0; m < 16; ++m) av[m] = gamma_entry(a, rg + 4 * m, c, (double)FLT_EPSILON * (double)(10 * a.R));
    ok = gj_invert(av, c, rg, prow, pcol, pinv);
    if (!ok) path = 2;
  }
#pragma unroll
  for (int m = 0; m < 16; ++m) a.Winv[(rg + 4 * m) * NE + c] = ok ? av[m] : 0.0;
  if (threadIdx.x == 0 && a.scal) a.scal[SC_STATUS] = fmax(a.scal[SC_STATUS], (double)path);
}

int normal_eq_inverse(tenrso_ctx* ctx, const double* G, int nd, int skip, i64 R, const GramReg& reg,
                      double* Winv, double* scal, hipStream_t stream) {
  if (R > SOLVE_RMAX) {
    set_error("rank %lld exceeds the on-device normal-equation solve bound %d", (long long)R, SOLVE_RMAX);
    return TENRSO_ERR_UNSUPPORTED;
  }
  NormalEqArgs a{G, nd, skip, (int)R, reg.l2, reg.tikhonov_order, reg.tikhonov_lambda, Winv, scal};
  hipLaunchKernelGGL(normal_eq_kernel, dim3(1), dim3(256), 0, stream ? stream : ctx->stream, a);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

// ---------------------------------------------------------------- row solve
// grid: rows (one workgroup per row); 256 threads = 4 split quarters x 64
// columns.  Each thread sums a quarter of the MTTKRP split partials of its
// column (all loads in flight, FP64), the quarters are added in fixed order,
// then A[i, c] = epi(sum_r M[i, r] Winv[r, c]) (the r-sum split over the same
// four thread groups) and the block's <M, A> partial is written.
constexpr int SOLVE_RB = 1;

__global__ __launch_bounds__(256) void solve_rows_kernel(const float* __restrict__ part, int splits,
                                                         const double* __restrict__ M64, i64 rows, int R,
                                                         const double* __restrict__ Winv, double l1, int nonneg,
                                                         float* __restrict__ A, double* __restrict__ ip_part) {
  __shared__ double sq[4 * NE];
  __shared__ double sM[NE];
  __shared__ double red[4];
  const int c = threadIdx.x & 63, qd = threadIdx.x >> 6;
  const i64 row = blockIdx.x;
  const bool colok = c < R;
  double m = 0.0;
  if (colok) {
    if (part) {
      const i64 n = rows * R, off = row * R + c;
      const int s0 = qd * splits / 4, s1 = (qd + 1) * splits / 4;
      double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
      int sp = s0;
      for (; sp + 4 <= s1; sp += 4) {
        a0 += (double)part[(i64)(sp + 0) * n + off];
        a1 += (double)part[(i64)(sp + 1) * n + off];
        a2 += (double)part[(i64)(sp + 2) * n + off];
        a3 += (double)part[(i64)(sp + 3) * n + off];
      }
      for (; sp < s1; ++sp) a0 += (double)part[(i64)sp * n + off];
      m = (a0 + a1) + (a2 + a3);
    } else if (qd == 0) {
      m = M64[row * R + c];
    }
  }
  sq[qd * NE + c] = m;
  __syncthreads();
  if (qd == 0) sM[c] = (sq[c] + sq[NE + c]) + (sq[2 * NE + c] + sq[3 * NE + c]);
  __syncthreads();
  // x[c] = sum_r M[r] Winv[r, c], r split over the four quarters
  double x = 0.0;
  if (colok) {
    const int r0 = qd * R / 4, r1 = (qd + 1) * R / 4;
    double x0 = 0, x1 = 0;
    int r = r0;
    for (; r + 2 <= r1; r += 2) {
      x0 += sM[r] * Winv[r * NE + c];
      x1 += sM[r + 1] * Winv[(r + 1) * NE + c];
    }
    if (r < r1) x0 += sM[r] * Winv[r * NE + c];
    x = x0 + x1;
  }
  sq[qd * NE + c] = x;
  __syncthreads();
  double ip = 0.0;
  if (qd == 0 && colok) {
    double xv = (sq[c] + sq[NE + c]) + (sq[2 * NE + c] + sq[3 * NE + c]);
    if (l1 > 0.0) xv = xv > l1 ? xv - l1 : (xv < -l1 ? xv + l1 : 0.0);  // soft_threshold (helpers.rs:610-623)
    if (nonneg) xv = xv > 0.0 ? xv : 0.0;                                  // core.rs:357-360
    const float x32 = (float)xv;
    A[row * R + c] = x32;
    ip = sM[c] * (double)x32;
  }
  if (qd == 0) {
    ip = wave_sum(ip);
    if (c == 0 && ip_part) ip_part[blockIdx.x] = ip;
  }
  (void)red;
}

int solve_rows(tenrso_ctx* ctx, const float* part, int splits, const double* M, i64 rows, i64 R,
               const double* Winv, const RowEpilogue& epi, float* A, double* ip_part, int* nblk_out) {
  const unsigned nblk = cdiv(rows, SOLVE_RB);
  if (nblk_out) *nblk_out = (int)nblk;
  if (rows == 0) return TENRSO_OK;
  if (R > SOLVE_RMAX) {
    set_error("rank %lld exceeds the on-device solve bound %d", (long long)R, SOLVE_RMAX);
    return TENRSO_ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL(solve_rows_kernel, dim3(nblk), dim3(256), 0, ctx->stream, part, splits, M, rows, (int)R, Winv,
                     epi.l1, epi.nonneg, A, ip_part);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

__global__ __launch_bounds__(256) void inner_rows_kernel(const double* __restrict__ M, const float* __restrict__ A,
                                                         i64 rows, i64 R, double* __restrict__ ip_part) {
  __shared__ double red[256];
  const i64 r0 = (i64)blockIdx.x * SOLVE_RB;
  const i64 r1 = (r0 + SOLVE_RB) < rows ? r0 + SOLVE_RB : rows;
  double ip = 0.0;
  for (i64 e = r0 * R + threadIdx.x; e < r1 * R; e += 256) ip += M[e] * (double)A[e];
  const double tot = block_sum_256(ip, red);
  if (threadIdx.x == 0) ip_part[blockIdx.x] = tot;
}

int inner_rows(tenrso_ctx* ctx, const double* M, const float* A, i64 rows, i64 R, double* ip_part, int* nblk_out) {
  const unsigned nblk = cdiv(rows, SOLVE_RB);
  if (nblk_out) *nblk_out = (int)nblk;
  if (rows == 0) return TENRSO_OK;
  hipLaunchKernelGGL(inner_rows_kernel, dim3(nblk), dim3(256), 0, ctx->stream, M, A, rows, R, ip_part);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

// ---------------------------------------------------------------- fit
__global__ __launch_bounds__(256) void fit_kernel(const double* __restrict__ G, int nd, int R,
                                                  const double* __restrict__ ip_part, int nblk,
                                                  double* __restrict__ scal) {
  __shared__ double red[4];
  // ||Xhat||^2 = sum_{r,s} prod_k G_k[r,s]   (cp/helpers.rs:533-555)
  double rec = 0.0;
  for (int e = threadIdx.x; e < R * R; e += 256) {
    double cross = 1.0;
    for (int k = 0; k < nd; ++k) cross = cross * G[(i64)k * R * R + e];
    rec += cross;
  }
  const double recon = block_sum_256(rec, red);
  double ipv = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 256) ipv += ip_part[b];
  const double ip = block_sum_256(ipv, red);
  if (threadIdx.x == 0) {
    const double nx = scal[SC_NORM_X];
    const double err_sq = nx + recon - 2.0 * ip;
    const double err = sqrt(err_sq > 0.0 ? err_sq : 0.0);
    double fit = 1.0 - err / sqrt(nx);
    fit = fit < 0.0 ? 0.0 : (fit > 1.0 ? 1.0 : fit);
    scal[SC_FIT] = fit;
    scal[SC_IP] = ip;
    scal[SC_RECON] = recon;
  }
}

int fit_from_grams(tenrso_ctx* ctx, const double* G, int nd, i64 R, const double* ip_part, int nblk, double* scal) {
  hipLaunchKernelGGL(fit_kernel, dim3(1), dim3(256), 0, ctx->stream, G, nd, (int)R, ip_part, nblk, scal);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

// ---------------------------------------------------------------- small utilities
__global__ __launch_bounds__(256) void clamp_nonneg_kernel(float* __restrict__ A, i64 n) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i < n) A[i] = fmaxf(A[i], 0.0f);
}
int clamp_nonneg(tenrso_ctx* ctx, float* A, i64 n) {
  if (n == 0) return TENRSO_OK;
  hipLaunchKernelGGL(clamp_nonneg_kernel, dim3(cdiv(n, 256)), dim3(256), 0, ctx->stream, A, n);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

__global__ __launch_bounds__(256) void fill_f64_kernel(double* __restrict__ p, i64 n, double v) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}
int fill_f64(tenrso_ctx* ctx, double* p, i64 n, double v) {
  if (n == 0) return TENRSO_OK;
  hipLaunchKernelGGL(fill_f64_kernel, dim3(cdiv(n, 256)), dim3(256), 0, ctx->stream, p, n, v);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

}  // namespace tenrso
