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

// block-wide sum of one double per thread (256 threads), fixed tree order
__device__ double block_sum_256(double v, double* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) red[t] += red[t + s];
    __syncthreads();
  }
  double r = red[0];
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
// grid (ceil(R*R/256), rsplit): partial over a row range, then fixed-order reduce.
__global__ __launch_bounds__(256) void gram_partial(const float* __restrict__ A, i64 rows, i64 R, int rsplit,
                                                    double* __restrict__ part) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e >= R * R) return;
  const i64 r = e / R, s = e - (e / R) * R;
  const i64 i0 = (i64)blockIdx.y * rows / rsplit, i1 = (i64)(blockIdx.y + 1) * rows / rsplit;
  double acc = 0.0;
  for (i64 i = i0; i < i1; ++i) acc += (double)A[i * R + r] * (double)A[i * R + s];
  part[(i64)blockIdx.y * R * R + e] = acc;
}

__global__ __launch_bounds__(256) void sum_parts(const double* __restrict__ part, int nparts, i64 n,
                                                 double* __restrict__ out) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  double s = 0.0;
  for (int p = 0; p < nparts; ++p) s += part[(i64)p * n + e];
  out[e] = s;
}

int gram_f64(tenrso_ctx* ctx, const float* A, i64 rows, i64 R, double* G) {
  const int rsplit = (int)std::max<i64>(1, std::min<i64>(64, rows / 64));
  void* buf = nullptr;
  TENRSO_TRY(ctx_buffer(ctx, SLOT_GRAM_PART, (size_t)rsplit * R * R * sizeof(double), &buf));
  double* part = (double*)buf;
  hipLaunchKernelGGL(gram_partial, dim3(cdiv(R * R, 256), rsplit), dim3(256), 0, ctx->stream, A, rows, R, rsplit,
                     part);
  hipLaunchKernelGGL(sum_parts, dim3(cdiv(R * R, 256)), dim3(256), 0, ctx->stream, part, rsplit, R * R, G);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

// ---------------------------------------------------------------- normal equations
// One workgroup: Gamma = hadamard_{k != skip} G_k (ascending k, from ones) + reg,
// then Gamma^{-1} by in-place Gauss-Jordan in LDS (R <= SOLVE_RMAX).
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

__device__ void build_gamma(double* A, const NormalEqArgs& a, double ridge) {
  const int R = a.R;
  for (int e = threadIdx.x; e < R * R; e += blockDim.x) {
    double v = 1.0;
    for (int k = 0; k < a.nd; ++k)
      if (k != a.skip) v = v * a.G[(i64)k * R * R + e];
    const int i = e / R, j = e - (e / R) * R;
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
    A[e] = v;
  }
}

// Gauss-Jordan inverse in place; returns false on a non-positive or non-finite pivot.
__device__ bool gauss_jordan_spd(double* A, int R) {
  constexpr int MAXE = SOLVE_RMAX * SOLVE_RMAX / 256;
  double tmp[MAXE];
  for (int j = 0; j < R; ++j) {
    __syncthreads();
    const double p = A[j * R + j];
    if (!(p > 0.0) || !isfinite(p)) return false;  // uniform: every thread reads the same p
    const double ip = 1.0 / p;
#pragma unroll
    for (int m = 0; m < MAXE; ++m) {
      const int e = threadIdx.x + 256 * m;
      if (e < R * R) {
        const int i = e / R, c = e - (e / R) * R;
        const double aij = A[i * R + j], ajc = A[j * R + c], aic = A[e];
        double v;
        if (i == j && c == j) v = ip;
        else if (i == j) v = ajc * ip;
        else if (c == j) v = -aij * ip;
        else v = aic - aij * ajc * ip;
        tmp[m] = v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < MAXE; ++m) {
      const int e = threadIdx.x + 256 * m;
      if (e < R * R) A[e] = tmp[m];
    }
  }
  __syncthreads();
  return true;
}

__global__ __launch_bounds__(256) void normal_eq_kernel(NormalEqArgs a) {
  __shared__ double A[SOLVE_RMAX * SOLVE_RMAX];
  const int R = a.R;
  int path = 0;
  build_gamma(A, a, 0.0);
  bool ok = gauss_jordan_spd(A, R);
  if (!ok) {
    path = 1;
    __syncthreads();
    build_gamma(A, a, (double)FLT_EPSILON * (double)(10 * R));
    ok = gauss_jordan_spd(A, R);
    if (!ok) path = 2;
  }
  for (int e = threadIdx.x; e < R * R; e += 256) a.Winv[e] = ok ? A[e] : 0.0;
  if (threadIdx.x == 0 && a.scal) a.scal[SC_STATUS] = fmax(a.scal[SC_STATUS], (double)path);
}

int normal_eq_inverse(tenrso_ctx* ctx, const double* G, int nd, int skip, i64 R, const GramReg& reg,
                      double* Winv, double* scal) {
  if (R > SOLVE_RMAX) {
    set_error("rank %lld exceeds the on-device normal-equation solve bound %d", (long long)R, SOLVE_RMAX);
    return TENRSO_ERR_UNSUPPORTED;
  }
  NormalEqArgs a{G, nd, skip, (int)R, reg.l2, reg.tikhonov_order, reg.tikhonov_lambda, Winv, scal};
  hipLaunchKernelGGL(normal_eq_kernel, dim3(1), dim3(256), 0, ctx->stream, a);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

// ---------------------------------------------------------------- row solve
// grid: ceil(rows / 16); 256 threads = 4 row groups x 64 columns.
constexpr int SOLVE_RB = 16;

__global__ __launch_bounds__(256) void solve_rows_kernel(const double* __restrict__ M, i64 rows, int R,
                                                         const double* __restrict__ Winv, double l1, int nonneg,
                                                         float* __restrict__ A, double* __restrict__ ip_part) {
  __shared__ double sM[SOLVE_RB * SOLVE_RMAX];
  __shared__ double red[256];
  const i64 r0 = (i64)blockIdx.x * SOLVE_RB;
  const int nrows = (int)((rows - r0) < SOLVE_RB ? (rows - r0) : SOLVE_RB);
  for (int e = threadIdx.x; e < nrows * R; e += 256) sM[e] = M[r0 * R + e];
  __syncthreads();
  const int c = threadIdx.x & 63, rg = threadIdx.x >> 6;
  double ip = 0.0;
  for (int cc = c; cc < R; cc += 64) {
    double acc[SOLVE_RB / 4];
#pragma unroll
    for (int k = 0; k < SOLVE_RB / 4; ++k) acc[k] = 0.0;
    for (int r = 0; r < R; ++r) {
      const double w = Winv[r * R + cc];
#pragma unroll
      for (int k = 0; k < SOLVE_RB / 4; ++k) acc[k] += sM[(rg + 4 * k) * R + r] * w;
    }
#pragma unroll
    for (int k = 0; k < SOLVE_RB / 4; ++k) {
      const int lr = rg + 4 * k;
      if (lr < nrows) {
        double x = acc[k];
        if (l1 > 0.0) x = x > l1 ? x - l1 : (x < -l1 ? x + l1 : 0.0);  // soft_threshold (helpers.rs:610-623)
        if (nonneg) x = x > 0.0 ? x : 0.0;                               // core.rs:357-360
        const float x32 = (float)x;
        A[(r0 + lr) * R + cc] = x32;
        ip += sM[lr * R + cc] * (double)x32;
      }
    }
  }
  const double tot = block_sum_256(ip, red);
  if (threadIdx.x == 0 && ip_part) ip_part[blockIdx.x] = tot;
}

int solve_rows(tenrso_ctx* ctx, const double* M, i64 rows, i64 R, const double* Winv, const RowEpilogue& epi,
               float* A, double* ip_part, int* nblk_out) {
  const unsigned nblk = cdiv(rows, SOLVE_RB);
  if (nblk_out) *nblk_out = (int)nblk;
  if (rows == 0) return TENRSO_OK;
  if (R > SOLVE_RMAX) {
    set_error("rank %lld exceeds the on-device solve bound %d", (long long)R, SOLVE_RMAX);
    return TENRSO_ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL(solve_rows_kernel, dim3(nblk), dim3(256), 0, ctx->stream, M, rows, (int)R, Winv, epi.l1,
                     epi.nonneg, A, ip_part);
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
  __shared__ double red[256];
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
