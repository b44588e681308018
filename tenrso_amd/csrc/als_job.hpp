// als_job.hpp -- the per-mode FP64 linear algebra of the ALS step as one
// workgroup of 256 threads ("the job"):
//
//   1. optionally the fit of the previous sweep from the Grams, <X, Xhat> and
//      ||X||^2 (cp/helpers.rs:513-555),
//   2. Gamma_n = hadamard_{k != n} G_k + regularisation (cp/helpers.rs:409-452,
//      630-702), padded to 64 x 64 with the identity,
//   3. Gamma_n^{-1} by block Gauss-Jordan (4 x 4 pivot blocks) on the FP64
//      matrix cores, written to Winv.
//
// The job runs either as its own kernel or as workgroup 0 of the MTTKRP launch
// of mode n, so that it shares the GPU with the MTTKRP instead of queueing
// behind it.
#pragma once
#include <cfloat>

#include "common.hpp"

// Optional phase marks for timing builds (off by default).
#ifndef TENRSO_JOB_MARK
#define TENRSO_JOB_MARK(id)
#endif

namespace tenrso {

constexpr int JOB_NE = 64;        // padded matrix order (rank bound)
constexpr int JOB_THREADS = 256;  // workgroup size of the job
// GJ: [2][16] pivot blocks, [2][4][64] pivot rows, [2][64][4] pivot columns,
// [64][8] regularisation band.  Thread t owns the 4 x 4 block (t / 16, t % 16)
// of every 64 x 64 matrix.
constexpr int JOB_GJ_P = 0, JOB_GJ_ROW = 32, JOB_GJ_Q = JOB_GJ_ROW + 2 * 4 * JOB_NE;
constexpr int JOB_GJ_REG = JOB_GJ_Q + 2 * 4 * JOB_NE;
constexpr int JOB_LDS_BYTES = (JOB_GJ_REG + JOB_NE * 8) * 8;  // LDS bytes of the job

// Device scalar block slots shared with the ALS driver.
constexpr int JOB_SC_NORM_X = 0, JOB_SC_FIT = 1, JOB_SC_IP = 2, JOB_SC_RECON = 3, JOB_SC_STATUS = 4;

struct AlsJob {
  double* G;             // [nd][R][R] Grams (FP64)
  int nd, R;
  int do_fit;            // evaluate the fit before forming Gamma
  const double* ip_part; // <X, Xhat> partials of the previous sweep
  int nblk;
  int skip;              // n of Gamma_n; < 0: no inverse
  double l2;
  int tik_order;
  double tik_lambda;
  double* Winv;          // 64 x 64, stride 64
  double* scal;          // scalar block
  int status_slot;       // scal index that receives max(path)
  int status_reset;      // write path instead of max(previous, path)
};

__device__ __forceinline__ double job_recip(double p) {
  double r = __builtin_amdgcn_rcp(p);
  r = fma(r, fma(-p, r, 1.0), r);
  r = fma(r, fma(-p, r, 1.0), r);
  return r;
}

// 1/sqrt(d): hardware estimate refined by Newton steps.
#ifndef TENRSO_JOB_RSQ_STEPS
#define TENRSO_JOB_RSQ_STEPS 2
#endif
__device__ __forceinline__ double job_rsqrt(double d) {
  double y = __builtin_amdgcn_rsq(d);
#pragma unroll
  for (int i = 0; i < TENRSO_JOB_RSQ_STEPS; ++i) y = y * fma(-0.5 * d * y, y, 1.5);
  return y;
}

__device__ __forceinline__ double job_block_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

// Regularisation added to Gamma[i][j] for i, j < R: lambda D^T D of first or
// second differences (cp/helpers.rs:630-702), or l2 on the diagonal, plus ridge.
__device__ __forceinline__ double job_reg(const AlsJob& a, int i, int j, double ridge) {
  double v = 0.0;
  const int n = a.R;
  const double lam = a.tik_lambda;
  if (a.tik_order == 1) {
    if (n >= 2) {
      if (i == j) v += lam * ((i == 0 || i == n - 1) ? 1.0 : 2.0);
      else if (j == i + 1 || i == j + 1) v -= lam;
    }
  } else if (a.tik_order >= 2) {
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
  } else if (a.l2 > 0.0 && i == j) {
    v += a.l2;
  }
  if (i == j) v += ridge;
  return v;
}

typedef double doublex4 __attribute__((ext_vector_type(4)));

// Block Gauss-Jordan inverse of the 64 x 64 matrix held as 4 x 4 blocks:
// thread t owns av[u][v] = A[4bi + u][4bj + v], bi = t / 16, bj = t % 16.
// Step s eliminates the pivot block J = {4s..4s+3}, P = A[J][J]:
//   W = P^-1 Arow, Arow = A[J][:] with columns J replaced by the identity
//       (each thread solves P w = Arow[:, c] for its four columns with a 4 x 4
//       Cholesky factorisation of P);
//   pivot rows:  A[J][:] <- W;
//   other rows:  A[i][c] <- Ahat[i][c] - sum_t A[i][J_t] W[t][c],
//                Ahat = A with columns J replaced by zeros.
// This leaves A - A[:,J] P^-1 A[J,:] off the block row and column, P^-1 A[J,:]
// on the pivot rows, -A[:,J] P^-1 on the pivot columns and P^-1 on the block.
// The threads of block row s publish the pivot rows, those of block column s
// the pivot columns.  Each thread reads 16 pivot-row and 16 pivot-column values
// per step; LDS is double-buffered, so one barrier per step suffices.  Returns
// false if a Cholesky pivot is not positive and finite (the matrix is not
// positive definite).
__device__ __forceinline__ bool job_block_gj(double (&av)[4][4], double* lds) {
  const int bi = threadIdx.x >> 4, bj = threadIdx.x & 15;
  bool ok = true;
#pragma unroll 1
  for (int s = 0; s < 16; ++s) {
    const int b = s & 1;
    double* P = lds + JOB_GJ_P + b * 16;
    double* Rw = lds + JOB_GJ_ROW + b * 4 * JOB_NE;
    double* Q = lds + JOB_GJ_Q + b * 4 * JOB_NE;
    const bool prow = bi == s, pcol = bj == s;
    if (prow) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        double2 lo = make_double2(av[u][0], av[u][1]), hi = make_double2(av[u][2], av[u][3]);
        if (pcol) {
          lo = make_double2(u == 0 ? 1.0 : 0.0, u == 1 ? 1.0 : 0.0);
          hi = make_double2(u == 2 ? 1.0 : 0.0, u == 3 ? 1.0 : 0.0);
        }
        *reinterpret_cast<double2*>(&Rw[u * JOB_NE + 4 * bj]) = lo;
        *reinterpret_cast<double2*>(&Rw[u * JOB_NE + 4 * bj + 2]) = hi;
      }
      if (pcol) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) P[u * 4 + v] = av[u][v];
      }
    }
    if (pcol) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        *reinterpret_cast<double2*>(&Q[(4 * bi + u) * 4]) = make_double2(av[u][0], av[u][1]);
        *reinterpret_cast<double2*>(&Q[(4 * bi + u) * 4 + 2]) = make_double2(av[u][2], av[u][3]);
#pragma unroll
        for (int v = 0; v < 4; ++v) av[u][v] = 0.0;
      }
    }
    __syncthreads();
    TENRSO_JOB_MARK(10 + s);
    // all LDS reads of the step are issued before the Cholesky chain
    const double p00 = P[0], p10 = P[4], p11 = P[5], p20 = P[8], p21 = P[9], p22 = P[10];
    const double p30 = P[12], p31 = P[13], p32 = P[14], p33 = P[15];
    double r[4][4], q[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double2 ra = *reinterpret_cast<const double2*>(&Rw[u * JOB_NE + 4 * bj]);
      const double2 rb = *reinterpret_cast<const double2*>(&Rw[u * JOB_NE + 4 * bj + 2]);
      r[u][0] = ra.x;
      r[u][1] = ra.y;
      r[u][2] = rb.x;
      r[u][3] = rb.y;
      const double2 qa = *reinterpret_cast<const double2*>(&Q[(4 * bi + u) * 4]);
      const double2 qb = *reinterpret_cast<const double2*>(&Q[(4 * bi + u) * 4 + 2]);
      q[u][0] = qa.x;
      q[u][1] = qa.y;
      q[u][2] = qb.x;
      q[u][3] = qb.y;
    }
    // Cholesky P = L L^T (lower triangle of P), redundantly in every thread
    const double i0 = job_rsqrt(p00);
    const double l10 = p10 * i0, l20 = p20 * i0, l30 = p30 * i0;
    const double d1 = fma(-l10, l10, p11);
    const double i1 = job_rsqrt(d1);
    const double l21 = fma(-l20, l10, p21) * i1, l31 = fma(-l30, l10, p31) * i1;
    const double d2 = fma(-l21, l21, fma(-l20, l20, p22));
    const double i2 = job_rsqrt(d2);
    const double l32 = fma(-l31, l21, fma(-l30, l20, p32)) * i2;
    const double d3 = fma(-l32, l32, fma(-l31, l31, fma(-l30, l30, p33)));
    const double i3 = job_rsqrt(d3);
    ok = ok && p00 > 0.0 && d1 > 0.0 && d2 > 0.0 && d3 > 0.0 && isfinite(i0 * i1 * i2 * i3);
    TENRSO_JOB_MARK(30 + (s & 3));
    // for each of the four columns: L y = Arow[:, c], L^T w = y
    double wv[4][4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double y0 = r[0][v] * i0;
      const double y1 = fma(-l10, y0, r[1][v]) * i1;
      const double y2 = fma(-l21, y1, fma(-l20, y0, r[2][v])) * i2;
      const double y3 = fma(-l32, y2, fma(-l31, y1, fma(-l30, y0, r[3][v]))) * i3;
      wv[3][v] = y3 * i3;
      wv[2][v] = fma(-l32, wv[3][v], y2) * i2;
      wv[1][v] = fma(-l31, wv[3][v], fma(-l21, wv[2][v], y1)) * i1;
      wv[0][v] = fma(-l30, wv[3][v], fma(-l20, wv[2][v], fma(-l10, wv[1][v], y0))) * i0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double upd =
            fma(-q[u][3], wv[3][v], fma(-q[u][2], wv[2][v], fma(-q[u][1], wv[1][v], fma(-q[u][0], wv[0][v], av[u][v]))));
        av[u][v] = prow ? wv[u][v] : upd;
      }
    TENRSO_JOB_MARK(40 + (s & 3));
  }
  __syncthreads();
  return ok;
}

// The whole job; `lds` must hold JOB_LDS_BYTES (16-byte aligned).  All 256
// threads of the workgroup must call it.
__device__ __forceinline__ void als_job_run(const AlsJob& a, double* lds) {
  const int R = a.R;
  // every Gram entry a thread touches is in its 4 x 4 block (t / 16, t % 16)
  const int bi = threadIdx.x >> 4, bj = threadIdx.x & 15;
  TENRSO_JOB_MARK(0);
  // entries of G_k in this thread's block
  auto gram_block = [&](int k, double (&g)[4][4]) {
    const double* Gk = a.G + (i64)k * R * R;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 4 * bi + u, j = 4 * bj + v;
        g[u][v] = (i < R && j < R) ? Gk[i * R + j] : 0.0;
      }
  };
  if (a.do_fit) {
    // ||Xhat||^2 = sum_{r,s} prod_k G_k[r,s]; <X, Xhat> = sum of the partials
    double cross[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) cross[u][v] = 1.0;
    for (int k = 0; k < a.nd; ++k) {
      double g[4][4];
      gram_block(k, g);
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) cross[u][v] *= g[u][v];
    }
    double rec = 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) rec += cross[u][v];
    double ipv = 0.0;
    for (int b = threadIdx.x; b < a.nblk; b += JOB_THREADS) ipv += a.ip_part[b];
    const double recon = job_block_sum(rec, lds);
    const double ip = job_block_sum(ipv, lds + 8);
    if (threadIdx.x == 0) {
      const double nx = a.scal[JOB_SC_NORM_X];
      const double err_sq = nx + recon - 2.0 * ip;
      const double err = sqrt(err_sq > 0.0 ? err_sq : 0.0);
      double fit = 1.0 - err / sqrt(nx);
      fit = fit < 0.0 ? 0.0 : (fit > 1.0 ? 1.0 : fit);
      a.scal[JOB_SC_FIT] = fit;
      a.scal[JOB_SC_IP] = ip;
      a.scal[JOB_SC_RECON] = recon;
    }
    __syncthreads();
  }
  TENRSO_JOB_MARK(2);
  if (a.skip < 0) return;
  // Gamma = hadamard_{k != skip} G_k (ascending k, from ones) + regularisation,
  // identity outside R x R, in the 4 x 4 block layout of job_block_gj
  double* band = lds + JOB_GJ_REG;  // [64][8]: regularisation of entries (i, i + d), d = -2..2
  int path = 0;
  bool ok = false;
  double av[4][4];
#pragma unroll 1
  for (int attempt = 0; attempt < 2 && !ok; ++attempt) {
    const double ridge = attempt ? (double)FLT_EPSILON * (double)(10 * R) : 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) av[u][v] = 1.0;
    for (int k = 0; k < a.nd; ++k) {
      if (k == a.skip) continue;
      double g[4][4];
      gram_block(k, g);
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) av[u][v] *= g[u][v];
    }
    __syncthreads();
    if (threadIdx.x < JOB_NE) {
      const int i = threadIdx.x;
#pragma unroll 1
      for (int d = -2; d <= 2; ++d) {
        const int col = i + d;
        band[i * 8 + d + 2] = (i < R && col >= 0 && col < R) ? job_reg(a, i, col, ridge) : 0.0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 4 * bi + u, col = 4 * bj + v;
        const int d = col - i;
        const double add = (d >= -2 && d <= 2) ? band[i * 8 + d + 2] : 0.0;
        av[u][v] = (i < R && col < R) ? av[u][v] + add : (i == col ? 1.0 : 0.0);
      }
    TENRSO_JOB_MARK(3);
    ok = job_block_gj(av, lds);
    TENRSO_JOB_MARK(4);
    path = attempt;
  }
  if (!ok) path = 2;
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) a.Winv[(4 * bi + u) * JOB_NE + 4 * bj + v] = ok ? av[u][v] : 0.0;
  TENRSO_JOB_MARK(5);
  if (threadIdx.x == 0 && a.scal) {
    const double prev = a.scal[a.status_slot];
    a.scal[a.status_slot] = a.status_reset ? (double)path : fmax(prev, (double)path);
  }
}

}  // namespace tenrso
