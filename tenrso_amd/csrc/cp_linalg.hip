// cp_linalg.hip -- FP64 kernels of the ALS step on gfx950 besides the MTTKRP.
//
//   Gram           G = A^T A                                   (cp/helpers.rs:409-452)
//   ALS job        Gamma_n^{-1} and the fit as one workgroup (als_job.hpp; ranks
//                  above 64 in the wide-rank job below)
//   row solve      A_n[i,:] = M[i,:] Gamma_n^{-1} for every row (cp/helpers.rs:455-495)
//
// All of it runs in FP64 on the device; the factors stay FP32.  Every reduction
// has a fixed order (per-block partials summed in block order), so results are
// run-to-run deterministic.
#include <cfloat>

#include "als_job.hpp"
#include "cp_linalg.hpp"
#include "mttkrp.hpp"

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
// Grid-stride over X with four independent 16-byte loads per thread and step
// (VEC: X 16-byte aligned), FP64 sums of the exact fp32 squares; one (sum, min)
// partial per workgroup, reduced by norm_min_final.
template <bool VEC>
__global__ __launch_bounds__(256) void norm_min_partial(const float* __restrict__ X, i64 n, double* __restrict__ part) {
  __shared__ double red[256];
  double s0 = 0.0, s1 = 0.0;
  float mn = FLT_MAX;
  const i64 stride = (i64)gridDim.x * 256;
  const i64 g = (i64)blockIdx.x * 256 + threadIdx.x;
  auto add = [&](float x) {
    s0 = fma((double)x, (double)x, s0);
    mn = fminf(mn, x);
  };
  if constexpr (VEC) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4* X4 = reinterpret_cast<const f4*>(X);
    const i64 n4 = n / 4;
    i64 i = g;
    for (; i + 3 * stride < n4; i += 4 * stride) {
      const f4 a = X4[i], b = X4[i + stride], c = X4[i + 2 * stride], d = X4[i + 3 * stride];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        add(a[q]);
        s1 = fma((double)b[q], (double)b[q], s1);
        mn = fminf(mn, b[q]);
        add(c[q]);
        s1 = fma((double)d[q], (double)d[q], s1);
        mn = fminf(mn, d[q]);
      }
    }
    for (; i < n4; i += stride) {
      const f4 a = X4[i];
#pragma unroll
      for (int q = 0; q < 4; ++q) add(a[q]);
    }
    for (i64 j = 4 * n4 + g; j < n; j += stride) add(X[j]);
  } else {
    for (i64 i = g; i < n; i += stride) add(X[i]);
  }
  const double tot = block_sum_256(s0 + s1, red);
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

// Adds the partials; also clears scal[0, clear) first (the scalar block of a
// new run), so no separate fill launch is needed.
__global__ __launch_bounds__(256) void norm_min_final(const double* __restrict__ part, int nblk,
                                                      double* __restrict__ scal, int clear) {
  __shared__ double red[256];
  double s = 0.0, mn = DBL_MAX;
  for (int b = threadIdx.x; b < nblk; b += 256) {
    s += part[b];
    mn = fmin(mn, part[nblk + b]);
  }
  for (int i = threadIdx.x; i < clear; i += 256)
    if (i != SC_NORM_X && i != SC_MIN_X) scal[i] = 0.0;
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

int norm_min_finish(tenrso_ctx* ctx, const double* part, int nblk, double* scal, int clear) {
  hipLaunchKernelGGL(norm_min_final, dim3(1), dim3(256), 0, ctx->stream, part, nblk, scal, clear);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

int tensor_norm_min(tenrso_ctx* ctx, const float* X, i64 n, double* tmp, double* scal, int clear) {
  const int nblk = (int)std::min<i64>(1024, std::max<i64>(1, cdiv(n, 256 * 16)));
  if ((uintptr_t)X % 16 == 0)
    hipLaunchKernelGGL(norm_min_partial<true>, dim3(nblk), dim3(256), 0, ctx->stream, X, n, tmp);
  else
    hipLaunchKernelGGL(norm_min_partial<false>, dim3(nblk), dim3(256), 0, ctx->stream, X, n, tmp);
  TENRSO_HIP_TRY(hipGetLastError());
  return norm_min_finish(ctx, tmp, nblk, scal, clear);
}

// ---------------------------------------------------------------- Gram A^T A
// grid: R workgroups of 1024 threads; workgroup r computes Gram row r.  Thread
// (ch, c) sums A[i][r] A[i][c] in FP64 over row chunk ch (1024 / NC chunks: 16
// at rank <= 64, 8 at ranks up to WIDE_NE), with all of its loads independent;
// the chunk sums are added in fixed order.
template <int NC>
__device__ __forceinline__ void gram_rows_body(const float* __restrict__ A, i64 rows, int R, double* __restrict__ G) {
  constexpr int NG = 1024 / NC;
  __shared__ double part[1024];
  const int c = threadIdx.x % NC, ch = threadIdx.x / NC;
  const int r = blockIdx.x;
  double acc0 = 0.0, acc1 = 0.0;
  if (c < R) {
    const i64 i0 = (i64)ch * rows / NG, i1 = (i64)(ch + 1) * rows / NG;
    i64 i = i0;
    for (; i + 2 <= i1; i += 2) {
      acc0 += (double)A[i * R + r] * (double)A[i * R + c];
      acc1 += (double)A[(i + 1) * R + r] * (double)A[(i + 1) * R + c];
    }
    if (i < i1) acc0 += (double)A[i * R + r] * (double)A[i * R + c];
  }
  part[ch * NC + c] = acc0 + acc1;
  __syncthreads();
  if (ch == 0 && c < R) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NG; ++k) s += part[k * NC + c];
    G[(i64)r * R + c] = s;
  }
}

template <int NC>
__global__ __launch_bounds__(1024) void gram_rows_kernel(const float* __restrict__ A, i64 rows, int R,
                                                         double* __restrict__ G) {
  gram_rows_body<NC>(A, rows, R, G);
}

// Tall factors (more than GRAM_TALL rows, the sparse modes of 10^4 rows): the
// row loop above runs each thread over rows / 16 rows, one memory round trip
// after another (62-102 us at 6000-10000 rows, rank 32).  Here the rows go in
// chunks of GRAM_CH, one workgroup per chunk: the chunk is staged in LDS and
// thread t forms entries t, t + 1024, .. of the chunk's R x R partial in FP64, in
// row order; a second launch adds each entry's chunk partials in chunk order
// (16 thread groups over contiguous chunk ranges, their sums added in group
// order).  Entries (r, c) and (c, r) add the same products in the same order, so
// the Gram stays exactly symmetric.  Above rank SOLVE_RMAX a chunk holds half as
// many rows, so that it fills the same LDS.
constexpr int GRAM_CH = 128;
constexpr i64 GRAM_TALL = 1024;
__global__ __launch_bounds__(1024) void gram_chunk_kernel(const float* __restrict__ A, i64 rows, int R, int chr,
                                                          double* __restrict__ part) {
  __shared__ float sA[GRAM_CH * SOLVE_RMAX];  // chr * R <= GRAM_CH * SOLVE_RMAX
  const i64 i0 = (i64)blockIdx.x * chr;
  const int n = (int)((rows - i0) < chr ? (rows - i0) : chr);
  for (int e = threadIdx.x; e < n * R; e += 1024) sA[e] = A[i0 * R + e];
  __syncthreads();
  const int RR = R * R;
  for (int e = threadIdx.x; e < RR; e += 1024) {
    const int r = e / R, c = e - (e / R) * R;
    double a0 = 0.0, a1 = 0.0;
    int i = 0;
    for (; i + 2 <= n; i += 2) {
      a0 += (double)sA[i * R + r] * (double)sA[i * R + c];
      a1 += (double)sA[(i + 1) * R + r] * (double)sA[(i + 1) * R + c];
    }
    if (i < n) a0 += (double)sA[i * R + r] * (double)sA[i * R + c];
    part[(i64)blockIdx.x * RR + e] = a0 + a1;
  }
}

__global__ __launch_bounds__(1024) void gram_chunk_sum_kernel(const double* __restrict__ part, int nch, int RR,
                                                              double* __restrict__ G) {
  __shared__ double sq[1024];
  const int t = threadIdx.x, j = t & 63, grp = t >> 6;
  const int e = blockIdx.x * 64 + j;
  const int k0 = grp * nch / 16, k1 = (grp + 1) * nch / 16;
  double s0 = 0.0, s1 = 0.0;
  if (e < RR) {
    int k = k0;
    for (; k + 2 <= k1; k += 2) {
      s0 += part[(i64)k * RR + e];
      s1 += part[(i64)(k + 1) * RR + e];
    }
    if (k < k1) s0 += part[(i64)k * RR + e];
  }
  sq[t] = s0 + s1;
  __syncthreads();
  if (grp == 0 && e < RR) {
    double s = 0.0;
#pragma unroll
    for (int g = 0; g < 16; ++g) s += sq[g * 64 + j];
    G[e] = s;
  }
}

static int gram_tall(tenrso_ctx* ctx, const float* A, i64 rows, i64 R, double* G, hipStream_t stream) {
  // rows per chunk: the chunk (chr x R fp32) fills the same LDS at every rank
  const int chr = R <= SOLVE_RMAX ? GRAM_CH : GRAM_CH * SOLVE_RMAX / WIDE_NE;
  const i64 nch = (rows + chr - 1) / chr;
  void* pp = nullptr;
  TENRSO_TRY(ctx_buffer(ctx, SLOT_GRAM_CHUNKS, (size_t)nch * R * R * sizeof(double), &pp));
  hipLaunchKernelGGL(gram_chunk_kernel, dim3((unsigned)nch), dim3(1024), 0, stream, A, rows, (int)R, chr,
                     (double*)pp);
  TENRSO_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(gram_chunk_sum_kernel, dim3(cdiv(R * R, 64)), dim3(1024), 0, stream, (const double*)pp,
                     (int)nch, (int)(R * R), G);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

// The Grams of several factors in one launch: grid (R, count), G_k at G + k*R*R.
struct GramMany {
  const float* A[MAXD];
  i64 rows[MAXD];
};
template <int NC>
__global__ __launch_bounds__(1024) void gram_many_kernel(GramMany m, int R, double* __restrict__ G) {
  const int k = blockIdx.y;
  gram_rows_body<NC>(m.A[k], m.rows[k], R, G + (i64)k * R * R);
}

int gram_f64_many(tenrso_ctx* ctx, int count, const float* const* A, const i64* rows, i64 R, double* G) {
  if (R > WIDE_NE) {
    set_error("rank %lld exceeds the on-device Gram bound %d", (long long)R, WIDE_NE);
    return TENRSO_ERR_UNSUPPORTED;
  }
  if (R == 0 || count == 0) return TENRSO_OK;
  bool tall = false;
  for (int k = 0; k < count; ++k) tall = tall || rows[k] > GRAM_TALL;
  if (tall) {
    // the chunked form per factor (its chunk partials share one buffer, in stream order)
    for (int k = 0; k < count; ++k) TENRSO_TRY(gram_f64(ctx, A[k], rows[k], R, G + (i64)k * R * R));
    return TENRSO_OK;
  }
  GramMany m{};
  for (int k = 0; k < count; ++k) {
    m.A[k] = A[k];
    m.rows[k] = rows[k];
  }
  if (R <= SOLVE_RMAX)
    hipLaunchKernelGGL(gram_many_kernel<64>, dim3((unsigned)R, (unsigned)count), dim3(1024), 0, ctx->stream, m,
                       (int)R, G);
  else
    hipLaunchKernelGGL(gram_many_kernel<WIDE_NE>, dim3((unsigned)R, (unsigned)count), dim3(1024), 0, ctx->stream, m,
                       (int)R, G);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

int gram_f64(tenrso_ctx* ctx, const float* A, i64 rows, i64 R, double* G, hipStream_t stream) {
  if (R > WIDE_NE) {
    set_error("rank %lld exceeds the on-device Gram bound %d", (long long)R, WIDE_NE);
    return TENRSO_ERR_UNSUPPORTED;
  }
  if (R == 0) return TENRSO_OK;
  if (rows > GRAM_TALL) return gram_tall(ctx, A, rows, R, G, stream ? stream : ctx->stream);
  if (R <= SOLVE_RMAX)
    hipLaunchKernelGGL(gram_rows_kernel<64>, dim3((unsigned)R), dim3(1024), 0, stream ? stream : ctx->stream, A, rows,
                       (int)R, G);
  else
    hipLaunchKernelGGL(gram_rows_kernel<WIDE_NE>, dim3((unsigned)R), dim3(1024), 0, stream ? stream : ctx->stream, A,
                       rows, (int)R, G);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

// ---------------------------------------------------------------- ALS job
__global__ __launch_bounds__(JOB_THREADS) void als_job_kernel(AlsJob job) {
  __shared__ __attribute__((aligned(16))) double lds[JOB_LDS_BYTES / 8];
  als_job_run(job, lds);
}

// ---------------------------------------------------------------- wide-rank job
// Ranks SOLVE_RMAX + 1 .. WIDE_NE: the steps of the job (the fit, Gamma_n plus
// regularisation, its inverse with the ridge retry of cp/helpers.rs:477-490) as a
// kernel of 1024 threads.  Thread t holds, in registers, column t % 128 of rows
// t / 128 + 8 m (m = 0..15) of the 128 x 128 matrix, the identity outside R x R.
// The inverse is Gauss-Jordan with scalar pivots in diagonal order: step k
// publishes row k and column k (double-buffered in LDS, one barrier a step) and
// every thread updates its entries from them.  For a symmetric matrix those
// pivots are the ratios of consecutive leading principal minors, so all of them
// are positive exactly when the matrix is positive definite: a pivot that is not
// positive (or not finite) fails the attempt, as a pivot block that is not
// positive definite fails it in the job of rank <= 64.
constexpr int WIDE_THREADS = 1024;
constexpr int WIDE_RPT = WIDE_NE * WIDE_NE / WIDE_THREADS;  // entries per thread

__device__ __forceinline__ double wide_block_sum(double v, double* red) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < WIDE_THREADS / 64; ++w) s += red[w];
  return s;
}

// 4 x 4 inverse by scalar Gauss-Jordan in registers (every index fixed at compile
// time); false if a pivot is not positive and finite.
constexpr int WB = 4;  // pivots per elimination step of the wide-rank inverse
__device__ __forceinline__ bool wide_inv4(double (&P)[WB][WB]) {
  bool ok = true;
#pragma unroll
  for (int k = 0; k < WB; ++k) {
    const double p = P[k][k];
    ok = ok && p > 0.0 && p <= DBL_MAX;
    const double piv = job_recip(p);
    double rj[WB], cc[WB];
#pragma unroll
    for (int q = 0; q < WB; ++q) {
      rj[q] = q == k ? 0.0 : P[k][q] * piv;
      cc[q] = P[q][k];
    }
#pragma unroll
    for (int r = 0; r < WB; ++r)
#pragma unroll
      for (int q = 0; q < WB; ++q) P[r][q] = fma(-cc[r], rj[q], P[r][q]);
#pragma unroll
    for (int q = 0; q < WB; ++q) {
      if (q != k) {
        P[k][q] = rj[q];
        P[q][k] = -cc[q] * piv;
      }
    }
    P[k][k] = piv;
  }
  return ok;
}

// In-place inverse of the matrix held as described above; false if a pivot
// block is not positive definite.  Every thread returns at the same step.  Block
// Gauss-Jordan with 4 x 4 pivot blocks J = {k0..k0+3} (the order padded to a
// multiple of 4 with the identity), one barrier a step.  With P^-1 = A[J][J]^-1
// and W = P^-1 A[J][:] (taken as 0 on the columns J):
//   A[i][j] -= sum_q A[i][J_q] W[q][j]    every entry (columns J stay as they were),
//   A[J][j]  = W[:, j]                    rows J (the four waves' rows holding them),
//   A[i][J]  = -A[i][J] P^-1, A[J][J] = P^-1   columns J (four threads per row group).
// Every thread inverts P itself from the published rows.  The pivot blocks are
// positive definite exactly when the matrix is (their leading minors are ratios
// of the matrix's), so the scalar pivots inside them decide the ridge retry.
// Rows J are published as rk[r][j]; columns J as ck[((i % 8) * 16 + i / 8) * 4 + q],
// each thread's 64 values contiguous.
__device__ bool wide_gauss_jordan(double (&a)[WIDE_RPT], int R, double* rowb, double* colb) {
  const int j = threadIdx.x & (WIDE_NE - 1), i0 = threadIdx.x >> 7;
  const int Rp = (R + WB - 1) / WB * WB;
  for (int k0 = 0; k0 < Rp; k0 += WB) {
    double* rk = rowb + ((k0 / WB) & 1) * WB * WIDE_NE;
    double* ck = colb + ((k0 / WB) & 1) * WB * WIDE_NE;
    const int mk = k0 >> 3;          // rows k0..k0+3 are entry mk of their threads
    const int r_own = i0 - (k0 & 7);
    const bool row_owner = r_own >= 0 && r_own < WB;  // uniform per wave
    const int q_own = j - k0;
    const bool col_owner = q_own >= 0 && q_own < WB;
    if (row_owner) {
      // row k0 + r_own, selected without a run-time register index
      double v = 0.0;
#pragma unroll
      for (int m = 0; m < WIDE_RPT; ++m) v = m == mk ? a[m] : v;
      rk[r_own * WIDE_NE + j] = v;
    }
    if (col_owner) {
#pragma unroll
      for (int m = 0; m < WIDE_RPT; ++m) ck[(i0 * WIDE_RPT + m) * WB + q_own] = a[m];
    }
    __syncthreads();
    double P[WB][WB], rr[WB];
#pragma unroll
    for (int r = 0; r < WB; ++r) {
#pragma unroll
      for (int q = 0; q < WB; ++q) P[r][q] = rk[r * WIDE_NE + k0 + q];
      rr[r] = rk[r * WIDE_NE + j];
    }
    if (!wide_inv4(P)) {
      __syncthreads();  // every thread has read the buffers before they are reused
      return false;
    }
    // one pass of four FMAs per entry: columns J start from 0 and take column
    // q_own of P^-1 (A[i][J] -> -A[i][J] P^-1), the others start from A[i][j] and
    // take W[:, j] = P^-1 A[J][j]
    double X[WB];
#pragma unroll
    for (int q = 0; q < WB; ++q) {
      double w = 0.0;
#pragma unroll
      for (int r = 0; r < WB; ++r) w = fma(P[q][r], rr[r], w);
      const double pq = q_own == 0 ? P[q][0] : (q_own == 1 ? P[q][1] : (q_own == 2 ? P[q][2] : P[q][3]));
      X[q] = col_owner ? pq : w;
    }
    const double* cl = ck + i0 * WIDE_RPT * WB;
#pragma unroll
    for (int m = 0; m < WIDE_RPT; ++m) {
      double v = col_owner ? 0.0 : a[m];
#pragma unroll
      for (int q = 0; q < WB; ++q) v = fma(-cl[m * WB + q], X[q], v);
      a[m] = v;
    }
    if (row_owner) {
      // rows J: W[r_own][j], or P^-1[r_own][q_own] on the columns J
      const double xo = r_own == 0 ? X[0] : (r_own == 1 ? X[1] : (r_own == 2 ? X[2] : X[3]));
#pragma unroll
      for (int m = 0; m < WIDE_RPT; ++m) a[m] = m == mk ? xo : a[m];
    }
  }
  return true;
}

__global__ __launch_bounds__(WIDE_THREADS) void als_job_wide_kernel(AlsJob a) {
  __shared__ double rowb[2 * WB * WIDE_NE], colb[2 * WB * WIDE_NE];
  __shared__ double red0[WIDE_THREADS / 64], red1[WIDE_THREADS / 64];
  const int R = a.R;
  const int j = threadIdx.x & (WIDE_NE - 1), i0 = threadIdx.x >> 7;
  double ipv = 0.0;
  if (a.do_fit)
    for (int b = threadIdx.x; b < a.nblk; b += WIDE_THREADS) ipv += a.ip_part[b];
  // two passes over the Grams, each with one accumulator per entry in registers:
  // the fit's hadamard of all k, then Gamma's hadamard of k != skip
  auto hadamard = [&](double (&h)[WIDE_RPT], int skip) {
#pragma unroll
    for (int m = 0; m < WIDE_RPT; ++m) h[m] = 1.0;
    for (int k = 0; k < a.nd; ++k) {
      if (k == skip) continue;
      const double* Gk = a.G + (i64)k * R * R + j;
#pragma unroll
      for (int m = 0; m < WIDE_RPT; ++m) {
        const int i = i0 + 8 * m;
        h[m] *= (i < R && j < R) ? Gk[i * R] : 0.0;
      }
    }
  };
  if (a.do_fit) {
    // ||Xhat||^2 = sum_{r,s} prod_k G_k[r,s] (entries outside R x R are 0)
    double cross[WIDE_RPT];
    hadamard(cross, -1);
    double rec = 0.0;
#pragma unroll
    for (int m = 0; m < WIDE_RPT; ++m) rec += cross[m];
    const double recon = wide_block_sum(rec, red0);
    const double ip = wide_block_sum(ipv, red1);
    if (threadIdx.x == 0) {
      double blk[JOB_SC_STAMP];
#pragma unroll
      for (int i = 0; i < JOB_SC_STAMP; ++i) blk[i] = a.scal[i];
      const double nx = blk[JOB_SC_NORM_X];
      const double err_sq = nx + recon - 2.0 * ip;
      const double err = sqrt(err_sq > 0.0 ? err_sq : 0.0);
      double fit = 1.0 - err / sqrt(nx);
      fit = fit < 0.0 ? 0.0 : (fit > 1.0 ? 1.0 : fit);
      blk[JOB_SC_FIT] = fit;
      blk[JOB_SC_IP] = ip;
      blk[JOB_SC_RECON] = recon;
      a.scal[JOB_SC_FIT] = fit;
      a.scal[JOB_SC_IP] = ip;
      a.scal[JOB_SC_RECON] = recon;
      if (a.host) {
#pragma unroll
        for (int i = 0; i < JOB_SC_STAMP; ++i) a.host[i] = blk[i];
        __threadfence_system();
        a.host[JOB_SC_STAMP] = a.stamp;
      }
    }
  }
  double gam[WIDE_RPT];
  if (a.skip >= 0) hadamard(gam, a.skip);
  if (a.gate) {
    __syncthreads();  // every thread has its Grams
    if (threadIdx.x == 0) __hip_atomic_store(a.gate, a.gate_stamp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (a.skip < 0) return;
  // Gamma waits in Winv for the ridge retry (each thread rereads its own entries)
  double w[WIDE_RPT];
  auto form = [&](double ridge) {
#pragma unroll
    for (int m = 0; m < WIDE_RPT; ++m) {
      const int i = i0 + 8 * m;
      const double g = a.Winv[i * WIDE_NE + j];
      w[m] = (i < R && j < R) ? g + job_reg(a, i, j, ridge) : (i == j ? 1.0 : 0.0);
    }
  };
#pragma unroll
  for (int m = 0; m < WIDE_RPT; ++m) a.Winv[(i0 + 8 * m) * WIDE_NE + j] = gam[m];
  form(0.0);
  bool ok = wide_gauss_jordan(w, R, rowb, colb);
  int path = 0;
  if (!ok) {
    form((double)FLT_EPSILON * (double)(10 * R));
    ok = wide_gauss_jordan(w, R, rowb, colb);
    path = ok ? 1 : 2;
  }
#pragma unroll
  for (int m = 0; m < WIDE_RPT; ++m) a.Winv[(i0 + 8 * m) * WIDE_NE + j] = ok ? w[m] : 0.0;
  if (threadIdx.x == 0 && a.scal) {
    const double prev = a.scal[a.status_slot];
    a.scal[a.status_slot] = a.status_reset ? (double)path : fmax(prev, (double)path);
  }
}

int job_join(tenrso_ctx* ctx) {
  if (!ctx->job_join_pending) return TENRSO_OK;
  ctx->job_join_pending = false;
  TENRSO_HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->job_join, 0));
  return TENRSO_OK;
}

int als_job_launch(tenrso_ctx* ctx, const AlsJob& job) {
  if (job.R > WIDE_NE) {
    set_error("rank %d exceeds the on-device normal-equation solve bound %d", job.R, WIDE_NE);
    return TENRSO_ERR_UNSUPPORTED;
  }
  if (job.R > JOB_NE && job.skip >= 0) {
    // The wide-rank job with a row solve behind it runs on the side stream, after
    // everything queued so far, beside the MTTKRP queued after it (the two are
    // independent); solve_rows joins it before reading Winv.
    if (!ctx->job_stream) {
      TENRSO_HIP_TRY(hipStreamCreateWithFlags(&ctx->job_stream, hipStreamNonBlocking));
      TENRSO_HIP_TRY(hipEventCreateWithFlags(&ctx->job_fork, hipEventDisableTiming));
      TENRSO_HIP_TRY(hipEventCreateWithFlags(&ctx->job_join, hipEventDisableTiming));
    }
    TENRSO_TRY(job_join(ctx));
    TENRSO_HIP_TRY(hipEventRecord(ctx->job_fork, ctx->stream));
    TENRSO_HIP_TRY(hipStreamWaitEvent(ctx->job_stream, ctx->job_fork, 0));
    hipLaunchKernelGGL(als_job_wide_kernel, dim3(1), dim3(WIDE_THREADS), 0, ctx->job_stream, job);
    TENRSO_HIP_TRY(hipGetLastError());
    TENRSO_HIP_TRY(hipEventRecord(ctx->job_join, ctx->job_stream));
    ctx->job_join_pending = true;
    return TENRSO_OK;
  }
  TENRSO_TRY(job_join(ctx));
  if (job.R > JOB_NE)
    hipLaunchKernelGGL(als_job_wide_kernel, dim3(1), dim3(WIDE_THREADS), 0, ctx->stream, job);
  else
    hipLaunchKernelGGL(als_job_kernel, dim3(1), dim3(JOB_THREADS), 0, ctx->stream, job);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

// ---------------------------------------------------------------- row solve
// grid: rows (one workgroup per row); 256 threads = 4 split quarters x 64
// columns.  Each thread sums a quarter of the MTTKRP split partials of its
// column (all loads in flight, FP64), the quarters are added in fixed order,
// then A[i, c] = epi(sum_r M[i, r] Winv[r, c]) (the r-sum split over the same
// four thread groups) and the block's <M, A> partial is written.
constexpr int NE = JOB_NE;  // stride of Winv
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
  // this thread's quarter of Winv column c, loaded beside the partials (it does
  // not depend on them): one memory round trip instead of two
  const int r0 = qd * R / 4, r1 = (qd + 1) * R / 4;
  double w[NE / 4];
#pragma unroll
  for (int k = 0; k < NE / 4; ++k) w[k] = (colok && r0 + k < r1) ? Winv[(r0 + k) * NE + c] : 0.0;
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
  // x[c] = sum_r M[r] Winv[r, c], r split over the four quarters (two
  // alternating sums per thread, as before the prefetch)
  double x = 0.0;
  if (colok) {
    double x0 = 0, x1 = 0;
#pragma unroll
    for (int k = 0; k < NE / 4; k += 2) {
      if (r0 + k < r1) x0 += sM[r0 + k] * w[k];
      if (r0 + k + 1 < r1) x1 += sM[r0 + k + 1] * w[k + 1];
    }
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

// Ranks SOLVE_RMAX + 1 .. WIDE_NE: the same row solve with Winv of stride
// WIDE_NE; 256 threads = 2 halves x 128 columns (each half sums half of the
// split partials, then half of the r-sum), one row per workgroup.
__global__ __launch_bounds__(256) void solve_rows_wide_kernel(const float* __restrict__ part, int splits,
                                                              const double* __restrict__ M64, i64 rows, int R,
                                                              const double* __restrict__ Winv, double l1, int nonneg,
                                                              float* __restrict__ A, double* __restrict__ ip_part) {
  __shared__ double sq[2 * WIDE_NE];
  __shared__ double sM[WIDE_NE];
  __shared__ double red[2];
  const int c = threadIdx.x & (WIDE_NE - 1), h = threadIdx.x >> 7;
  const i64 row = blockIdx.x;
  const bool colok = c < R;
  double m = 0.0;
  if (colok) {
    if (part) {
      const i64 n = rows * R, off = row * R + c;
      const int s0 = h * splits / 2, s1 = (h + 1) * splits / 2;
      double a0 = 0, a1 = 0;
      int sp = s0;
      for (; sp + 2 <= s1; sp += 2) {
        a0 += (double)part[(i64)sp * n + off];
        a1 += (double)part[(i64)(sp + 1) * n + off];
      }
      if (sp < s1) a0 += (double)part[(i64)sp * n + off];
      m = a0 + a1;
    } else if (h == 0) {
      m = M64[row * R + c];
    }
  }
  sq[h * WIDE_NE + c] = m;
  __syncthreads();
  if (h == 0) sM[c] = sq[c] + sq[WIDE_NE + c];
  __syncthreads();
  double x = 0.0;
  if (colok) {
    const int r0 = h * R / 2, r1 = (h + 1) * R / 2;
    double x0 = 0, x1 = 0;
    int r = r0;
    for (; r + 2 <= r1; r += 2) {
      x0 += sM[r] * Winv[(i64)r * WIDE_NE + c];
      x1 += sM[r + 1] * Winv[(i64)(r + 1) * WIDE_NE + c];
    }
    if (r < r1) x0 += sM[r] * Winv[(i64)r * WIDE_NE + c];
    x = x0 + x1;
  }
  sq[h * WIDE_NE + c] = x;  // the reads of sq above finished before the last barrier
  __syncthreads();
  double ip = 0.0;
  if (h == 0 && colok) {
    double xv = sq[c] + sq[WIDE_NE + c];
    if (l1 > 0.0) xv = xv > l1 ? xv - l1 : (xv < -l1 ? xv + l1 : 0.0);  // soft_threshold (helpers.rs:610-623)
    if (nonneg) xv = xv > 0.0 ? xv : 0.0;                                  // core.rs:357-360
    const float x32 = (float)xv;
    A[row * R + c] = x32;
    ip = sM[c] * (double)x32;
  }
  if (h == 0) {
    ip = wave_sum(ip);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ip;
  }
  __syncthreads();
  if (threadIdx.x == 0 && ip_part) ip_part[blockIdx.x] = red[0] + red[1];
}

int solve_rows(tenrso_ctx* ctx, const float* part, int splits, const double* M, i64 rows, i64 R,
               const double* Winv, const RowEpilogue& epi, float* A, double* ip_part, int* nblk_out) {
  const unsigned nblk = cdiv(rows, SOLVE_RB);
  if (nblk_out) *nblk_out = (int)nblk;
  if (rows == 0) return TENRSO_OK;
  if (R > WIDE_NE) {
    set_error("rank %lld exceeds the on-device solve bound %d", (long long)R, WIDE_NE);
    return TENRSO_ERR_UNSUPPORTED;
  }
  TENRSO_TRY(job_join(ctx));  // Winv of a wide-rank job on the side stream
  if (R > SOLVE_RMAX)
    hipLaunchKernelGGL(solve_rows_wide_kernel, dim3(nblk), dim3(256), 0, ctx->stream, part, splits, M, rows, (int)R,
                       Winv, epi.l1, epi.nonneg, A, ip_part);
  else
    hipLaunchKernelGGL(solve_rows_kernel, dim3(nblk), dim3(256), 0, ctx->stream, part, splits, M, rows, (int)R, Winv,
                       epi.l1, epi.nonneg, A, ip_part);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

// ---------------------------------------------------------------- orthonormalisation
// Q of the QR factorisation of a rows x R fp32 factor (orthonormalize_factor,
// cp/helpers.rs:582-607), one workgroup, FP64: Householder reflections
// H_k = I - 2 v v^T / (v^T v), v = x - alpha e_k, alpha = -sign(x_k) ||x||
// (the oracle's restatement), then Q = H_0 ... H_{R-1} [I; 0], with column k
// negated where alpha < 0 so that R has a positive diagonal.  A column that is
// already zero below and on the diagonal takes no reflection (rank-deficient
// factors get an orthonormal completion, as Householder QR gives).
// W: 2 * rows * R doubles of scratch (the reduced matrix with the reflectors
// below its diagonal, then Q).  Thread t works on column t % NC and row slice
// t / NC (4 slices; NC = 64, or 128 above rank 64); partial dot products are
// added in slice order.
template <int NC>
__global__ __launch_bounds__(4 * NC) void householder_q_kernel(float* __restrict__ A, i64 rows, int R,
                                                               double* __restrict__ W) {
  constexpr int NT = 4 * NC;  // 4 row slices x NC columns
  __shared__ double red[4][NC];
  __shared__ double dots[NC];
  __shared__ double v0s[NC], vns[NC], sgn[NC];
  const int tid = threadIdx.x, col = tid % NC, slice = tid / NC;
  double* Q = W + rows * R;
  const i64 n = rows * R;
  for (i64 e = tid; e < n; e += NT) W[e] = (double)A[e];
  __syncthreads();
  const int kmax = (int)(rows < R ? rows : R);
  // row range of this slice within [lo, rows)
  auto slice_rows = [&](i64 lo, i64& r0, i64& r1) {
    const i64 len = rows - lo;
    r0 = lo + len * slice / 4;
    r1 = lo + len * (slice + 1) / 4;
  };
  for (int k = 0; k < kmax; ++k) {
    i64 r0, r1;
    slice_rows(k, r0, r1);
    // ||x||^2, x = W[k:, k]
    double p = 0.0;
    if (col == 0)
      for (i64 i = r0; i < r1; ++i) p = fma(W[i * R + k], W[i * R + k], p);
    if (col == 0) red[slice][0] = p;
    __syncthreads();
    const double nrm2 = (red[0][0] + red[1][0]) + (red[2][0] + red[3][0]);
    const double a0 = W[(i64)k * R + k];
    const double nrm = sqrt(nrm2);
    const double alpha = a0 > 0.0 ? -nrm : nrm;
    const double v0 = a0 - alpha;
    const double vn = (nrm2 - a0 * a0) + v0 * v0;
    if (tid == 0) {
      v0s[k] = v0;
      vns[k] = vn;
      sgn[k] = alpha < 0.0 ? -1.0 : 1.0;
    }
    // d_j = v^T W[:, j] for j > k (v_k = v0, v_i = W[i][k] below)
    double d = 0.0;
    if (vn > 0.0 && col > k && col < R) {
      for (i64 i = r0; i < r1; ++i) {
        const double vi = (i == k) ? v0 : W[i * R + k];
        d = fma(vi, W[i * R + col], d);
      }
    }
    __syncthreads();  // red[.][0] read by everyone before it is reused
    red[slice][col] = d;
    __syncthreads();
    if (slice == 0) dots[col] = 2.0 * ((red[0][col] + red[1][col]) + (red[2][col] + red[3][col])) / (vn > 0.0 ? vn : 1.0);
    __syncthreads();
    if (vn > 0.0 && col > k && col < R) {
      const double dj = dots[col];
      for (i64 i = r0; i < r1; ++i) {
        const double vi = (i == k) ? v0 : W[i * R + k];
        W[i * R + col] -= dj * vi;
      }
    }
    __syncthreads();
  }
  // Q = H_0 ... H_{kmax-1} applied to the first R columns of the identity
  for (i64 e = tid; e < n; e += NT) Q[e] = (e / R == e % R) ? 1.0 : 0.0;
  __syncthreads();
  for (int k = kmax - 1; k >= 0; --k) {
    const double vn = vns[k], v0 = v0s[k];
    if (!(vn > 0.0)) continue;  // uniform across the workgroup
    i64 r0, r1;
    slice_rows(k, r0, r1);
    double d = 0.0;
    if (col >= k && col < R) {
      for (i64 i = r0; i < r1; ++i) {
        const double vi = (i == k) ? v0 : W[i * R + k];
        d = fma(vi, Q[i * R + col], d);
      }
    }
    red[slice][col] = d;
    __syncthreads();
    if (slice == 0) dots[col] = 2.0 * ((red[0][col] + red[1][col]) + (red[2][col] + red[3][col])) / vn;
    __syncthreads();
    if (col >= k && col < R) {
      const double dj = dots[col];
      for (i64 i = r0; i < r1; ++i) {
        const double vi = (i == k) ? v0 : W[i * R + k];
        Q[i * R + col] -= dj * vi;
      }
    }
    __syncthreads();
  }
  for (i64 e = tid; e < n; e += NT) {
    const int j = (int)(e % R);
    A[e] = (float)(Q[e] * (j < kmax ? sgn[j] : 1.0));
  }
}

int orthonormalize_rows(tenrso_ctx* ctx, float* A, i64 rows, i64 R) {
  if (R > WIDE_NE) {
    set_error("rank %lld exceeds the on-device orthonormalisation bound %d", (long long)R, WIDE_NE);
    return TENRSO_ERR_UNSUPPORTED;
  }
  if (rows == 0 || R == 0) return TENRSO_OK;
  void* pw = nullptr;
  TENRSO_TRY(ctx_buffer(ctx, SLOT_ORTHO, (size_t)2 * rows * R * sizeof(double), &pw));
  if (R <= SOLVE_RMAX)
    hipLaunchKernelGGL(householder_q_kernel<64>, dim3(1), dim3(256), 0, ctx->stream, A, rows, (int)R, (double*)pw);
  else
    hipLaunchKernelGGL(householder_q_kernel<WIDE_NE>, dim3(1), dim3(4 * WIDE_NE), 0, ctx->stream, A, rows, (int)R,
                       (double*)pw);
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
