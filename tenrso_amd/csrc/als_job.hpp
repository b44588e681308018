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
// GJ (doubles): [2][4][64] pivot rows, [2][4][4] pivot-block inverses, then
// one int flag.
constexpr int JOB_GJ_ROW = 0, JOB_GJ_PINV = 2 * 4 * JOB_NE, JOB_GJ_FLAG = JOB_GJ_PINV + 2 * 16;
constexpr int JOB_LDS_BYTES = (JOB_GJ_FLAG + 2) * 8;  // LDS bytes of the job

// Device scalar block slots shared with the ALS driver.
constexpr int JOB_SC_NORM_X = 0, JOB_SC_FIT = 1, JOB_SC_IP = 2, JOB_SC_RECON = 3, JOB_SC_STATUS = 4;
constexpr int JOB_SC_STAMP = 7;  // host copy only: index of the sweep whose fit it holds
constexpr int JOB_SC_COUNT = 8;  // size of the scalar block

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
  double* Winv;          // 64 x 64, stride 64 (128 x 128, stride 128, in the wide-rank job)
  double* scal;          // scalar block
  int status_slot;       // scal index that receives max(path)
  int status_reset;      // write path instead of max(previous, path)
  double* host;          // with do_fit: mapped host copy of scal[0..JOB_SC_COUNT) (or null)
  double stamp;          // written last into host[JOB_SC_STAMP]
  unsigned* gate;        // or null: set to gate_stamp once the Gram loads have returned
  unsigned gate_stamp;
};

// Polls of the start gate by an MTTKRP workgroup beside the job (each poll is
// a memory round trip, then a short sleep): the wait is bounded, so a gate
// that never opens delays the launch by a few microseconds and nothing else.
constexpr int JOB_GATE_POLLS = 12;

__device__ __forceinline__ double job_recip(double p) {
  double r = __builtin_amdgcn_rcp(p);
  r = fma(r, fma(-p, r, 1.0), r);
  r = fma(r, fma(-p, r, 1.0), r);
  return r;
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

// v from the lane of the same quad given by the quad permutation `ctrl`
// (0xB1: lane ^ 1, 0x4E: lane ^ 2), one DPP move per 32-bit half.
template <int CTRL>
__device__ __forceinline__ double job_quad_perm(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Matrix layout of the job: wave w (of 4) holds rows 16w..16w+15 of the
// 64 x 64 matrix as four 16 x 16 tiles in the accumulator layout of
// v_mfma_f64_16x16x4f64: lane l, tile t, element e holds
//   A[16w + 4e + l/16][16t + l%16].
// In that layout rows 4m..4m+3 of a tile (element e = m) are exactly a 4 x 16
// B operand of the same instruction (lane l holds B[l/16][l%16]).
__device__ __forceinline__ int job_row(int e) { return 16 * (threadIdx.x >> 6) + 4 * e + ((threadIdx.x & 63) >> 4); }
__device__ __forceinline__ int job_col(int t) { return 16 * t + (threadIdx.x & 15); }

// Block Gauss-Jordan inverse (4 x 4 pivot blocks) of a symmetric positive
// definite 64 x 64 matrix in the job layout, on the FP64 matrix cores.
// Step s eliminates J = {4s..4s+3}, P = A[J][J]:
//   W = P^-1 Arow', Arow' = A[J][:] with columns J replaced by the identity;
//   A <- Ahat - A[:,J] W, Ahat = A with rows J and columns J zeroed, and
//   with A[J][J] taken as -I in that product,
// which leaves W on the pivot rows, -A[:,J] P^-1 on the pivot columns, P^-1
// on the block and A - A[:,J] P^-1 A[J,:] elsewhere.  The matrix stays
// symmetric up to sign (A[i][J] = A[J][i] for rows i not yet eliminated,
// -A[J][i] for eliminated ones), so the pivot column operand is read from the
// published pivot rows.
// Per step, after the barrier, every wave forms W for all 64 columns with four
// v_mfma_f64_4x4x4f64 (block b of tile t = columns 16t + 4b..+3; lane l holds
// A[l&3][l>>4], B[l>>4][l&3], D[l>>4][l&3], so D comes out in the B operand
// layout of the 16x16x4 instruction) and updates its four tiles with four
// v_mfma_f64_16x16x4f64, the tile holding the next pivot block first.  The
// wave holding the next pivot rows publishes them (one 8-byte store per lane
// and tile) and 16 of its lanes invert the next P (adjugate, below) while the
// remaining MFMAs run.  LDS is double-buffered: one barrier per step.  Returns
// false if a pivot block is not positive definite.
__device__ __forceinline__ bool job_block_gj(doublex4 (&acc)[4], double* lds) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int q = lane >> 4, j = lane & 15;
  int* flag = reinterpret_cast<int*>(lds + JOB_GJ_FLAG);
  if (threadIdx.x == 0) *flag = 0;
  // Inverse of the pivot block P = A[J][J] by the wave holding rows J, from the
  // published rows: lane (r, c) = (lane / 4 % 4, lane % 4) forms the cofactor
  // C[r][c] from its 3 x 3 minor (nine LDS reads at offsets fixed per lane);
  // det = sum_c P[r][c] C[r][c], summed over the lane's quad; lanes 0..15 write
  // P^-1[c][r] = C[r][c] / det.  Positive definiteness is Sylvester's criterion:
  // p00, p00 p11 - p01^2, C[3][3] (the leading 3 x 3 minor) and det all > 0
  // (NaN fails every comparison).  Sixteen lanes share what one lane did
  // serially (a 4 x 4 Schur-complement inverse, about 50 dependent FP64 steps).
  const int lr = (lane >> 2) & 3, lc = lane & 3;
  int moff[9];
  {
    const int rr[3] = {lr == 0 ? 1 : 0, lr <= 1 ? 2 : 1, lr <= 2 ? 3 : 2};
    const int cc[3] = {lc == 0 ? 1 : 0, lc <= 1 ? 2 : 1, lc <= 2 ? 3 : 2};
#pragma unroll
    for (int u = 0; u < 3; ++u)
#pragma unroll
      for (int v = 0; v < 3; ++v) moff[3 * u + v] = rr[u] * JOB_NE + cc[v];
  }
  const int self_off = lr * JOB_NE + lc;
  const double csign = ((lr + lc) & 1) ? -1.0 : 1.0;
  auto invert_pivot = [&](int s) {
    const double* Pb = lds + JOB_GJ_ROW + (s & 1) * 4 * JOB_NE + 4 * s;  // P[u][v] = Pb[u * JOB_NE + v]
    double* Pi = lds + JOB_GJ_PINV + (s & 1) * 16;
    double m[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) m[k] = Pb[moff[k]];
    const double pself = Pb[self_off];
    const double p00 = Pb[0], p01 = Pb[1], p11 = Pb[JOB_NE + 1];
    const double d3 = m[0] * fma(m[4], m[8], -(m[5] * m[7])) - m[1] * fma(m[3], m[8], -(m[5] * m[6])) +
                      m[2] * fma(m[3], m[7], -(m[4] * m[6]));
    const double cof = csign * d3;
    double det = pself * cof;
    det += job_quad_perm<0xB1>(det);
    det += job_quad_perm<0x4E>(det);
    const double det2 = fma(p00, p11, -(p01 * p01));
    bool bad = !(p00 > 0.0 && det2 > 0.0 && det > 0.0 && det <= DBL_MAX);
    if (lane == 15) bad = bad || !(cof > 0.0);
    const double rd = job_recip(det);
    if (lane < 16) {
      Pi[lc * 4 + lr] = cof * rd;
      if (bad) *flag = 1;
    }
  };
  // step 0's pivot rows and inverse
  if (wave == 0) {
#pragma unroll
    for (int t = 0; t < 4; ++t) lds[JOB_GJ_ROW + q * JOB_NE + 16 * t + j] = acc[t][0];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    invert_pivot(0);
  }
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int b = s & 1, T = s >> 2, e = s & 3;
    const double* Rw = lds + JOB_GJ_ROW + b * 4 * JOB_NE;
    const double* Pi = lds + JOB_GJ_PINV + b * 16;
    __syncthreads();
    TENRSO_JOB_MARK(10 + s);
    const double pa = Pi[(lane & 3) * 4 + q];  // P^-1[l & 3][l >> 4]
    double rb[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) rb[t] = Rw[q * JOB_NE + 16 * t + j];  // Arow[q][16t + j]
    const double rown = Rw[q * JOB_NE + 16 * wave + j];            // Arow[q][16 wave + j]
    // Arow': identity on the pivot columns
    const bool pj = (j >> 2) == e;  // column 16T + j is a pivot column of tile T
    rb[T] = pj ? ((j & 3) == q ? 1.0 : 0.0) : rb[T];
    // W[q][16t + j] for the four tiles
    double wv[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) wv[t] = __builtin_amdgcn_mfma_f64_4x4x4f64(pa, rb[t], 0.0, 0, 0, 0);
    // A operand: -A[16 wave + j][J_q] (rows of this wave), -I -> +I on the pivot rows
    const int c = 16 * wave + j;
    double aop;
    if (wave == T && pj) aop = (j & 3) == q ? 1.0 : 0.0;
    else aop = c < 4 * s ? rown : -rown;
    // Ahat: zero the pivot columns (tile T) and, in wave T, the pivot rows
    if (pj) acc[T] = doublex4{0.0, 0.0, 0.0, 0.0};
    if (wave == T) {
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[t][e] = 0.0;
    }
    if (s + 1 < 16) {
      const int s1 = s + 1, T1 = s1 >> 2, e1 = s1 & 3;
      double* Rn = lds + JOB_GJ_ROW + (s1 & 1) * 4 * JOB_NE;
      // the tile holding the next pivot block first
      acc[T1] = __builtin_amdgcn_mfma_f64_16x16x4f64(aop, wv[T1], acc[T1], 0, 0, 0);
#pragma unroll
      for (int i = 1; i < 4; ++i) {
        const int t = (T1 + i) & 3;
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(aop, wv[t], acc[t], 0, 0, 0);
      }
      if (wave == T1) {
        Rn[q * JOB_NE + 16 * T1 + j] = acc[T1][e1];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        invert_pivot(s1);
#pragma unroll
        for (int i = 1; i < 4; ++i) {
          const int t = (T1 + i) & 3;
          Rn[q * JOB_NE + 16 * t + j] = acc[t][e1];
        }
      }
    } else {
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(aop, wv[t], acc[t], 0, 0, 0);
    }
    TENRSO_JOB_MARK(40 + (s & 3));
  }
  __syncthreads();
  const bool ok = *flag == 0;
  __syncthreads();  // every thread has read the flag before a retry resets it
  return ok;
}

// The whole job; `lds` must hold JOB_LDS_BYTES (16-byte aligned).  All 256
// threads of the workgroup must call it.
__device__ __forceinline__ void als_job_run(const AlsJob& a, double* lds) {
  const int R = a.R;
  // entries of this thread: (job_row(e), job_col(t)), t, e = 0..3
  TENRSO_JOB_MARK(0);
  const bool want_fit = a.do_fit != 0;
  // <X, Xhat> partials: loads issued ahead of the Gram loads
  double ipv = 0.0;
  if (want_fit)
    for (int b = threadIdx.x; b < a.nblk; b += JOB_THREADS) ipv += a.ip_part[b];
  // one pass over the Grams: cross = hadamard of all (fit), gam = hadamard of
  // k != skip (Gamma, ascending k from ones)
  double cross[4][4], gam[4][4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      cross[t][e] = 1.0;
      gam[t][e] = 1.0;
    }
  // Gram blocks of up to four modes loaded together: one round trip for a
  // tensor of order <= 4 (the MTTKRP beside the job keeps the memory system
  // busy, so each round trip is long)
  auto load_block = [&](int k, double (&g)[4][4]) {
    const double* Gk = a.G + (i64)k * R * R;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = job_row(e), j = job_col(t);
        g[t][e] = (k < a.nd && i < R && j < R) ? Gk[i * R + j] : 0.0;
      }
  };
  auto fold = [&](int k, const double (&g)[4][4]) {
    if (k >= a.nd) return;
    const bool in_gam = k != a.skip;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        cross[t][e] *= g[t][e];
        gam[t][e] = in_gam ? gam[t][e] * g[t][e] : gam[t][e];
      }
  };
  for (int k = 0; k < a.nd; k += 4) {
    double g0[4][4], g1[4][4], g2[4][4], g3[4][4];
    load_block(k, g0);
    load_block(k + 1, g1);
    load_block(k + 2, g2);
    load_block(k + 3, g3);
    fold(k, g0);
    fold(k + 1, g1);
    fold(k + 2, g2);
    fold(k + 3, g3);
  }
  if (a.gate) {
    // the Grams (and <X, Xhat> partials) of this wave are in registers: open
    // the gate of the MTTKRP workgroups, whose stream of X would otherwise
    // queue ahead of these round trips
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    if (threadIdx.x == 0) __hip_atomic_store(a.gate, a.gate_stamp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (want_fit) {
    // ||Xhat||^2 = sum_{r,s} prod_k G_k[r,s]; <X, Xhat> = sum of the partials
    double rec = 0.0;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) rec += cross[t][e];
    const double recon = job_block_sum(rec, lds);
    const double ip = job_block_sum(ipv, lds + 8);
    if (threadIdx.x == 0) {
      // the scalar block as it stands (the solve status of the sweep just
      // evaluated included), read in one round trip before any store to it
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
        // straight into host memory, then the stamp the host polls for (no copy
        // launch, no event); the fence orders the block before the stamp, and
        // nothing here waits for the stamp itself to land
#pragma unroll
        for (int i = 0; i < JOB_SC_STAMP; ++i) a.host[i] = blk[i];
        __threadfence_system();
        a.host[JOB_SC_STAMP] = a.stamp;
      }
    }
    __syncthreads();
  }
  TENRSO_JOB_MARK(2);
  if (a.skip < 0) return;
  // Gamma + regularisation, identity outside R x R, in the 4 x 4 block layout
  int path = 0;
  bool ok = false;
  // Gamma waits in Winv (global) for the ridge retry instead of in registers
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int e = 0; e < 4; ++e) a.Winv[job_row(e) * JOB_NE + job_col(t)] = gam[t][e];
  doublex4 acc[4];
  // acc = Gamma (first attempt) or Gamma from Winv (retry), then the identity
  // outside R x R and the regularisation band, each in a uniform branch taken
  // only when it changes something.  The retry is written out as a second
  // attempt rather than a loop: in a loop the compiler hoists the band
  // comparisons of the retry into the loop header, where the first attempt
  // pays for them.
  auto pad_and_band = [&](double ridge) {
    if (R < JOB_NE) {
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int i = job_row(e), col = job_col(t);
          acc[t][e] = (i < R && col < R) ? acc[t][e] : (i == col ? 1.0 : 0.0);
        }
    }
    if (ridge > 0.0 || a.tik_order > 0 || a.l2 > 0.0) {
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int i = job_row(e), col = job_col(t);
          if (i < R && col < R) acc[t][e] += job_reg(a, i, col, ridge);
        }
    }
  };
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int e = 0; e < 4; ++e) acc[t][e] = gam[t][e];
  pad_and_band(0.0);
  TENRSO_JOB_MARK(3);
  ok = job_block_gj(acc, lds);
  TENRSO_JOB_MARK(4);
  if (!ok) {
    // ridge retry (cp/helpers.rs:477-490): Gamma + eps 10R I
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[t][e] = a.Winv[job_row(e) * JOB_NE + job_col(t)];
    pad_and_band((double)FLT_EPSILON * (double)(10 * R));
    ok = job_block_gj(acc, lds);
    path = 1;
  }
  if (!ok) path = 2;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int e = 0; e < 4; ++e) a.Winv[job_row(e) * JOB_NE + job_col(t)] = ok ? acc[t][e] : 0.0;
  TENRSO_JOB_MARK(5);
  if (threadIdx.x == 0 && a.scal) {
    const double prev = a.scal[a.status_slot];
    a.scal[a.status_slot] = a.status_reset ? (double)path : fmax(prev, (double)path);
  }
}

}  // namespace tenrso
