// cp_init.hip -- the SVD-based CP initialisations (InitStrategy::Svd, Nnsvd,
// LeverageScore of tenrso-decomp cp/helpers.rs:232-373) on the device.
//
// Per mode k (X viewed as [P, I, Q], I = I_k, J = P*Q columns of the unfolding):
//   1. K = X_(k) X_(k)^T (I x I, FP64): the left singular vectors of X_(k) are the
//      eigenvectors of K and its singular values the square roots of K's
//      eigenvalues.  Split-K tiled kernel over the column range, FP64 sums, then a
//      reduction that also mirrors the upper triangle.
//   2. One workgroup: cyclic parallel Jacobi (round-robin pairing, all n/2
//      rotations of a round applied at once: rows, barrier, columns and V) until
//      the off-diagonal mass is below 1e-14 of the diagonal, then the top R
//      eigenpairs in descending order.  The sign of each eigenvector is fixed so
//      that its entry of largest magnitude is positive (an SVD's signs are
//      arbitrary; the third-party SVD behind the reference does not document one).
//   3. Nnsvd only: ||max(v_r, 0)|| and ||max(-v_r, 0)|| of the right singular
//      vectors v_r = X_(k)^T u_r / s_r, streamed over the columns of X_(k).
//   4. The factor: Svd u_r; LeverageScore u_r sqrt(s_r) sqrt(l_i I) with the
//      normalised leverage scores l_i; Nnsvd the larger-product sign part of u_r
//      scaled by sqrt(s_r * prod) (helpers.rs:265-310).  Columns past
//      min(R, I, J) come from the caller's N(0, 0.01) draws (|.| for Nnsvd,
//      times l_i for LeverageScore), as in the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <random>
#include <vector>

#include "common.hpp"
#include "cp_init.hpp"
#include "mttkrp.hpp"

namespace tenrso {

constexpr int INIT_RMAX = 128;  // rank bound of the SVD-based initialisations (that of the ALS)

namespace {

constexpr int GT = 32;          // Gram tile
constexpr int EIG_THREADS = 1024;

// Gram partials: workgroup (tile pair, split) sums the columns [j0, j1) of its
// split for a 32 x 32 tile (ti <= tj) of K.  Thread t computes row t / 8 and
// columns 4 (t % 8) .. + 3 of the tile in FP64.
__global__ __launch_bounds__(256) void unfold_gram_partial_kernel(const float* __restrict__ X, i64 P, i64 I, i64 Q,
                                                                  int nt, int splits, double* __restrict__ part) {
  __shared__ float sA[GT][GT + 1], sB[GT][GT + 1];
  int t = blockIdx.x, ti = 0;
  while (t >= nt - ti) {
    t -= nt - ti;
    ++ti;
  }
  const int tj = ti + t;
  const int split = blockIdx.y;
  const i64 J = P * Q;
  const i64 j0 = J * split / splits, j1 = J * (split + 1) / splits;
  const int tid = threadIdx.x, r = tid >> 3, c4 = (tid & 7) * 4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (i64 jb = j0; jb < j1; jb += GT) {
    // stage 32 rows x 32 columns of both row tiles: element (row, col) of the
    // unfolding is X[p, i, q] with (p, q) = (j / Q, j % Q)
    for (int e = tid; e < GT * GT; e += 256) {
      const int rr = e / GT, cc = e % GT;
      const i64 j = jb + cc;
      float a = 0.f, b = 0.f;
      if (j < j1) {
        const i64 p = j / Q, q = j - (j / Q) * Q;
        const i64 ia = (i64)ti * GT + rr, ib = (i64)tj * GT + rr;
        if (ia < I) a = X[(p * I + ia) * Q + q];
        if (ib < I) b = X[(p * I + ib) * Q + q];
      }
      sA[rr][cc] = a;
      sB[rr][cc] = b;
    }
    __syncthreads();
#pragma unroll 4
    for (int cc = 0; cc < GT; ++cc) {
      const double a = (double)sA[r][cc];
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u] = fma(a, (double)sB[c4 + u][cc], acc[u]);
    }
    __syncthreads();
  }
  double* out = part + (i64)split * I * I;
  const i64 i = (i64)ti * GT + r;
  if (i < I)
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const i64 i2 = (i64)tj * GT + c4 + u;
      if (i2 < I) out[i * I + i2] = acc[u];
    }
}

// K[i][i2] = sum of the split partials, for tile(i) <= tile(i2), mirrored below.
__global__ __launch_bounds__(256) void unfold_gram_reduce_kernel(const double* __restrict__ part, int splits, i64 I,
                                                                 double* __restrict__ K) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e >= I * I) return;
  const i64 i = e / I, i2 = e - (e / I) * I;
  if (i / GT > i2 / GT) return;
  double s = 0.0;
  for (int k = 0; k < splits; ++k) s += part[(i64)k * I * I + e];
  K[i * I + i2] = s;
  K[i2 * I + i] = s;
}

// One workgroup: eigen-decomposition of the symmetric n x n matrix A (in place,
// global memory), eigenvectors in V; then the top `R` eigenpairs: U (n x R,
// column r = eigenvector of the r-th largest eigenvalue, largest-magnitude entry
// positive) and lam (R).  red: EIG_THREADS doubles of scratch is in LDS.
__global__ __launch_bounds__(EIG_THREADS) void jacobi_eig_kernel(double* __restrict__ A, double* __restrict__ V,
                                                                 int n, int R, int max_sweeps,
                                                                 double* __restrict__ U, double* __restrict__ lam,
                                                                 int* __restrict__ sweeps_out) {
  __shared__ double cs[2 * 512];  // (c, s) of each pair of a round (n <= 1024)
  __shared__ double red[EIG_THREADS / 64];
  __shared__ int done;
  const int tid = threadIdx.x;
  const int m = n + (n & 1);  // even count; index n is a dummy when n is odd
  const int half = m / 2;
  for (i64 e = tid; e < (i64)n * n; e += EIG_THREADS) V[e] = (e / n == e % n) ? 1.0 : 0.0;
  auto block_sum = [&](double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < EIG_THREADS / 64; ++w) s += red[w];
    __syncthreads();
    return s;
  };
  // pair k of round r (circle method): (r, m - 1) for k = 0, else (r + k, r - k) mod m - 1
  auto pair_of = [&](int r, int k, int& a, int& b) {
    if (k == 0) {
      a = r;
      b = m - 1;
    } else {
      a = (r + k) % (m - 1);
      b = (r - k + (m - 1)) % (m - 1);
    }
  };
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (i64 e = tid; e < (i64)n * n; e += EIG_THREADS) {
      const double v = A[e];
      if (e / n == e % n) dia += v * v;
      else off += v * v;
    }
    off = block_sum(off);
    dia = block_sum(dia);
    if (tid == 0) done = off <= 1e-28 * dia || off == 0.0;
    __syncthreads();
    if (done) break;
    for (int r = 0; r < m - 1; ++r) {
      // rotation of every pair from the matrix before the round (NR jacobi form)
      for (int k = tid; k < half; k += EIG_THREADS) {
        int a, b;
        pair_of(r, k, a, b);
        double c = 1.0, s = 0.0;
        if (a < n && b < n) {
          const double apq = A[(i64)a * n + b];
          if (apq != 0.0) {
            const double th = (A[(i64)b * n + b] - A[(i64)a * n + a]) / (2.0 * apq);
            const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(fma(th, th, 1.0)));
            c = 1.0 / sqrt(fma(t, t, 1.0));
            s = t * c;
          }
        }
        cs[2 * k] = c;
        cs[2 * k + 1] = s;
      }
      __syncthreads();
      // rows: A[a][:], A[b][:] <- c A[a] - s A[b], s A[a] + c A[b]
      for (i64 it = tid; it < (i64)half * n; it += EIG_THREADS) {
        const int k = (int)(it / n), j = (int)(it - (it / n) * n);
        int a, b;
        pair_of(r, k, a, b);
        if (a >= n || b >= n) continue;
        const double c = cs[2 * k], s = cs[2 * k + 1];
        const double xa = A[(i64)a * n + j], xb = A[(i64)b * n + j];
        A[(i64)a * n + j] = c * xa - s * xb;
        A[(i64)b * n + j] = s * xa + c * xb;
      }
      __syncthreads();
      // columns of A and V
      for (i64 it = tid; it < (i64)half * n; it += EIG_THREADS) {
        const int k = (int)(it / n), i = (int)(it - (it / n) * n);
        int a, b;
        pair_of(r, k, a, b);
        if (a >= n || b >= n) continue;
        const double c = cs[2 * k], s = cs[2 * k + 1];
        const double xa = A[(i64)i * n + a], xb = A[(i64)i * n + b];
        A[(i64)i * n + a] = c * xa - s * xb;
        A[(i64)i * n + b] = s * xa + c * xb;
        const double va = V[(i64)i * n + a], vb = V[(i64)i * n + b];
        V[(i64)i * n + a] = c * va - s * vb;
        V[(i64)i * n + b] = s * va + c * vb;
      }
      __syncthreads();
    }
  }
  // descending order: rank of i = #{j : lam_j > lam_i or (lam_j == lam_i and j < i)}
  for (int i = tid; i < n; i += EIG_THREADS) {
    const double li = A[(i64)i * n + i];
    int rk = 0;
    for (int j = 0; j < n; ++j) {
      const double lj = A[(i64)j * n + j];
      rk += (lj > li || (lj == li && j < i)) ? 1 : 0;
    }
    if (rk < R) {
      lam[rk] = li;
      // sign: largest-magnitude entry positive (first such entry)
      double best = 0.0;
      for (int row = 0; row < n; ++row) {
        const double v = V[(i64)row * n + i];
        if (fabs(v) > fabs(best)) best = v;
      }
      const double sg = best < 0.0 ? -1.0 : 1.0;
      for (int row = 0; row < n; ++row) U[(i64)row * R + rk] = sg * V[(i64)row * n + i];
    }
  }
  if (tid == 0 && sweeps_out) *sweeps_out = sweep;
}

// Nnsvd: for columns j of X_(k) and rank columns r0 .. r0 + 16 of U (n x R):
// z = X_(k)^T u_r; partial[block][r] = (sum of z^2 over z > 0, over z < 0).
constexpr int NN_RC = 16;
__global__ __launch_bounds__(256) void nnsvd_norms_kernel(const float* __restrict__ X, i64 P, i64 I, i64 Q,
                                                          const double* __restrict__ U, int R, int ar,
                                                          double* __restrict__ part) {
  __shared__ double su[256][NN_RC];
  __shared__ double red[2 * NN_RC][4];
  const int tid = threadIdx.x;
  const int r0 = blockIdx.y * NN_RC;
  const i64 J = P * Q;
  const i64 j = (i64)blockIdx.x * 256 + tid;
  const bool live = j < J;
  const i64 p = live ? j / Q : 0, q = live ? j - (j / Q) * Q : 0;
  double z[NN_RC];
#pragma unroll
  for (int u = 0; u < NN_RC; ++u) z[u] = 0.0;
  for (i64 i0 = 0; i0 < I; i0 += 256) {
    for (int e = tid; e < 256 * NN_RC; e += 256) {
      const int ii = e / NN_RC, rr = e % NN_RC;
      su[ii][rr] = (i0 + ii < I && r0 + rr < ar) ? U[(i0 + ii) * R + r0 + rr] : 0.0;
    }
    __syncthreads();
    const i64 iend = std::min<i64>(256, I - i0);
    if (live)
      for (i64 ii = 0; ii < iend; ++ii) {
        const double x = (double)X[(p * I + i0 + ii) * Q + q];
#pragma unroll
        for (int u = 0; u < NN_RC; ++u) z[u] = fma(x, su[ii][u], z[u]);
      }
    __syncthreads();
  }
  // per rank column: sums of positive and negative squares over the block
#pragma unroll
  for (int u = 0; u < NN_RC; ++u) {
    double ps = z[u] > 0.0 ? z[u] * z[u] : 0.0, ns = z[u] < 0.0 ? z[u] * z[u] : 0.0;
    for (int off = 32; off > 0; off >>= 1) {
      ps += __shfl_down(ps, off, 64);
      ns += __shfl_down(ns, off, 64);
    }
    if ((tid & 63) == 0) {
      red[2 * u][tid >> 6] = ps;
      red[2 * u + 1][tid >> 6] = ns;
    }
  }
  __syncthreads();
  if (tid < 2 * NN_RC) {
    const int u = tid >> 1;
    const double s = (red[tid][0] + red[tid][1]) + (red[tid][2] + red[tid][3]);
    if (r0 + u < ar) part[((i64)blockIdx.x * ar + r0 + u) * 2 + (tid & 1)] = s;
  }
}

// One workgroup: the factor (n x R, fp32) of mode k from U (n x R, FP64; columns
// < ar valid), lam, the Nnsvd partials (nblk x ar x 2) and the N(0, 0.01) draws
// rnd (n x (R - ar)) for the columns past ar.
__global__ __launch_bounds__(1024) void svd_init_finish_kernel(const double* __restrict__ U,
                                                               const double* __restrict__ lam, int n, int R, int ar,
                                                               int kind, const double* __restrict__ nn_part, int nblk,
                                                               const float* __restrict__ rnd, float* __restrict__ F) {
  __shared__ double red[16];
  __shared__ double lev_total;
  __shared__ double colv[INIT_RMAX][4];  // Nnsvd: ||u+||^2, ||u-||^2, ||z+||^2, ||z-||^2 per column
  const int tid = threadIdx.x;
  auto block_sum = [&](double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < 16; ++w) s += red[w];
    __syncthreads();
    return s;
  };
  if (kind == TENRSO_INIT_LEVERAGE) {
    // l_i = sum_{r < ar} u_ir^2 / ar, normalised by their total when it exceeds
    // f32 epsilon (helpers.rs:329-343)
    double part = 0.0;
    for (int i = tid; i < n; i += 1024) {
      double sc = 0.0;
      for (int r = 0; r < ar; ++r) sc += U[(i64)i * R + r] * U[(i64)i * R + r];
      part += sc / ar;
    }
    const double tot = block_sum(part);
    if (tid == 0) lev_total = tot;
    __syncthreads();
  }
  if (kind == TENRSO_INIT_NNSVD) {
    for (int r = 0; r < ar; ++r) {
      double up = 0.0, un = 0.0;
      for (int i = tid; i < n; i += 1024) {
        const double u = U[(i64)i * R + r];
        if (u > 0.0) up += u * u;
        else un += u * u;
      }
      up = block_sum(up);
      un = block_sum(un);
      double zp = 0.0, zn = 0.0;
      for (int b = tid; b < nblk; b += 1024) {
        zp += nn_part[((i64)b * ar + r) * 2];
        zn += nn_part[((i64)b * ar + r) * 2 + 1];
      }
      zp = block_sum(zp);
      zn = block_sum(zn);
      if (tid == 0) {
        colv[r][0] = up;
        colv[r][1] = un;
        colv[r][2] = zp;
        colv[r][3] = zn;
      }
    }
    __syncthreads();
  }
  for (i64 e = tid; e < (i64)n * R; e += 1024) {
    const int i = (int)(e / R), r = (int)(e - (e / R) * R);
    float out;
    if (r < ar) {
      const double u = U[(i64)i * R + r];
      const double s = sqrt(lam[r] > 0.0 ? lam[r] : 0.0);
      if (kind == TENRSO_INIT_SVD) {
        out = (float)u;
      } else if (kind == TENRSO_INIT_LEVERAGE) {
        double sc = 0.0;
        for (int k = 0; k < ar; ++k) sc += U[(i64)i * R + k] * U[(i64)i * R + k];
        double l = sc / ar;
        if (lev_total > (double)FLT_EPSILON) l /= lev_total;
        out = (float)(u * sqrt(s) * sqrt(l * n));
      } else {
        // ||v+|| = ||z+|| / s (v = z / s); with s == 0 both products are zero
        const double vp = s > 0.0 ? sqrt(colv[r][2]) / s : 0.0, vn = s > 0.0 ? sqrt(colv[r][3]) / s : 0.0;
        const double pp = sqrt(colv[r][0]) * vp, np = sqrt(colv[r][1]) * vn;
        if (pp >= np) out = (float)((u > 0.0 ? u : 0.0) * sqrt(s * pp));
        else out = (float)((u > 0.0 ? 0.0 : -u) * sqrt(s * np));
      }
    } else {
      const float d = rnd[(i64)i * (R - ar) + (r - ar)];
      if (kind == TENRSO_INIT_SVD) {
        out = d;
      } else if (kind == TENRSO_INIT_NNSVD) {
        out = fabsf(d);
      } else {
        double sc = 0.0;
        for (int k = 0; k < ar; ++k) sc += U[(i64)i * R + k] * U[(i64)i * R + k];
        double l = ar > 0 ? sc / ar : 0.0;
        if (lev_total > (double)FLT_EPSILON) l /= lev_total;
        out = (float)(d * l);
      }
    }
    F[e] = out;
  }
}

}  // namespace

int cp_init_svd(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, i64 R, int kind, uint64_t seed,
                float* const* F, int* sweeps) {
  if (kind != TENRSO_INIT_SVD && kind != TENRSO_INIT_NNSVD && kind != TENRSO_INIT_LEVERAGE) {
    set_error("unknown initialisation kind %d", kind);
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  if (nd < 1 || nd > MAXD) {
    set_error("tensors with %d modes are not supported", nd);
    return TENRSO_ERR_UNSUPPORTED;
  }
  if (R <= 0 || R > INIT_RMAX) {
    set_error("rank %lld outside the on-device SVD initialisation bound 1..%d", (long long)R, INIT_RMAX);
    return TENRSO_ERR_UNSUPPORTED;
  }
  i64 total = 1;
  for (int k = 0; k < nd; ++k) total *= shape[k];
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> noise(0.0, 0.01);  // helpers.rs: Normal::new(0.0, 0.01)
  for (int k = 0; k < nd; ++k) {
    const i64 I = shape[k];
    if (I > 1024) {
      set_error("mode %d of size %lld exceeds the on-device SVD initialisation bound 1024", k, (long long)I);
      return TENRSO_ERR_UNSUPPORTED;
    }
    i64 P = 1, Q = 1;
    for (int m = 0; m < k; ++m) P *= shape[m];
    for (int m = k + 1; m < nd; ++m) Q *= shape[m];
    const i64 J = P * Q;
    const int ar = (int)std::min<i64>(R, std::min<i64>(I, J));
    const int nt = (int)((I + GT - 1) / GT);
    const int ntp = nt * (nt + 1) / 2;
    int splits = (int)std::max<i64>(1, std::min<i64>(64, 512 / ntp));
    splits = (int)std::max<i64>(1, std::min<i64>(splits, (64ll << 20) / (I * I * 8)));
    void *pK, *pV, *pPart, *pW;
    TENRSO_TRY(ctx_buffer(ctx, SLOT_EIG_A, (size_t)I * I * 8, &pK));
    TENRSO_TRY(ctx_buffer(ctx, SLOT_EIG_V, (size_t)I * I * 8, &pV));
    TENRSO_TRY(ctx_buffer(ctx, SLOT_EIG_PART, (size_t)splits * I * I * 8, &pPart));
    // U (I x R doubles), lam (R), the Nnsvd partials, the draws (I x (R - ar) floats), sweeps
    const i64 nn_blk = (J + 255) / 256;
    const size_t u_bytes = (size_t)I * R * 8, lam_bytes = (size_t)R * 8;
    const size_t nn_bytes = kind == TENRSO_INIT_NNSVD ? (size_t)nn_blk * ar * 2 * 8 : 0;
    const size_t rnd_bytes = (size_t)I * (R - ar) * 4 + 16;
    TENRSO_TRY(ctx_buffer(ctx, SLOT_EIG_W, u_bytes + lam_bytes + nn_bytes + rnd_bytes + 16, &pW));
    double* K = (double*)pK;
    double* U = (double*)pW;
    double* lam = U + (size_t)I * R;
    double* nn = lam + R;
    float* rnd = (float*)((char*)nn + nn_bytes);
    hipLaunchKernelGGL(unfold_gram_partial_kernel, dim3((unsigned)ntp, (unsigned)splits), dim3(256), 0, ctx->stream, X,
                       P, I, Q, nt, splits, (double*)pPart);
    TENRSO_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(unfold_gram_reduce_kernel, dim3((unsigned)((I * I + 255) / 256)), dim3(256), 0, ctx->stream,
                       (const double*)pPart, splits, I, K);
    TENRSO_HIP_TRY(hipGetLastError());
    int* dsweeps = (int*)ctx->pinned_dev + 64;  // scratch int in mapped host memory
    hipLaunchKernelGGL(jacobi_eig_kernel, dim3(1), dim3(EIG_THREADS), 0, ctx->stream, K, (double*)pV, (int)I,
                       (int)R, 60, U, lam, dsweeps);
    TENRSO_HIP_TRY(hipGetLastError());
    if (kind == TENRSO_INIT_NNSVD && ar > 0) {
      hipLaunchKernelGGL(nnsvd_norms_kernel, dim3((unsigned)nn_blk, (unsigned)((ar + NN_RC - 1) / NN_RC)), dim3(256),
                         0, ctx->stream, X, P, I, Q, (const double*)U, (int)R, ar, nn);
      TENRSO_HIP_TRY(hipGetLastError());
    }
    if (ar < R) {
      std::vector<float> h((size_t)I * (R - ar));
      for (auto& v : h) v = (float)noise(rng);
      TENRSO_HIP_TRY(hipMemcpyAsync(rnd, h.data(), h.size() * 4, hipMemcpyHostToDevice, ctx->stream));
      TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));  // h leaves scope
    }
    hipLaunchKernelGGL(svd_init_finish_kernel, dim3(1), dim3(1024), 0, ctx->stream, (const double*)U,
                       (const double*)lam, (int)I, (int)R, ar, kind, (const double*)nn, (int)nn_blk,
                       (const float*)rnd, F[k]);
    TENRSO_HIP_TRY(hipGetLastError());
    if (sweeps) {
      TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
      sweeps[k] = ((volatile int*)ctx->pinned)[64];
    }
  }
  (void)total;
  return TENRSO_OK;
}

}  // namespace tenrso
