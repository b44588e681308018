// ad.hip -- reverse-mode rules of the CP path on gfx950.
//
//   * MTTKRP VJP.  M = mttkrp(X, A, n) is linear in X and in each A_m (m != n).
//     For a cotangent Mbar (I_n x R):
//       dA_m = mttkrp(X, A with A_n := Mbar, m)          (m != n)
//       dX   = cp_reconstruct(A with A_n := Mbar)
//     Both reuse the forward kernels; Khatri-Rao rows are recomputed, never stored.
//   * CP reconstruction VJP (tenrso-ad/src/grad.rs:81-120, CpReconstructionGrad):
//       dA_n = mttkrp(dL/dXhat, A, n) * diag(weights)
//   * CP loss L = 0.5 ||X - [[A]]||^2 and its gradient
//       dL/dA_n = A_n Gamma_n - mttkrp(X, A, n),  Gamma_n = hadamard_{k != n} A_k^T A_k
//       L = 0.5 (||X||^2 - 2 <X, Xhat> + ||Xhat||^2)
//     with the Grams, Gamma_n, the product A_n Gamma_n and the loss in FP64.
//   * Gradient-descent CP fit: simultaneous steps A_n -= lr dL/dA_n.
#include <cmath>
#include <vector>

#include "ad.hpp"
#include "cp_linalg.hpp"
#include "kr_reconstruct.hpp"
#include "mttkrp.hpp"

namespace tenrso {

static inline unsigned cdiv(i64 a, i64 b) { return (unsigned)((a + b - 1) / b); }

// Gamma = hadamard_{k != skip} G_k (FP64, R x R)
__global__ __launch_bounds__(256) void hadamard_gram_kernel(const double* __restrict__ G, int nd, int skip, int R,
                                                            double* __restrict__ Gamma) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= R * R) return;
  double v = 1.0;
  for (int k = 0; k < nd; ++k)
    if (k != skip) v = v * G[(i64)k * R * R + e];
  Gamma[e] = v;
}

// grad[i, c] = sum_r A[i, r] Gamma[r, c] - M[i, c]   (one workgroup per row, FP64)
__global__ __launch_bounds__(64) void grad_rows_kernel(const float* __restrict__ A, const double* __restrict__ Gamma,
                                                       const double* __restrict__ M, i64 rows, int R,
                                                       float* __restrict__ grad) {
  __shared__ double sa[64];
  const i64 i = blockIdx.x;
  const int c = threadIdx.x;
  for (int r = c; r < R; r += 64) sa[r] = (double)A[i * R + r];
  __syncthreads();
  for (int cc = c; cc < R; cc += 64) {
    double s0 = 0.0, s1 = 0.0;
    int r = 0;
    for (; r + 2 <= R; r += 2) {
      s0 += sa[r] * Gamma[r * R + cc];
      s1 += sa[r + 1] * Gamma[(r + 1) * R + cc];
    }
    if (r < R) s0 += sa[r] * Gamma[r * R + cc];
    grad[i * R + cc] = (float)((s0 + s1) - M[i * R + cc]);
  }
}

// partial <M, A> per row block (64 rows), FP64
__global__ __launch_bounds__(256) void dot_rows_kernel(const double* __restrict__ M, const float* __restrict__ A,
                                                       i64 n, double* __restrict__ part) {
  __shared__ double red[256];
  const i64 b0 = (i64)blockIdx.x * 4096;
  double s = 0.0;
  for (i64 e = b0 + threadIdx.x; e < n && e < b0 + 4096; e += 256) s += M[e] * (double)A[e];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// loss = 0.5 (||X||^2 - 2 sum(part) + sum_rs prod_k G_k[r,s]); also ||grad||_F scale for lr
__global__ __launch_bounds__(256) void loss_kernel(const double* __restrict__ G, int nd, int R,
                                                   const double* __restrict__ part, int npart,
                                                   const double* __restrict__ scal, double* __restrict__ out) {
  __shared__ double red[256];
  double rec = 0.0;
  for (int e = threadIdx.x; e < R * R; e += 256) {
    double c = 1.0;
    for (int k = 0; k < nd; ++k) c = c * G[(i64)k * R * R + e];
    rec += c;
  }
  red[threadIdx.x] = rec;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  const double recon = red[0];
  __syncthreads();
  double ip = 0.0;
  for (int b = threadIdx.x; b < npart; b += 256) ip += part[b];
  red[threadIdx.x] = ip;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = 0.5 * (scal[SC_NORM_X] - 2.0 * red[0] + recon);
    out[1] = red[0];
    out[2] = recon;
  }
}

// ||Gamma_n||_F for every mode (used by the automatic step size)
__global__ __launch_bounds__(256) void gamma_norms_kernel(const double* __restrict__ G, int nd, int R,
                                                          double* __restrict__ out) {
  __shared__ double red[256];
  for (int n = 0; n < nd; ++n) {
    double s = 0.0;
    for (int e = threadIdx.x; e < R * R; e += 256) {
      double v = 1.0;
      for (int k = 0; k < nd; ++k)
        if (k != n) v = v * G[(i64)k * R * R + e];
      s += v * v;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
      if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = sqrt(red[0]);
    __syncthreads();
  }
}

// A -= step * grad, step = lr (lr > 0) or 1 / max_n ||Gamma_n||_F read from gnorm
__global__ __launch_bounds__(256) void axpy_step_kernel(float* __restrict__ A, const float* __restrict__ g, i64 n,
                                                        double lr, const double* __restrict__ gnorm, int nd) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double step = lr;
  if (lr <= 0.0) {
    double mx = 0.0;
    for (int k = 0; k < nd; ++k) mx = fmax(mx, gnorm[k]);
    step = mx > 0.0 ? 1.0 / mx : 0.0;
  }
  A[i] = (float)((double)A[i] - step * (double)g[i]);
}

__global__ __launch_bounds__(256) void scale_cols_kernel(float* __restrict__ A, i64 rows, i64 R,
                                                         const float* __restrict__ w) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e >= rows * R) return;
  A[e] = A[e] * w[e % R];
}

// ---------------------------------------------------------------- host side
int mttkrp_vjp(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, const float* const* F, i64 R, int mode,
               const float* mbar, float* const* dF, float* dX) {
  std::vector<const float*> Fp(F, F + nd);
  Fp[mode] = mbar;
  for (int m = 0; m < nd; ++m) {
    if (m == mode || !dF || !dF[m]) continue;
    TENRSO_TRY(mttkrp_dense(ctx, TENRSO_PREC_FP32, X, nullptr, nd, shape, Fp.data(), R, m, MttkrpOut::F32, dF[m],
                            nullptr));
  }
  if (dX) TENRSO_TRY(cp_reconstruct_f32(ctx, nd, shape, Fp.data(), R, nullptr, dX));
  return TENRSO_OK;
}

int cp_factor_grads(tenrso_ctx* ctx, const float* Gt, int nd, const i64* shape, const float* const* F, i64 R,
                    const float* w, float* const* grads) {
  for (int n = 0; n < nd; ++n) {
    TENRSO_TRY(mttkrp_dense(ctx, TENRSO_PREC_FP32, Gt, nullptr, nd, shape, F, R, n, MttkrpOut::F32, grads[n],
                            nullptr));
    if (w) {
      hipLaunchKernelGGL(scale_cols_kernel, dim3(cdiv(shape[n] * R, 256)), dim3(256), 0, ctx->stream, grads[n],
                         shape[n], R, w);
      TENRSO_HIP_TRY(hipGetLastError());
    }
  }
  return TENRSO_OK;
}

namespace {
struct LossWork {
  double* G;      // [nd][R][R]
  double* Gamma;  // [R][R]
  double* M;      // [maxI][R]
  double* part;
  double* scal;
  double* out;    // loss, ip, recon, gamma norms [3 .. 3+nd)
};

int loss_work(tenrso_ctx* ctx, int nd, const i64* shape, i64 R, LossWork* w) {
  i64 maxI = 0;
  for (int k = 0; k < nd; ++k) maxI = std::max(maxI, shape[k]);
  void *pG, *pGa, *pM, *pp, *ps;
  TENRSO_TRY(ctx_buffer(ctx, SLOT_GRAMS, (size_t)(nd + 1) * R * R * sizeof(double), &pG));
  TENRSO_TRY(ctx_buffer(ctx, SLOT_SOLVE, (size_t)std::max<i64>(R * R, 4096) * sizeof(double), &pGa));
  TENRSO_TRY(ctx_buffer(ctx, SLOT_M64, (size_t)maxI * R * sizeof(double), &pM));
  TENRSO_TRY(ctx_buffer(ctx, SLOT_IP, (size_t)std::max<i64>(4096, maxI * R / 4096 + 16) * sizeof(double), &pp));
  TENRSO_TRY(ctx_buffer(ctx, SLOT_SCALARS, 64 * sizeof(double), &ps));
  w->G = (double*)pG;
  w->Gamma = (double*)pGa;
  w->M = (double*)pM;
  w->part = (double*)pp;
  w->scal = (double*)ps;
  w->out = w->scal + SC_COUNT;
  return TENRSO_OK;
}
}  // namespace

// grads[n] = A_n Gamma_n - mttkrp(X, A, n); loss into w.out[0] (device), with ||X||^2
// already in w.scal[SC_NORM_X].
static int loss_grad_device(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, const float* const* F, i64 R,
                            float* const* grads, LossWork& w) {
  for (int k = 0; k < nd; ++k) TENRSO_TRY(gram_f64(ctx, F[k], shape[k], R, w.G + (i64)k * R * R));
  int npart = 1;
  for (int n = 0; n < nd; ++n) {
    TENRSO_TRY(mttkrp_dense(ctx, TENRSO_PREC_FP32, X, nullptr, nd, shape, F, R, n, MttkrpOut::F64, w.M, nullptr));
    if (n == nd - 1) {
      npart = (int)cdiv(shape[n] * R, 4096);
      hipLaunchKernelGGL(dot_rows_kernel, dim3(npart), dim3(256), 0, ctx->stream, w.M, F[n], shape[n] * R, w.part);
    }
    if (grads && grads[n]) {
      hipLaunchKernelGGL(hadamard_gram_kernel, dim3(cdiv(R * R, 256)), dim3(256), 0, ctx->stream, w.G, nd, n,
                         (int)R, w.Gamma);
      hipLaunchKernelGGL(grad_rows_kernel, dim3((unsigned)shape[n]), dim3(64), 0, ctx->stream, F[n], w.Gamma, w.M,
                         shape[n], (int)R, grads[n]);
    }
    TENRSO_HIP_TRY(hipGetLastError());
  }
  hipLaunchKernelGGL(loss_kernel, dim3(1), dim3(256), 0, ctx->stream, w.G, nd, (int)R, w.part, npart, w.scal, w.out);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

int cp_loss_grad(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, const float* const* F, i64 R,
                 float* const* grads, double* loss_host) {
  if (R > SOLVE_RMAX) {
    set_error("rank %lld exceeds the on-device Gram bound %d", (long long)R, SOLVE_RMAX);
    return TENRSO_ERR_UNSUPPORTED;
  }
  LossWork w;
  TENRSO_TRY(loss_work(ctx, nd, shape, R, &w));
  i64 total = 1;
  for (int k = 0; k < nd; ++k) total *= shape[k];
  TENRSO_TRY(fill_f64(ctx, w.scal, 64, 0.0));
  TENRSO_TRY(tensor_norm_min(ctx, X, total, w.part, w.scal));
  TENRSO_TRY(loss_grad_device(ctx, X, nd, shape, F, R, grads, w));
  double* hs = (double*)ctx->pinned;
  TENRSO_HIP_TRY(hipMemcpyAsync(hs, w.out, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (loss_host) *loss_host = hs[0];
  return TENRSO_OK;
}

int cp_gd(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, float* const* F, i64 R, int steps, double lr,
          double* loss_hist) {
  if (R > SOLVE_RMAX) {
    set_error("rank %lld exceeds the on-device Gram bound %d", (long long)R, SOLVE_RMAX);
    return TENRSO_ERR_UNSUPPORTED;
  }
  LossWork w;
  TENRSO_TRY(loss_work(ctx, nd, shape, R, &w));
  i64 total = 1, ftot = 0;
  for (int k = 0; k < nd; ++k) {
    total *= shape[k];
    ftot += shape[k] * R;
  }
  void* pg = nullptr;
  TENRSO_TRY(ctx_buffer(ctx, SLOT_STAGE_OUT, (size_t)ftot * sizeof(float), &pg));
  std::vector<float*> grads(nd);
  i64 off = 0;
  for (int k = 0; k < nd; ++k) {
    grads[k] = (float*)pg + off;
    off += shape[k] * R;
  }
  TENRSO_TRY(fill_f64(ctx, w.scal, 64, 0.0));
  TENRSO_TRY(tensor_norm_min(ctx, X, total, w.part, w.scal));
  std::vector<double> hist_dev;
  double* hs = (double*)ctx->pinned;
  for (int s = 0; s < steps; ++s) {
    TENRSO_TRY(loss_grad_device(ctx, X, nd, shape, (const float* const*)F, R, grads.data(), w));
    if (lr <= 0.0)
      hipLaunchKernelGGL(gamma_norms_kernel, dim3(1), dim3(256), 0, ctx->stream, w.G, nd, (int)R, w.out + 3);
    for (int k = 0; k < nd; ++k)
      hipLaunchKernelGGL(axpy_step_kernel, dim3(cdiv(shape[k] * R, 256)), dim3(256), 0, ctx->stream, F[k], grads[k],
                         shape[k] * R, lr, w.out + 3, nd);
    TENRSO_HIP_TRY(hipGetLastError());
    if (loss_hist) {
      TENRSO_HIP_TRY(hipMemcpyAsync(hs, w.out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
      loss_hist[s] = hs[0];  // loss at the point the step was taken from
    }
  }
  TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return TENRSO_OK;
}

}  // namespace tenrso
