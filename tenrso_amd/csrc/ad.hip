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

// One workgroup of 1024 threads per row i of mode n; thread t is column
// c = t % CW of thread group t / CW (CW = 32 for R <= 32, else 64; NG = 1024 / CW
// groups):
//   M[i, c]    = FP64 sum of the MTTKRP split partials of (i, c): group g sums a
//                contiguous range of the splits (four independent chains, all
//                loads in flight), the NG group sums added in group order;
//   grad[i, c] = sum_r A[i, r] Gamma_n[r, c] - M[i, c], Gamma_n = hadamard_{k != n}
//                G_k formed on the fly: group g forms the terms of its range of
//                r (the Gram loads issued beside the partials), the group sums
//                added in order;
//   ip_part[i] = sum_c M[i, c] A[i, c]  (the loss's <X, Xhat> for the last mode).
// grad and ip_part may be null.  This replaces a reduction launch, a Gamma launch
// and a dot-product launch per mode; every sum has a fixed order.  (A first form
// of 256 threads in four split quarters ran each quarter's ~40 loads in chains
// of four: 15 us per call at 96^4 r32 with 170 splits.)
//
// Tree form (TREE): M comes from a two-mode partial MTTKRP Y (split partials
// [splits][yrows][R]) instead of the full one: M[i, c] = sum_j Y[i sa + j sb, c]
// B[j, c], j = 0 .. J-1, each group a contiguous range of j.
constexpr int GR_NE = 64;  // rank bound (SOLVE_RMAX)
constexpr int GR_T = 1024;
//
// One launch can serve two modes (the two modes of one tree pass read the same
// Y): workgroups [0, g0.rows) take g0, the rest g1 (g1.rows = 0: one mode).
struct GradTree {
  i64 yrows, J, sa, sb;
  const float* B;
};
struct GradRows {
  int n;            // mode
  i64 rows;         // I_n, one workgroup per row
  const float* A;   // A_n
  float* grad;      // or null
  double* ip_part;  // or null: <M, A> per row (the last mode)
  GradTree tr;      // TREE only
};
template <bool TREE>
__global__ __launch_bounds__(GR_T) void grad_rows_kernel(const float* __restrict__ part, int splits, int R,
                                                         const double* __restrict__ G, int nd, GradRows g0,
                                                         GradRows g1) {
  __shared__ double sq[GR_T];
  __shared__ double sM[GR_NE];
  __shared__ double sA[GR_NE];
  const bool second = (i64)blockIdx.x >= g0.rows;
  const GradRows g = second ? g1 : g0;
  const int n = g.n;
  const i64 rows = g.rows;
  const float* __restrict__ A = g.A;
  float* __restrict__ grad = g.grad;
  double* __restrict__ ip_part = g.ip_part;
  const GradTree& tr = g.tr;
  const int cw = R <= 32 ? 32 : 64, ng = GR_T / cw;
  const int t = threadIdx.x, c = t % cw, grp = t / cw;
  const i64 row = second ? (i64)blockIdx.x - g0.rows : (i64)blockIdx.x;
  const bool colok = c < R;
  // this group's rows r of Gamma_n's column c (at most four: R / ng <= 4)
  const int r0 = grp * R / ng, r1 = (grp + 1) * R / ng;
  double w[4] = {0.0, 0.0, 0.0, 0.0};
  if (grad && colok) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (r0 + k < r1) {
        double v = 1.0;
        for (int m = 0; m < nd; ++m)
          if (m != n) v = v * G[((i64)m * R + r0 + k) * R + c];
        w[k] = v;
      }
    }
  }
  double msum = 0.0;
  if (TREE && colok) {
    const i64 nn = tr.yrows * R;
    const i64 j0 = grp * tr.J / ng, j1 = (grp + 1) * tr.J / ng;
    double a0 = 0, a1 = 0;
    i64 j = j0;
    for (; j + 2 <= j1; j += 2) {
      const i64 y0 = (row * tr.sa + j * tr.sb) * R + c, y1 = y0 + tr.sb * R;
      double v0 = 0.0, v1 = 0.0;
      for (int sp = 0; sp < splits; ++sp) {
        v0 += (double)part[(i64)sp * nn + y0];
        v1 += (double)part[(i64)sp * nn + y1];
      }
      a0 += v0 * (double)tr.B[j * R + c];
      a1 += v1 * (double)tr.B[(j + 1) * R + c];
    }
    if (j < j1) {
      const i64 y0 = (row * tr.sa + j * tr.sb) * R + c;
      double v0 = 0.0;
      for (int sp = 0; sp < splits; ++sp) v0 += (double)part[(i64)sp * nn + y0];
      a0 += v0 * (double)tr.B[j * R + c];
    }
    msum = a0 + a1;
  } else if (colok) {
    const i64 nn = rows * R, off = row * R + c;
    const int s0 = (int)((i64)grp * splits / ng), s1 = (int)((i64)(grp + 1) * splits / ng);
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    int sp = s0;
    for (; sp + 4 <= s1; sp += 4) {
      a0 += (double)part[(i64)(sp + 0) * nn + off];
      a1 += (double)part[(i64)(sp + 1) * nn + off];
      a2 += (double)part[(i64)(sp + 2) * nn + off];
      a3 += (double)part[(i64)(sp + 3) * nn + off];
    }
    for (; sp < s1; ++sp) a0 += (double)part[(i64)sp * nn + off];
    msum = (a0 + a1) + (a2 + a3);
  }
  if (t < R) sA[t] = (double)A[row * R + t];
  sq[t] = msum;
  __syncthreads();
  if (t < R) {
    double s = 0.0;
    for (int g = 0; g < ng; ++g) s += sq[g * cw + t];
    sM[t] = s;
  }
  __syncthreads();
  if (grad) {
    double x = 0.0;
    if (colok) {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (r0 + k < r1) x += sA[r0 + k] * w[k];
    }
    sq[t] = x;
    __syncthreads();
    if (t < R) {
      double s = 0.0;
      for (int g = 0; g < ng; ++g) s += sq[g * cw + t];
      grad[row * R + t] = (float)(s - sM[t]);
    }
  }
  if (ip_part && t < 64) {
    double ip = t < R ? sM[t] * sA[t] : 0.0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) ip += __shfl_xor(ip, o, 64);
    if (t == 0) ip_part[row] = ip;
  }
}

// loss = 0.5 (||X||^2 - 2 sum(part) + sum_rs prod_k G_k[r,s]) into out[0] and, when
// hist is not null, into *hist (the device-side loss history of a descent)
__device__ void gamma_norm_block(const double* __restrict__ G, int nd, int R, int n, double* __restrict__ out);

// Workgroup 0: the loss.  With gnorm set, workgroups 1 .. nd: ||Gamma_n||_F of
// mode n = blockIdx.x - 1 into gnorm[n] (the step size of a descent step).
__global__ __launch_bounds__(256) void loss_kernel(const double* __restrict__ G, int nd, int R,
                                                   const double* __restrict__ part, int npart,
                                                   const double* __restrict__ scal, double* __restrict__ out,
                                                   double* __restrict__ hist, double* __restrict__ gnorm) {
  if (blockIdx.x > 0) {
    gamma_norm_block(G, nd, R, (int)blockIdx.x - 1, gnorm);
    return;
  }
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
    const double loss = 0.5 * (scal[SC_NORM_X] - 2.0 * red[0] + recon);
    out[0] = loss;
    out[1] = red[0];
    out[2] = recon;
    if (hist) *hist = loss;
  }
}

// ||Gamma_n||_F for every mode (used by the automatic step size)
// ||Gamma_n||_F of mode n into out[n], by one workgroup of 256 threads
__device__ void gamma_norm_block(const double* __restrict__ G, int nd, int R, int n, double* __restrict__ out) {
  __shared__ double red[256];
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
}

// A_k -= step * grad_k for every factor in one launch (the factors' elements are
// laid end to end: element e belongs to the factor k with off[k] <= e < off[k+1]),
// step = lr (lr > 0) or 1 / max_n ||Gamma_n||_F read from gnorm
struct AxpyMany {
  float* A[MAXD];
  const float* g[MAXD];
  i64 off[MAXD + 1];
};
__global__ __launch_bounds__(256) void axpy_step_kernel(AxpyMany a, int nd, double lr,
                                                        const double* __restrict__ gnorm) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e >= a.off[nd]) return;
  int k = 0;
  while (e >= a.off[k + 1]) ++k;
  const i64 i = e - a.off[k];
  double step = lr;
  if (lr <= 0.0) {
    double mx = 0.0;
    for (int m = 0; m < nd; ++m) mx = fmax(mx, gnorm[m]);
    step = mx > 0.0 ? 1.0 / mx : 0.0;
  }
  a.A[k][i] = (float)((double)a.A[k][i] - step * (double)a.g[k][i]);
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
    if (!grads[n]) continue;
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
  // <M, A> partials (one per row of the last mode) and the norm's partials
  TENRSO_TRY(ctx_buffer(ctx, SLOT_IP, (size_t)std::max<i64>(4096, maxI + 16) * sizeof(double), &pp));
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

// grads[n] = A_n Gamma_n - mttkrp(X, A, n); loss into w.out[0] (device, and into
// *hist when not null), with ||X||^2 already in w.scal[SC_NORM_X].  Per mode: the
// MTTKRP as split partials, then one fused row kernel; modes without a gradient
// to form are skipped except the last, whose MTTKRP gives <X, Xhat>.  With `tree`
// (four modes) two partial MTTKRPs serve all four modes instead.
static int loss_grad_device(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, const float* const* F, i64 R,
                            float* const* grads, LossWork& w, double* hist, const float* Xt = nullptr,
                            double* gnorm = nullptr, bool tree = false) {
  TENRSO_TRY(gram_f64_many(ctx, nd, F, shape, R, w.G));
  if (tree) {
    // 4-way dimension tree: two passes over X instead of four.
    //   Y01[i0 I1 + i1, r] = sum_{i2,i3} X A2[i2,r] A3[i3,r]: mode 0 of the view
    //     [I0 I1, I2, I3];  M_0 = sum_{i1} Y01 A1,  M_1 = sum_{i0} Y01 A0;
    //   Y23[i2 I3 + i3, r] = sum_{i0,i1} X A0[i0,r] A1[i1,r]: mode 2 (the last) of
    //     the view [I0, I1, I2 I3];  M_2 = sum_{i3} Y23 A3,  M_3 = sum_{i2} Y23 A2.
    // Each Y stays as the MTTKRP's split partials, consumed before the next pass
    // reuses that buffer (stream order).
    const i64 I0 = shape[0], I1 = shape[1], I2 = shape[2], I3 = shape[3];
    for (int half = 0; half < 2; ++half) {
      const i64 v[3] = {half == 0 ? I0 * I1 : I0, half == 0 ? I2 : I1, half == 0 ? I3 : I2 * I3};
      const float* Fv[3] = {half == 0 ? F[0] : F[0], half == 0 ? F[2] : F[1], half == 0 ? F[3] : F[2]};
      MttkrpResult mr;
      TENRSO_TRY(mttkrp_dense(ctx, TENRSO_PREC_FP32, X, nullptr, 3, v, Fv, R, half == 0 ? 0 : 2,
                              MttkrpOut::PARTIAL, nullptr, &mr));
      const i64 ya = half == 0 ? I0 : I2, yb = half == 0 ? I1 : I3;  // Y rows = ya x yb
      // both modes of the pass in one launch (the first: M[i] sums over the
      // second index of Y; the second: over the first)
      GradRows gr[2] = {};
      int ng = 0;
      for (int k = 0; k < 2; ++k) {
        const int n = 2 * half + k;
        float* gn = (grads && grads[n]) ? grads[n] : nullptr;
        const bool last = n == nd - 1;
        if (!gn && !last) continue;
        gr[ng++] = GradRows{n, shape[n], F[n], gn, last ? w.part : nullptr,
                            GradTree{ya * yb, k == 0 ? yb : ya, k == 0 ? yb : 1, k == 0 ? 1 : yb,
                                     F[k == 0 ? n + 1 : n - 1]}};
      }
      if (ng > 0) {
        hipLaunchKernelGGL(grad_rows_kernel<true>, dim3((unsigned)(gr[0].rows + gr[1].rows)), dim3(GR_T), 0,
                           ctx->stream, mr.partial, mr.splits, (int)R, w.G, nd, gr[0], gr[1]);
        TENRSO_HIP_TRY(hipGetLastError());
      }
    }
    hipLaunchKernelGGL(loss_kernel, dim3(gnorm ? 1 + nd : 1), dim3(256), 0, ctx->stream, w.G, nd, (int)R, w.part,
                       (int)shape[nd - 1], w.scal, w.out, hist, gnorm);
    TENRSO_HIP_TRY(hipGetLastError());
    return TENRSO_OK;
  }
  // tensors of order >= 4: every mode's Khatri-Rao PRE rows in one launch
  const float* pre[MAXD] = {};
  if (nd >= 4) {
    int modes[MAXD];
    for (int n = 0; n < nd; ++n) modes[n] = n;
    TENRSO_TRY(mttkrp_pre_batch(ctx, nd, shape, F, R, nd, modes, Xt != nullptr, pre));
  }
  for (int n = 0; n < nd; ++n) {
    float* gn = (grads && grads[n]) ? grads[n] : nullptr;
    const bool last = n == nd - 1;
    if (!gn && !last) continue;
    MttkrpResult mr;
    TENRSO_TRY(mttkrp_dense(ctx, TENRSO_PREC_FP32, X, Xt, nd, shape, F, R, n, MttkrpOut::PARTIAL, nullptr, &mr,
                            nullptr, false, nullptr, 0, 0, pre[n]));
    if (shape[n] == 0) continue;
    hipLaunchKernelGGL(grad_rows_kernel<false>, dim3((unsigned)shape[n]), dim3(GR_T), 0, ctx->stream, mr.partial,
                       mr.splits, (int)R, w.G, nd, GradRows{n, shape[n], F[n], gn, last ? w.part : nullptr, {}},
                       GradRows{});
    TENRSO_HIP_TRY(hipGetLastError());
  }
  hipLaunchKernelGGL(loss_kernel, dim3(gnorm ? 1 + nd : 1), dim3(256), 0, ctx->stream, w.G, nd, (int)R, w.part,
                     (int)shape[nd - 1], w.scal, w.out, hist, gnorm);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

int cp_loss_grad_dev(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, const float* const* F, i64 R,
                     float* const* grads, const double** loss_dev) {
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
  TENRSO_TRY(loss_grad_device(ctx, X, nd, shape, F, R, grads, w, nullptr, nullptr, nullptr, nd == 4 && total > 0));
  if (loss_dev) *loss_dev = w.out;
  return TENRSO_OK;
}

int cp_loss_grad(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, const float* const* F, i64 R,
                 float* const* grads, double* loss_host) {
  const double* out = nullptr;
  TENRSO_TRY(cp_loss_grad_dev(ctx, X, nd, shape, F, R, grads, &out));
  double* hs = (double*)ctx->pinned;
  TENRSO_HIP_TRY(hipMemcpyAsync(hs, out, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (loss_host) *loss_host = hs[0];
  return TENRSO_OK;
}

// Gradient descent: every step is queued without a host round trip (the loss of
// step s goes to a device history, copied to the host once at the end).
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
  void *pg = nullptr, *ph = nullptr;
  TENRSO_TRY(ctx_buffer(ctx, SLOT_STAGE_OUT, (size_t)ftot * sizeof(float), &pg));
  if (loss_hist && steps > 0) TENRSO_TRY(ctx_buffer(ctx, SLOT_TMP2, (size_t)steps * sizeof(double), &ph));
  double* hist_dev = (double*)ph;
  std::vector<float*> grads(nd);
  AxpyMany am{};
  i64 off = 0;
  for (int k = 0; k < nd; ++k) {
    grads[k] = (float*)pg + off;
    am.A[k] = F[k];
    am.g[k] = grads[k];
    am.off[k] = off;
    off += shape[k] * R;
  }
  am.off[nd] = off;
  TENRSO_TRY(fill_f64(ctx, w.scal, 64, 0.0));
  // X does not change over the steps: one last-to-front copy lets the last
  // mode's MTTKRP of every step stream X with 16-byte q-contiguous loads
  // (measured at 96^4 r32: that mode from 82 to the ~63 us of mode 0).  The
  // copy and ||X||^2 come from one pass over X where the shape allows it.
  // Four modes take the dimension tree instead (two passes over X per step).
  const bool tree = nd == 4 && total > 0;
  const float* Xt = nullptr;
  bool normed = false;
  if (steps > 1 && nd >= 3 && total > 0 && !tree) {
    void* pt = nullptr;
    TENRSO_TRY(ctx_buffer(ctx, SLOT_XBF16_T, (size_t)total * sizeof(float), &pt));
    int nblk = 0;
    if (prepare_bf16_als(ctx, X, nd, shape, nullptr, (float*)pt, w.part, &nblk) == TENRSO_OK) {
      TENRSO_TRY(norm_min_finish(ctx, w.part, nblk, w.scal, 0));
      normed = true;
    } else {
      TENRSO_TRY(transpose_last_to_front_f32(ctx, X, nd, shape, (float*)pt));
    }
    Xt = (const float*)pt;
  }
  if (!normed) TENRSO_TRY(tensor_norm_min(ctx, X, total, w.part, w.scal));
  for (int s = 0; s < steps; ++s) {
    // loss at the point the step is taken from
    // with lr <= 0 the loss launch also leaves ||Gamma_n||_F (the step size)
    TENRSO_TRY(loss_grad_device(ctx, X, nd, shape, (const float* const*)F, R, grads.data(), w,
                                hist_dev ? hist_dev + s : nullptr, Xt, lr <= 0.0 ? w.out + 3 : nullptr, tree));
    if (ftot > 0)
      hipLaunchKernelGGL(axpy_step_kernel, dim3(cdiv(ftot, 256)), dim3(256), 0, ctx->stream, am, nd, lr, w.out + 3);
    TENRSO_HIP_TRY(hipGetLastError());
  }
  if (hist_dev)
    TENRSO_HIP_TRY(hipMemcpyAsync(loss_hist, hist_dev, (size_t)steps * sizeof(double), hipMemcpyDeviceToHost,
                                  ctx->stream));
  TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return TENRSO_OK;
}

}  // namespace tenrso
