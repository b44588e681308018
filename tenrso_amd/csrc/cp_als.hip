// cp_als.hip -- host driver of CP-ALS on the device.
//
// Restates the control flow of tenrso_decomp::cp_als (cp/core.rs:59-182) and
// cp_als_constrained (cp/core.rs:225-408): the same validation order, the same
// Gauss-Seidel mode sweep, the same fit / oscillation / tolerance / time-limit
// rules.  The arithmetic runs on the GPU:
//   per mode : MTTKRP (fp32 or bf16 MFMA, split partials summed in FP64)
//              -> Gamma^{-1} (FP64, one workgroup)
//              -> rows A_n = epi(M Gamma^{-1}) (FP64, written as fp32)
//              -> Gram A_n^T A_n (FP64) kept for the next modes and the fit
//   per iter : fit from the Grams and <X, Xhat> = sum M_{N-1} . A_{N-1}
//              (the last mode's MTTKRP, which equals the reference's separate
//              mode-0 MTTKRP of cp/helpers.rs:558-579 in exact arithmetic).
// The only host<->device traffic per iteration is the scalar block read back
// for the convergence test.
#include <chrono>
#include <cmath>
#include <cstring>

#include "cp_als.hpp"
#include "cp_linalg.hpp"
#include "mttkrp.hpp"

namespace tenrso {

int cp_als_run(tenrso_ctx* ctx, AlsProblem& pb, float* const* F, const tenrso_cp_options* o, tenrso_cp_result* res,
               double* hist) {
  const int nd = pb.nd;
  const i64* shape = pb.shape;
  const i64 R = o->rank;
  // validation order of cp/core.rs:82-95, 250-275
  if (R == 0) {
    set_error("Invalid rank: %lld", (long long)R);
    return TENRSO_ERR_INVALID_RANK;
  }
  for (int k = 0; k < nd; ++k)
    if (R > shape[k] || R < 0) {
      set_error("Invalid rank: %lld", (long long)R);
      return TENRSO_ERR_INVALID_RANK;
    }
  if (!(o->tol >= 0.0 && o->tol < 1.0)) {
    set_error("Invalid tolerance: %g", o->tol);
    return TENRSO_ERR_INVALID_TOLERANCE;
  }
  if (o->l2 < 0.0 || o->l1 < 0.0 || o->tikhonov_lambda < 0.0) {
    set_error("Invalid regularization parameter: lambda must be >= 0");
    return TENRSO_ERR_INVALID_REGULARIZATION;
  }
  if (nd < 2) {
    set_error("Shape mismatch: Need at least 2 factor matrices for MTTKRP");
    return TENRSO_ERR_SHAPE_MISMATCH;
  }
  if (nd > MAXD) {
    set_error("tensors with more than %d modes are not supported", MAXD);
    return TENRSO_ERR_UNSUPPORTED;
  }
  if (R > SOLVE_RMAX) {
    set_error("rank %lld exceeds the on-device normal-equation solve bound %d", (long long)R, SOLVE_RMAX);
    return TENRSO_ERR_UNSUPPORTED;
  }
  if (o->orthogonal) {
    set_error("orthogonal constraint is not supported by the device CP-ALS");
    return TENRSO_ERR_UNSUPPORTED;
  }
  TENRSO_TRY(set_device(ctx));
  i64 maxI = 0;
  for (int k = 0; k < nd; ++k) maxI = std::max(maxI, shape[k]);

  // ---- device workspace
  void *p_scal, *p_G, *p_Winv, *p_M, *p_tmp;
  TENRSO_TRY(ctx_buffer(ctx, SLOT_SCALARS, SC_COUNT * sizeof(double), &p_scal));
  TENRSO_TRY(ctx_buffer(ctx, SLOT_GRAMS, (size_t)(nd + 1) * R * R * sizeof(double), &p_G));
  TENRSO_TRY(ctx_buffer(ctx, SLOT_SOLVE, (size_t)SOLVE_RMAX * SOLVE_RMAX * sizeof(double), &p_Winv));
  TENRSO_TRY(ctx_buffer(ctx, SLOT_M64, (size_t)maxI * R * sizeof(double), &p_M));
  TENRSO_TRY(ctx_buffer(ctx, SLOT_IP, (size_t)std::max<i64>(4096, maxI + 16) * sizeof(double), &p_tmp));
  double* scal = (double*)p_scal;
  double* G = (double*)p_G;
  double* Winv = (double*)p_Winv;
  double* M = (double*)p_M;
  double* ip_part = (double*)p_tmp;

  TENRSO_TRY(fill_f64(ctx, scal, SC_COUNT, 0.0));
  TENRSO_TRY(pb.norm_min(ip_part, scal));

  const bool nonneg = o->nonnegative != 0;
  if (nonneg && o->validate_nonneg_input) {
    double* hs = (double*)ctx->pinned;
    TENRSO_HIP_TRY(hipMemcpyAsync(hs, scal, SC_COUNT * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (hs[SC_MIN_X] < 0.0) {
      set_error("Non-negative constraint violated: input tensor contains negative values");
      return TENRSO_ERR_NONNEGATIVE;
    }
  }
  if (nonneg)
    for (int k = 0; k < nd; ++k) TENRSO_TRY(clamp_nonneg(ctx, F[k], shape[k] * R));  // core.rs:281-285
  if (pb.prepare) TENRSO_TRY(pb.prepare());

  for (int k = 0; k < nd; ++k) TENRSO_TRY(gram_f64(ctx, F[k], shape[k], R, G + (i64)k * R * R));

  GramReg reg;
  reg.l2 = o->l2;
  reg.tikhonov_order = o->tikhonov_order;
  reg.tikhonov_lambda = o->tikhonov_lambda;
  RowEpilogue epi;
  epi.l1 = o->l1;
  epi.nonneg = nonneg ? 1 : 0;

  double prev_fit = 0.0, fit = 0.0, final_change = 0.0;
  int iters = 0, osc = 0, reason = TENRSO_REASON_MAX_ITERATIONS, worst = 0;
  const float* const* Fc = (const float* const*)F;
  const auto t0 = std::chrono::steady_clock::now();
  for (int iter = 0; iter < o->max_iters; ++iter) {
    if (o->time_limit_s >= 0.0) {
      const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (el > o->time_limit_s) {
        reason = TENRSO_REASON_TIME_LIMIT;
        break;
      }
    }
    iters = iter + 1;
    int nblk = 0;
    for (int mode = 0; mode < nd; ++mode) {
      const i64 I = shape[mode];
      // Side stream: the Gram of the factor updated just before (mode - 1) and
      // Gamma_n^{-1}; both only need factors that are final for this mode, so
      // they run while the MTTKRP of mode n runs on the main stream.
      TENRSO_HIP_TRY(hipEventRecord(ctx->ev_g, ctx->stream));
      hipStream_t side = ctx->serial_solve ? ctx->stream : ctx->aux;
      if (!ctx->serial_solve) TENRSO_HIP_TRY(hipStreamWaitEvent(ctx->aux, ctx->ev_g, 0));
      if (mode > 0) TENRSO_TRY(gram_f64(ctx, F[mode - 1], shape[mode - 1], R, G + (i64)(mode - 1) * R * R, side));
      TENRSO_TRY(normal_eq_inverse(ctx, G, nd, mode, R, reg, Winv, scal, side));
      TENRSO_HIP_TRY(hipEventRecord(ctx->ev_w, side));
      MttkrpResult mr;
      TENRSO_TRY(pb.mttkrp(mode, Fc, &mr));
      if (!ctx->serial_solve) TENRSO_HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->ev_w, 0));
      TENRSO_TRY(solve_rows(ctx, mr.partial, mr.splits, nullptr, I, R, Winv, epi, F[mode], ip_part, &nblk));
    }
    // Gram of the last factor (needed by the fit and by the next iteration)
    TENRSO_TRY(gram_f64(ctx, F[nd - 1], shape[nd - 1], R, G + (i64)(nd - 1) * R * R));
    if (o->fit_from_mode0) {
      // cp/helpers.rs:558-579: <X, Xhat> from a mode-0 MTTKRP with every updated factor
      MttkrpResult mr;
      TENRSO_TRY(pb.mttkrp(0, Fc, &mr));
      TENRSO_TRY(reduce_partials(ctx, mr.partial, mr.splits, shape[0] * R, nullptr, M));
      TENRSO_TRY(inner_rows(ctx, M, F[0], shape[0], R, ip_part, &nblk));
    }
    TENRSO_TRY(fit_from_grams(ctx, G, nd, R, ip_part, nblk, scal));
    double* hs = (double*)ctx->pinned;
    TENRSO_HIP_TRY(hipMemcpyAsync(hs, scal, SC_COUNT * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
    worst = std::max(worst, (int)hs[SC_STATUS]);
    if (worst >= 2) {
      set_error("Linear algebra error: the Gram-Hadamard matrix is not positive definite after the ridge retry");
      return TENRSO_ERR_LINALG;
    }
    // fit bookkeeping in FP64 (the fit itself is an FP64 reduction on the device)
    fit = hs[SC_FIT];
    if (hist) hist[iter] = fit;
    if (iter > 0 && fit < prev_fit) osc += 1;
    const double change = std::fabs(fit - prev_fit);
    final_change = change;
    if (iter > 0 && change < o->tol) {
      reason = TENRSO_REASON_FIT_TOLERANCE;
      break;
    }
    if (osc > 5 && iter > 10) {
      reason = TENRSO_REASON_OSCILLATION;
      break;
    }
    prev_fit = fit;
  }
  TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (res) {
    res->iters = iters;
    res->reason = reason;
    res->oscillation_count = osc;
    res->oscillated = osc > 0;
    res->fit = fit;
    res->final_fit_change = final_change;
    res->solve_path_max = worst;
    res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return TENRSO_OK;
}

int cp_als_device(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, float* const* F,
                  const tenrso_cp_options* o, tenrso_cp_result* res, double* hist) {
  const bool bf16 = o->precision == TENRSO_PREC_BF16;
  if (o->precision != TENRSO_PREC_FP32 && !bf16) {
    set_error("unknown precision %d", o->precision);
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  i64 total = 1;
  for (int k = 0; k < nd; ++k) total *= shape[k];
  const void* Xk = X;
  const void* Xt = nullptr;
  const i64 R = o->rank;
  AlsProblem pb;
  pb.nd = nd;
  pb.shape = shape;
  pb.norm_min = [&](double* tmp, double* scal) { return tensor_norm_min(ctx, X, total, tmp, scal); };
  pb.prepare = [&]() -> int {
    if (!bf16) return TENRSO_OK;
    void *pb16, *pt;
    TENRSO_TRY(ctx_buffer(ctx, SLOT_XBF16, (size_t)total * 2, &pb16));
    TENRSO_TRY(ctx_buffer(ctx, SLOT_XBF16_T, (size_t)total * 2, &pt));
    TENRSO_TRY(convert_bf16(ctx, X, nd, shape, false, (uint16_t*)pb16));
    TENRSO_TRY(convert_bf16(ctx, X, nd, shape, true, (uint16_t*)pt));
    Xk = pb16;
    Xt = pt;
    return TENRSO_OK;
  };
  pb.mttkrp = [&](int mode, const float* const* Fc, MttkrpResult* mr) {
    return mttkrp_dense(ctx, bf16 ? TENRSO_PREC_BF16 : TENRSO_PREC_FP32, Xk, Xt, nd, shape, Fc, R, mode,
                        MttkrpOut::PARTIAL, nullptr, mr);
  };
  return cp_als_run(ctx, pb, F, o, res, hist);
}

}  // namespace tenrso
