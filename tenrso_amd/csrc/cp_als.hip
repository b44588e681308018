// cp_als.hip -- host driver of CP-ALS on the device.
//
// Restates the control flow of tenrso_decomp::cp_als (cp/core.rs:59-182) and
// cp_als_constrained (cp/core.rs:225-408): the same validation order, the same
// Gauss-Seidel mode sweep, the same fit / oscillation / tolerance / time-limit
// rules.  The arithmetic runs on the GPU, per mode n of a sweep:
//   one launch   MTTKRP of mode n (fp32 or bf16 MFMA, split partials) whose
//                workgroup 0 runs the FP64 job: the previous sweep's fit (at
//                n = 0), Gamma_n and its inverse -- so the latency-bound linear
//                algebra shares the GPU with the bandwidth-bound MTTKRP instead
//                of waiting behind it;
//   one launch   rows A_n = epi(M Gamma_n^{-1}) (FP64, written as fp32) and the
//                <M, A_n> partials;
//   one launch   the Gram A_n^T A_n (FP64) for the modes that follow.
// Fit: from the Grams and <X, Xhat> = sum M_{N-1} . A_{N-1} (the last mode's
// MTTKRP, equal in exact arithmetic to the reference's separate mode-0 MTTKRP of
// cp/helpers.rs:558-579, which fit_from_mode0 selects).  In BF16 mode that last
// MTTKRP contracts the fp32 tensor exactly, so the fit is the fit to X.
//
// Sweep k reads factor set k % 2 and writes set (k + 1) % 2 (set 0 is the
// caller's factors, set 1 context scratch).  So the host can queue sweep k + 1
// before it has read the fit of sweep k: if sweep k turns out to be the last,
// sweep k + 1 has not touched its result.  The only host<->device traffic per
// sweep is the scalar block copied back for the convergence test.
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>

#include "als_job.hpp"
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
  // ranks above JOB_NE take the wide-rank job and row solve (cp_linalg.hip)
  if (R > WIDE_NE) {
    set_error("rank %lld exceeds the on-device normal-equation solve bound %d", (long long)R, WIDE_NE);
    return TENRSO_ERR_UNSUPPORTED;
  }
  TENRSO_TRY(set_device(ctx));
  i64 maxI = 0, ftot = 0;
  for (int k = 0; k < nd; ++k) {
    maxI = std::max(maxI, shape[k]);
    ftot += shape[k] * R;
  }

  // ---- device workspace
  void *p_scal, *p_G, *p_Winv, *p_tmp, *p_set;
  TENRSO_TRY(ctx_buffer(ctx, SLOT_SCALARS, SC_COUNT * sizeof(double), &p_scal));
  TENRSO_TRY(ctx_buffer(ctx, SLOT_GRAMS, (size_t)(nd + 1) * R * R * sizeof(double), &p_G));
  const i64 ne = R > JOB_NE ? WIDE_NE : JOB_NE;  // order (and stride) of Winv
  TENRSO_TRY(ctx_buffer(ctx, SLOT_SOLVE, (size_t)ne * ne * sizeof(double), &p_Winv));
  TENRSO_TRY(ctx_buffer(ctx, SLOT_IP, (size_t)std::max<i64>(4096, maxI + 16) * sizeof(double), &p_tmp));
  TENRSO_TRY(ctx_buffer(ctx, SLOT_FSET, (size_t)ftot * sizeof(float), &p_set));
  double* scal = (double*)p_scal;
  double* G = (double*)p_G;
  double* Winv = (double*)p_Winv;
  double* ip_part = (double*)p_tmp;
  double* M = nullptr;
  // the orthogonal constraint changes A_{N-1} after its solve, so the fit takes
  // <X, Xhat> from a mode-0 MTTKRP of the final factors, as the reference does
  const bool fit0 = o->fit_from_mode0 || o->orthogonal;
  if (fit0) {
    void* p_M;
    TENRSO_TRY(ctx_buffer(ctx, SLOT_M64, (size_t)maxI * R * sizeof(double), &p_M));
    M = (double*)p_M;
  }
  float* set[2][MAXD];
  {
    i64 off = 0;
    for (int k = 0; k < nd; ++k) {
      set[0][k] = F[k];
      set[1][k] = (float*)p_set + off;
      off += shape[k] * R;
    }
  }

  TENRSO_TRY(pb.norm_min(ip_part, scal));  // also clears the rest of the scalar block

  const bool nonneg = o->nonnegative != 0;
  double* pinned = (double*)ctx->pinned;  // [2][SC_COUNT]
  if (nonneg && o->validate_nonneg_input) {
    TENRSO_HIP_TRY(hipMemcpyAsync(pinned, scal, SC_COUNT * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (pinned[SC_MIN_X] < 0.0) {
      set_error("Non-negative constraint violated: input tensor contains negative values");
      return TENRSO_ERR_NONNEGATIVE;
    }
  }
  if (nonneg)
    for (int k = 0; k < nd; ++k) TENRSO_TRY(clamp_nonneg(ctx, F[k], shape[k] * R));  // core.rs:281-285
  if (pb.prepare) TENRSO_TRY(pb.prepare());

  TENRSO_TRY(gram_f64_many(ctx, nd, (const float* const*)F, shape, R, G));

  RowEpilogue epi;
  epi.l1 = o->l1;
  epi.nonneg = nonneg ? 1 : 0;
  int nblk = 0;  // <M, A> partials written by the last row solve

  auto base_job = [&]() {
    AlsJob j{};
    j.G = G;
    j.nd = nd;
    j.R = (int)R;
    j.ip_part = ip_part;
    j.nblk = nblk;
    j.skip = -1;
    j.l2 = o->l2;
    j.tik_order = o->tikhonov_order;
    j.tik_lambda = o->tikhonov_lambda;
    j.Winv = Winv;
    j.scal = scal;
    j.status_slot = SC_STATUS;
    return j;
  };
  // The job evaluating the fit of sweep k writes the scalar block into
  // pinned[k % 2] (mapped, coherent host memory) and then stamps it with k; the
  // host polls the stamp.  (An event between the launches would cost the GPU
  // an idle gap per sweep.)
  static_assert(JOB_SC_COUNT == SC_COUNT && JOB_SC_STAMP < SC_COUNT, "scalar block layout");
  auto host_block = [&](int k) { return ctx->pinned_dev + (k & 1) * SC_COUNT; };
  pinned[JOB_SC_STAMP] = -1.0;
  pinned[SC_COUNT + JOB_SC_STAMP] = -1.0;
  auto wait_fit = [&](int k) -> int {
    volatile double* stamp = pinned + (k & 1) * SC_COUNT + JOB_SC_STAMP;
    const auto t_spin = std::chrono::steady_clock::now();
    while (*stamp != (double)k) {
      if (std::chrono::steady_clock::now() - t_spin > std::chrono::milliseconds(2)) {
        // long sweep (or a failed launch): block on the stream, then the stamp must be there
        TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (*stamp != (double)k) {
          set_error("internal: the fit of sweep %d was not published", k);
          return TENRSO_ERR_DEVICE;
        }
        break;
      }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return TENRSO_OK;
  };
  // rows of mode n, then their Gram
  auto solve_and_gram = [&](int n, const MttkrpResult& mr, float* An) -> int {
    TENRSO_TRY(solve_rows(ctx, mr.partial, mr.splits, nullptr, shape[n], R, Winv, epi, An, ip_part, &nblk));
    // orthogonal constraint (core.rs:362-365): A_n <- Q of its QR
    if (o->orthogonal) TENRSO_TRY(orthonormalize_rows(ctx, An, shape[n], R));
    TENRSO_TRY(gram_f64(ctx, An, shape[n], R, G + (i64)n * R * R));
    return TENRSO_OK;
  };
  // queue sweep k; with fit_prev its first job also evaluates the fit of sweep k - 1
  auto enqueue_sweep = [&](int k, bool fit_prev) -> int {
    float* const* olds = set[k & 1];
    float* const* news = set[(k + 1) & 1];
    for (int n = 0; n < nd; ++n) {
      const float* Fc[MAXD];
      for (int m = 0; m < nd; ++m) Fc[m] = m < n ? news[m] : olds[m];
      AlsJob job = base_job();
      job.do_fit = (n == 0 && fit_prev) ? 1 : 0;
      if (job.do_fit) {
        job.host = host_block(k - 1);
        job.stamp = (double)(k - 1);
      }
      job.skip = n;
      job.status_slot = (k & 1) ? SC_STATUS_ODD : SC_STATUS;
      job.status_reset = n == 0 ? 1 : 0;
      MttkrpResult mr;
      TENRSO_TRY(pb.mttkrp(n, Fc, &job, &mr));
      TENRSO_TRY(solve_and_gram(n, mr, news[n]));
    }
    return TENRSO_OK;
  };
  // fit of sweep k - 1 on its own (no further sweep queued behind it)
  auto enqueue_fit = [&](int k) -> int {
    float* const* cur = set[k & 1];
    if (fit0) {
      // cp/helpers.rs:558-579: <X, Xhat> from a mode-0 MTTKRP with every updated factor
      MttkrpResult mr;
      TENRSO_TRY(pb.mttkrp(0, (const float* const*)cur, nullptr, &mr));
      TENRSO_TRY(reduce_partials(ctx, mr.partial, mr.splits, shape[0] * R, nullptr, M));
      TENRSO_TRY(inner_rows(ctx, M, cur[0], shape[0], R, ip_part, &nblk));
    }
    AlsJob job = base_job();
    job.do_fit = 1;
    job.host = host_block(k - 1);
    job.stamp = (double)(k - 1);
    return als_job_launch(ctx, job);
  };

  // speculation (queue sweep k before reading the fit of k - 1) keeps the GPU
  // busy through the host's convergence test; with a time limit or the
  // reference fit form each sweep is queued only after that test.
  const bool spec = !fit0 && !(o->time_limit_s >= 0.0);
  double prev_fit = 0.0, fit = 0.0, final_change = 0.0;
  int iters = 0, osc = 0, reason = TENRSO_REASON_MAX_ITERATIONS, worst = 0;
  const auto t0 = std::chrono::steady_clock::now();
  auto time_up = [&]() {
    if (!(o->time_limit_s >= 0.0)) return false;
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > o->time_limit_s;
  };
  int status = TENRSO_OK;
  if (o->max_iters > 0) {
    if (time_up()) {
      reason = TENRSO_REASON_TIME_LIMIT;
    } else {
      TENRSO_TRY(enqueue_sweep(0, false));
      for (int k = 1;; ++k) {
        // sweep k - 1 is queued
        const bool more = spec && k < o->max_iters;
        if (more) TENRSO_TRY(enqueue_sweep(k, true));
        else TENRSO_TRY(enqueue_fit(k));
        TENRSO_TRY(wait_fit(k - 1));
        const double* hs = pinned + ((k - 1) & 1) * SC_COUNT;
        iters = k;
        const int it = k - 1;
        worst = std::max(worst, (int)hs[(it & 1) ? SC_STATUS_ODD : SC_STATUS]);
        if (worst >= 2) {
          set_error("Linear algebra error: the Gram-Hadamard matrix is not positive definite after the ridge retry");
          status = TENRSO_ERR_LINALG;
          break;
        }
        fit = hs[SC_FIT];
        if (hist) hist[it] = fit;
        if (it > 0 && fit < prev_fit) osc += 1;
        const double change = std::fabs(fit - prev_fit);
        final_change = change;
        if (it > 0 && change < o->tol) {
          reason = TENRSO_REASON_FIT_TOLERANCE;
          break;
        }
        if (osc > 5 && it > 10) {
          reason = TENRSO_REASON_OSCILLATION;
          break;
        }
        prev_fit = fit;
        if (k >= o->max_iters) {
          reason = TENRSO_REASON_MAX_ITERATIONS;
          break;
        }
        if (!spec) {
          if (time_up()) {
            reason = TENRSO_REASON_TIME_LIMIT;
            break;
          }
          TENRSO_TRY(enqueue_sweep(k, false));
        }
      }
    }
  }
  TENRSO_TRY(job_join(ctx));  // a wide-rank job forked for a solve that an error skipped
  TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));  // drains a speculative sweep
  if (status != TENRSO_OK) return status;

  // the result of the last committed sweep (iters - 1) is in set iters % 2
  if ((iters & 1) == 1) {
    for (int k = 0; k < nd; ++k)
      TENRSO_HIP_TRY(hipMemcpyAsync(F[k], set[1][k], (size_t)shape[k] * R * sizeof(float), hipMemcpyDeviceToDevice,
                                    ctx->stream));
    TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
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
  // BF16: modes 0..N-2 contract the bf16 copy of X; the last mode, whose <M, A>
  // gives the fit, contracts the exact fp32 tensor through its last-to-front copy
  // (bf16x3 split on the MFMA path), so the fit is the fit to X itself.  The bf16
  // copy, the fp32 last-to-front copy and ||X||^2 come from one pass over X when
  // the shape allows it (prepare_bf16_als); otherwise norm, copy and transposed
  // copy are three passes
  bool prepared = false;
  // A last mode that is no multiple of 8 (bf16 copy) or a mode N-2 that is no
  // multiple of 4 (the fp32 last-to-front copy, whose last mode it is) would
  // leave rows off the MFMA kernels' load vectors: those copies get the mode
  // padded with zeros (ragged shapes such as 250^3), so every mode still runs
  // on the matrix cores.
  const i64 Kx = shape[nd - 1], Bt = nd >= 2 ? shape[nd - 2] : 1;
  // (the last-to-front copy's mode goes to a multiple of 32 where that adds at
  // most 1/8: 128-byte rows, and no masked loads)
  const i64 Kxp = (Kx + 7) / 8 * 8, Bt32a = (Bt + 31) / 32 * 32;
  const i64 Btp = Bt32a * 8 <= Bt * 9 ? Bt32a : (Bt + 3) / 4 * 4;
  const bool padded = bf16 && nd >= 2 && total > 0 && (Kxp != Kx || Btp != Bt);
  auto bf16_buffers = [&](void** pb16, void** pt) -> int {
    const size_t nb = padded ? (size_t)(total / Kx * Kxp) : (size_t)total;
    const size_t nt = padded ? (size_t)(total / Bt * Btp) : (size_t)total;
    TENRSO_TRY(ctx_buffer(ctx, SLOT_XBF16, nb * 2, pb16));
    TENRSO_TRY(ctx_buffer(ctx, SLOT_XBF16_T, nt * sizeof(float), pt));
    return TENRSO_OK;
  };
  pb.norm_min = [&](double* tmp, double* scal) -> int {
    if (bf16 && !padded) {
      void *pb16, *pt;
      TENRSO_TRY(bf16_buffers(&pb16, &pt));
      int nblk = 0;
      if (prepare_bf16_als(ctx, X, nd, shape, (uint16_t*)pb16, (float*)pt, tmp, &nblk) == TENRSO_OK) {
        Xk = pb16;
        Xt = pt;
        prepared = true;
        return norm_min_finish(ctx, tmp, nblk, scal, SC_COUNT);
      }
    }
    return tensor_norm_min(ctx, X, total, tmp, scal, SC_COUNT);
  };
  pb.prepare = [&]() -> int {
    if (!bf16 || prepared) return TENRSO_OK;
    void *pb16, *pt;
    TENRSO_TRY(bf16_buffers(&pb16, &pt));
    if (padded) {
      TENRSO_TRY(convert_bf16_padded(ctx, X, nd, shape, Kxp, (uint16_t*)pb16));
      TENRSO_TRY(transpose_last_to_front_f32_padded(ctx, X, nd, shape, Btp, (float*)pt));
    } else {
      TENRSO_TRY(convert_bf16(ctx, X, nd, shape, false, (uint16_t*)pb16));
      TENRSO_TRY(transpose_last_to_front_f32(ctx, X, nd, shape, (float*)pt));
    }
    Xk = pb16;
    Xt = pt;
    return TENRSO_OK;
  };
  // FP32 with a last mode of no multiple of 32 floats (250^3, say): its rows are
  // not 128-byte aligned, so every 16-byte vector run straddles cache lines, and
  // the last mode is i-contiguous and ragged.  Where padding adds at most 1/8,
  // modes 0..N-2 contract a copy with the last mode padded to a multiple of 32
  // (zeros) and the last mode the last-to-front copy, its own last mode padded
  // the same way.  One pass each per call.
  const i64 Kx32 = (Kx + 31) / 32 * 32, Bt32 = (Bt + 31) / 32 * 32;
  const bool fpad = !bf16 && nd >= 2 && total > 0 && Kx32 != Kx && Kx32 * 8 <= Kx * 9 && Bt32 * 8 <= Bt * 9;
  if (fpad) {
    pb.prepare = [&]() -> int {
      void *pp, *pt;
      TENRSO_TRY(ctx_buffer(ctx, SLOT_XPAD, (size_t)(total / Kx * Kx32) * sizeof(float), &pp));
      TENRSO_TRY(ctx_buffer(ctx, SLOT_XBF16_T, (size_t)(total / Bt * Bt32) * sizeof(float), &pt));
      TENRSO_TRY(copy_f32_padded(ctx, X, nd, shape, Kx32, (float*)pp));
      TENRSO_TRY(transpose_last_to_front_f32_padded(ctx, X, nd, shape, Bt32, (float*)pt));
      Xk = pp;
      Xt = pt;
      return TENRSO_OK;
    };
  }
  const i64 x_pad = padded ? Kxp : (fpad ? Kx32 : 0), xt_pad = padded ? Btp : (fpad ? Bt32 : 0);
  pb.mttkrp = [&](int mode, const float* const* Fc, const AlsJob* job, MttkrpResult* mr) {
    return mttkrp_dense(ctx, bf16 ? TENRSO_PREC_BF16 : TENRSO_PREC_FP32, Xk, Xt, nd, shape, Fc, R, mode,
                        MttkrpOut::PARTIAL, nullptr, mr, job, /*xt_f32=*/true, nullptr, x_pad, xt_pad);
  };
  return cp_als_run(ctx, pb, F, o, res, hist);
}

}  // namespace tenrso
