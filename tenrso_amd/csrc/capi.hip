// capi.hip -- extern "C" entry points (include/tenrso_hip.h).
//
// Argument validation reproduces the reference's checks and messages:
//   mttkrp      tenrso-kernels/src/mttkrp.rs:85-122 (+ 194-196 for < 2 factors)
//   cp_als      tenrso-decomp/src/cp/core.rs:82-95 (CpError Display strings, cp/types.rs:13-35)
#include <cstring>
#include <vector>

#include "cp_als.hpp"
#include "kr_reconstruct.hpp"
#include "mttkrp.hpp"
#include "cp_linalg.hpp"
#include "sparse.hpp"
#include "als_job.hpp"
#include "ad.hpp"
#include "cp_init.hpp"

using namespace tenrso;

namespace {

bool bad_ctx(tenrso_ctx* ctx) {
  if (!ctx) {
    set_error("null context");
    return true;
  }
  return false;
}

i64 prod(const int64_t* s, int n) {
  i64 p = 1;
  for (int k = 0; k < n; ++k) p *= s[k];
  return p;
}

}  // namespace

// the checks every MTTKRP entry point makes before touching the device
int tenrso::mttkrp_check_args(int ndim, const int64_t* shape, int64_t rank, int mode) {
  if (ndim <= 0 || !shape) {
    set_error("tensor must have at least one mode");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  if (mode < 0 || mode >= ndim) {
    set_error("Mode %d out of bounds for tensor with rank %d", mode, ndim);
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  if (ndim > MAXD) {
    set_error("tensors with more than %d modes are not supported", MAXD);
    return TENRSO_ERR_UNSUPPORTED;
  }
  for (int k = 0; k < ndim; ++k)
    if (shape[k] < 0) {
      set_error("negative mode size %lld", (long long)shape[k]);
      return TENRSO_ERR_INVALID_ARGUMENT;
    }
  if (rank < 0) {
    set_error("negative rank %lld", (long long)rank);
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  if (ndim < 2) {
    set_error("Need at least 2 factor matrices for MTTKRP");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  return TENRSO_OK;
}

namespace {

int check_mttkrp_basic(int ndim, const int64_t* shape, int64_t rank, int mode) {
  return mttkrp_check_args(ndim, shape, rank, mode);
}

int timed_begin(tenrso_ctx* ctx) {
  if (ctx->timing) TENRSO_HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
  return TENRSO_OK;
}
int timed_end(tenrso_ctx* ctx) {
  if (ctx->timing) {
    TENRSO_HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    TENRSO_HIP_TRY(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    TENRSO_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_kernel_ms = ms;
  }
  return TENRSO_OK;
}

}  // namespace

extern "C" {

int tenrso_mttkrp_check(int ndim, const int64_t* shape, int nfactors, const int64_t* factor_rows,
                        const int64_t* factor_cols, int mode) {
  clear_error();
  if (ndim <= 0 || !shape) {
    set_error("tensor must have at least one mode");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  if (mode < 0 || mode >= ndim) {
    set_error("Mode %d out of bounds for tensor with rank %d", mode, ndim);
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  if (nfactors != ndim) {
    set_error("Number of factor matrices (%d) must match tensor rank (%d)", nfactors, ndim);
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  const int64_t cp_rank = factor_cols[0];
  for (int i = 0; i < nfactors; ++i) {
    if (factor_cols[i] != cp_rank) {
      set_error("Factor matrix %d has %lld columns, expected %lld", i, (long long)factor_cols[i],
                (long long)cp_rank);
      return TENRSO_ERR_INVALID_ARGUMENT;
    }
    if (factor_rows[i] != shape[i]) {
      set_error("Factor matrix %d has %lld rows, expected %lld (tensor mode-%d size)", i,
                (long long)factor_rows[i], (long long)shape[i], i);
      return TENRSO_ERR_INVALID_ARGUMENT;
    }
  }
  if (ndim < 2) {
    set_error("Need at least 2 factor matrices for MTTKRP");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  return TENRSO_OK;
}

int tenrso_mttkrp_f32(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape, const float* const* factors,
                      int64_t rank, int mode, float* out) {
  clear_error();
  if (bad_ctx(ctx)) return TENRSO_ERR_INVALID_ARGUMENT;
  TENRSO_TRY(check_mttkrp_basic(ndim, shape, rank, mode));
  if (rank == 0 || shape[mode] == 0) return TENRSO_OK;
  TENRSO_TRY(set_device(ctx));
  TENRSO_TRY(timed_begin(ctx));
  TENRSO_TRY(mttkrp_dense(ctx, TENRSO_PREC_FP32, X, nullptr, ndim, shape, factors, rank, mode, MttkrpOut::F32, out,
                          nullptr));
  return timed_end(ctx);
}

int tenrso_mttkrp_f32_f64out(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape,
                             const float* const* factors, int64_t rank, int mode, double* out) {
  clear_error();
  if (bad_ctx(ctx)) return TENRSO_ERR_INVALID_ARGUMENT;
  TENRSO_TRY(check_mttkrp_basic(ndim, shape, rank, mode));
  if (rank == 0 || shape[mode] == 0) return TENRSO_OK;
  TENRSO_TRY(set_device(ctx));
  TENRSO_TRY(timed_begin(ctx));
  TENRSO_TRY(mttkrp_dense(ctx, TENRSO_PREC_FP32, X, nullptr, ndim, shape, factors, rank, mode, MttkrpOut::F64, out,
                          nullptr));
  return timed_end(ctx);
}

int tenrso_convert_f32_to_bf16(tenrso_ctx* ctx, const float* src, int64_t n, uint16_t* dst) {
  clear_error();
  if (bad_ctx(ctx)) return TENRSO_ERR_INVALID_ARGUMENT;
  if (n < 0) {
    set_error("negative length");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  TENRSO_TRY(set_device(ctx));
  const i64 shape1[1] = {n};
  return convert_bf16(ctx, src, 1, shape1, false, dst);
}

int tenrso_mttkrp_bf16(tenrso_ctx* ctx, const uint16_t* X_bf16, int ndim, const int64_t* shape,
                       const float* const* factors, int64_t rank, int mode, float* out) {
  clear_error();
  if (bad_ctx(ctx)) return TENRSO_ERR_INVALID_ARGUMENT;
  TENRSO_TRY(check_mttkrp_basic(ndim, shape, rank, mode));
  if (rank == 0 || shape[mode] == 0) return TENRSO_OK;
  TENRSO_TRY(set_device(ctx));
  const void* Xt = nullptr;
  if (mode == ndim - 1) {
    // last mode: move it to the front once (bf16 -> bf16 transpose)
    const i64 total = prod(shape, ndim);
    void* pt = nullptr;
    TENRSO_TRY(ctx_buffer(ctx, SLOT_XBF16_T, (size_t)total * 2 + 16, &pt));
    TENRSO_TRY(transpose_last_to_front_u16(ctx, X_bf16, ndim, shape, (uint16_t*)pt));
    Xt = pt;
  }
  TENRSO_TRY(timed_begin(ctx));
  TENRSO_TRY(mttkrp_dense(ctx, TENRSO_PREC_BF16, X_bf16, Xt, ndim, shape, factors, rank, mode, MttkrpOut::F32, out,
                          nullptr));
  return timed_end(ctx);
}

int tenrso_mttkrp_profile(tenrso_ctx* ctx, int precision, const void* X, int ndim, const int64_t* shape,
                          const float* const* factors, int64_t rank, int mode, int reps, double* ms) {
  clear_error();
  if (bad_ctx(ctx)) return TENRSO_ERR_INVALID_ARGUMENT;
  TENRSO_TRY(check_mttkrp_basic(ndim, shape, rank, mode));
  if (reps <= 0 || !ms) {
    set_error("reps must be positive and ms non-null");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  if (rank == 0 || shape[mode] == 0) {
    *ms = 0.0;
    return TENRSO_OK;
  }
  TENRSO_TRY(set_device(ctx));
  // precision 2: the three-way-split kernel of the BF16 ALS's last mode (fp32 X)
  const bool x3 = precision == 2;
  if (x3 && mode != ndim - 1) {
    set_error("precision 2 profiles the last mode only");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  const bool bf16 = precision == TENRSO_PREC_BF16 || x3;
  const void* Xt = nullptr;
  if (bf16 && mode == ndim - 1) {
    const i64 total = prod(shape, ndim);
    void* pt = nullptr;
    TENRSO_TRY(ctx_buffer(ctx, SLOT_XBF16_T, (size_t)total * (x3 ? 4 : 2) + 16, &pt));
    if (x3) TENRSO_TRY(transpose_last_to_front_f32(ctx, (const float*)X, ndim, shape, (float*)pt));
    else TENRSO_TRY(transpose_last_to_front_u16(ctx, (const uint16_t*)X, ndim, shape, (uint16_t*)pt));
    Xt = pt;
  }
  ctx->profile_reps = reps;
  int rc = mttkrp_dense(ctx, bf16 ? TENRSO_PREC_BF16 : TENRSO_PREC_FP32, X, Xt, ndim, shape, factors, rank, mode,
                        MttkrpOut::PARTIAL, nullptr, nullptr, nullptr, x3);
  ctx->profile_reps = 0;
  if (rc != TENRSO_OK) return rc;
  *ms = ctx->last_kernel_ms;
  return TENRSO_OK;
}

int tenrso_mttkrp_all_modes_f32(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape,
                                const float* const* factors, int64_t rank, float* const* outs) {
  for (int m = 0; m < ndim; ++m) TENRSO_TRY(tenrso_mttkrp_f32(ctx, X, ndim, shape, factors, rank, m, outs[m]));
  return TENRSO_OK;
}

int tenrso_mttkrp_f32_host(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape,
                           const float* const* factors, int64_t rank, int mode, float* out) {
  clear_error();
  if (bad_ctx(ctx)) return TENRSO_ERR_INVALID_ARGUMENT;
  TENRSO_TRY(check_mttkrp_basic(ndim, shape, rank, mode));
  if (rank == 0 || shape[mode] == 0) return TENRSO_OK;
  TENRSO_TRY(set_device(ctx));
  const i64 total = prod(shape, ndim);
  i64 ftot = 0;
  for (int k = 0; k < ndim; ++k) ftot += shape[k] * rank;
  void *dx, *df, *dout;
  TENRSO_TRY(ctx_buffer(ctx, SLOT_STAGE_X, (size_t)total * sizeof(float), &dx));
  TENRSO_TRY(ctx_buffer(ctx, SLOT_STAGE_F, (size_t)ftot * sizeof(float), &df));
  TENRSO_TRY(ctx_buffer(ctx, SLOT_STAGE_OUT, (size_t)shape[mode] * rank * sizeof(float), &dout));
  TENRSO_HIP_TRY(hipMemcpyAsync(dx, X, (size_t)total * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  std::vector<const float*> dF(ndim);
  i64 off = 0;
  for (int k = 0; k < ndim; ++k) {
    float* p = (float*)df + off;
    TENRSO_HIP_TRY(hipMemcpyAsync(p, factors[k], (size_t)shape[k] * rank * sizeof(float), hipMemcpyHostToDevice,
                                  ctx->stream));
    dF[k] = p;
    off += shape[k] * rank;
  }
  TENRSO_TRY(mttkrp_dense(ctx, TENRSO_PREC_FP32, dx, nullptr, ndim, shape, dF.data(), rank, mode, MttkrpOut::F32,
                          dout, nullptr));
  TENRSO_HIP_TRY(hipMemcpyAsync(out, dout, (size_t)shape[mode] * rank * sizeof(float), hipMemcpyDeviceToHost,
                                ctx->stream));
  TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return TENRSO_OK;
}

int tenrso_khatri_rao_f32(tenrso_ctx* ctx, const float* A, int64_t I, const float* B, int64_t J, int64_t K,
                          float* C) {
  clear_error();
  if (bad_ctx(ctx)) return TENRSO_ERR_INVALID_ARGUMENT;
  if (I < 0 || J < 0 || K < 0) {
    set_error("Number of columns must match and sizes must be non-negative");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  TENRSO_TRY(set_device(ctx));
  return khatri_rao_f32(ctx, A, I, B, J, K, C);
}

int tenrso_khatri_rao_f32_host(tenrso_ctx* ctx, const float* A, int64_t I, const float* B, int64_t J, int64_t K,
                               float* C) {
  clear_error();
  if (bad_ctx(ctx)) return TENRSO_ERR_INVALID_ARGUMENT;
  if (I < 0 || J < 0 || K < 0) {
    set_error("Number of columns must match and sizes must be non-negative");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  if (I * J * K == 0) return TENRSO_OK;
  TENRSO_TRY(set_device(ctx));
  void *da, *db, *dc;
  TENRSO_TRY(ctx_buffer(ctx, SLOT_STAGE_X, (size_t)(I + J) * K * sizeof(float), &da));
  db = (float*)da + I * K;
  TENRSO_TRY(ctx_buffer(ctx, SLOT_STAGE_OUT, (size_t)I * J * K * sizeof(float), &dc));
  TENRSO_HIP_TRY(hipMemcpyAsync(da, A, (size_t)I * K * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  TENRSO_HIP_TRY(hipMemcpyAsync(db, B, (size_t)J * K * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  TENRSO_TRY(khatri_rao_f32(ctx, (const float*)da, I, (const float*)db, J, K, (float*)dc));
  TENRSO_HIP_TRY(hipMemcpyAsync(C, dc, (size_t)I * J * K * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return TENRSO_OK;
}

int tenrso_cp_reconstruct_f32(tenrso_ctx* ctx, int ndim, const int64_t* shape, const float* const* factors,
                              int64_t rank, const float* weights, float* out) {
  clear_error();
  if (bad_ctx(ctx)) return TENRSO_ERR_INVALID_ARGUMENT;
  if (ndim <= 0) {
    set_error("Need at least one factor matrix for CP reconstruction");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  if (ndim > MAXD_RECON) {
    set_error("tensors with more than %d modes are not supported", MAXD_RECON);
    return TENRSO_ERR_UNSUPPORTED;
  }
  TENRSO_TRY(set_device(ctx));
  return cp_reconstruct_f32(ctx, ndim, shape, factors, rank, weights, out);
}

int tenrso_cp_extract_weights_f32(tenrso_ctx* ctx, int ndim, const int64_t* shape, float* const* factors,
                                  int64_t rank, float* weights) {
  clear_error();
  if (bad_ctx(ctx)) return TENRSO_ERR_INVALID_ARGUMENT;
  if (ndim <= 0 || !shape || !factors || !weights) {
    set_error("Need at least one factor matrix and a weights buffer to extract weights");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  if (ndim > MAXD_RECON) {
    set_error("tensors with more than %d modes are not supported", MAXD_RECON);
    return TENRSO_ERR_UNSUPPORTED;
  }
  if (rank < 0) {
    set_error("negative rank %lld", (long long)rank);
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  for (int k = 0; k < ndim; ++k)
    if (shape[k] < 0 || (shape[k] > 0 && rank > 0 && !factors[k])) {
      set_error("factor %d: negative size or null pointer", k);
      return TENRSO_ERR_INVALID_ARGUMENT;
    }
  TENRSO_TRY(set_device(ctx));
  return cp_extract_weights_f32(ctx, ndim, shape, factors, rank, weights);
}

int tenrso_cp_als_f32(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape, float* const* factors,
                      const tenrso_cp_options* opts, tenrso_cp_result* result, double* fit_history) {
  clear_error();
  if (bad_ctx(ctx)) return TENRSO_ERR_INVALID_ARGUMENT;
  if (!opts || ndim <= 0 || !shape) {
    set_error("null options or empty shape");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  return cp_als_device(ctx, X, ndim, shape, factors, opts, result, fit_history);
}

int tenrso_cp_init_svd(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape, int64_t rank, int kind,
                       uint64_t seed, float* const* factors) {
  clear_error();
  if (bad_ctx(ctx)) return TENRSO_ERR_INVALID_ARGUMENT;
  if (!X || !shape || !factors || ndim <= 0) {
    set_error("null tensor, shape or factors");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  if (rank <= 0) {
    set_error("Invalid rank: %lld", (long long)rank);
    return TENRSO_ERR_INVALID_RANK;
  }
  TENRSO_TRY(set_device(ctx));
  return cp_init_svd(ctx, X, ndim, shape, rank, kind, seed, factors, nullptr);
}

int tenrso_cp_als_f32_host(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape, float* const* factors,
                           const tenrso_cp_options* opts, tenrso_cp_result* result, double* fit_history) {
  clear_error();
  if (bad_ctx(ctx)) return TENRSO_ERR_INVALID_ARGUMENT;
  if (!opts || ndim <= 0 || !shape || ndim > MAXD) {
    set_error("null options, empty shape or too many modes");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  TENRSO_TRY(set_device(ctx));
  const i64 total = prod(shape, ndim);
  const i64 R = opts->rank;
  i64 ftot = 0;
  for (int k = 0; k < ndim; ++k) ftot += shape[k] * (R > 0 ? R : 0);
  void *dx, *df;
  TENRSO_TRY(ctx_buffer(ctx, SLOT_STAGE_X, (size_t)total * sizeof(float), &dx));
  TENRSO_TRY(ctx_buffer(ctx, SLOT_STAGE_F, (size_t)ftot * sizeof(float) + 16, &df));
  TENRSO_HIP_TRY(hipMemcpyAsync(dx, X, (size_t)total * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  std::vector<float*> dF(ndim);
  i64 off = 0;
  for (int k = 0; k < ndim; ++k) {
    dF[k] = (float*)df + off;
    if (R > 0)
      TENRSO_HIP_TRY(hipMemcpyAsync(dF[k], factors[k], (size_t)shape[k] * R * sizeof(float), hipMemcpyHostToDevice,
                                    ctx->stream));
    off += shape[k] * (R > 0 ? R : 0);
  }
  int rc = cp_als_device(ctx, (const float*)dx, ndim, shape, dF.data(), opts, result, fit_history);
  if (rc != TENRSO_OK) return rc;
  for (int k = 0; k < ndim; ++k)
    TENRSO_HIP_TRY(hipMemcpyAsync(factors[k], dF[k], (size_t)shape[k] * R * sizeof(float), hipMemcpyDeviceToHost,
                                  ctx->stream));
  TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return TENRSO_OK;
}

// ---------------------------------------------------------------- sparse
int tenrso_sp_create(tenrso_ctx* ctx, int ndim, const int64_t* shape, int64_t nnz, const int32_t* indices,
                     const float* values, int host_pointers, tenrso_sptensor** out) {
  clear_error();
  if (bad_ctx(ctx) || !out || !shape) {
    if (ctx) set_error("null argument");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  TENRSO_TRY(set_device(ctx));
  return sp_create(ctx, ndim, shape, nnz, indices, values, host_pointers != 0, out);
}

int tenrso_sp_destroy(tenrso_sptensor* t) {
  sp_destroy(t);
  return TENRSO_OK;
}

int tenrso_sp_mode_arrays(tenrso_ctx* ctx, const tenrso_sptensor* t, int mode, int64_t* rowptr, int32_t* row,
                          int32_t* others, float* values) {
  clear_error();
  if (bad_ctx(ctx) || !t) return TENRSO_ERR_INVALID_ARGUMENT;
  if (mode < 0 || mode >= t->nd) {
    set_error("Mode %d out of bounds for tensor with rank %d", mode, t->nd);
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  TENRSO_TRY(set_device(ctx));
  const SpMode& md = t->modes[mode];
  const i64 nnz = t->nnz;
  if (rowptr)
    TENRSO_HIP_TRY(hipMemcpyAsync(rowptr, md.rowptr, (size_t)(md.rows + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
  if (nnz > 0) {
    if (row) TENRSO_HIP_TRY(hipMemcpyAsync(row, md.row, (size_t)nnz * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (others && t->nd > 1)
      TENRSO_HIP_TRY(hipMemcpyAsync(others, md.oth, (size_t)nnz * (t->nd - 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (values) TENRSO_HIP_TRY(hipMemcpyAsync(values, md.vals, (size_t)nnz * 4, hipMemcpyDeviceToHost, ctx->stream));
  }
  TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return TENRSO_OK;
}

int tenrso_sp_csf(tenrso_ctx* ctx, const tenrso_sptensor* t, int mode, int64_t* nfib, int64_t* const* fptr,
                  int32_t* const* fids, float* values) {
  clear_error();
  if (bad_ctx(ctx) || !t || !nfib) return TENRSO_ERR_INVALID_ARGUMENT;
  if (mode < 0 || mode >= t->nd) {
    set_error("Mode %d out of bounds for tensor with rank %d", mode, t->nd);
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  TENRSO_TRY(set_device(ctx));
  std::vector<std::vector<int64_t>> fp;
  std::vector<std::vector<int32_t>> fi;
  std::vector<float> vals;
  TENRSO_TRY(sp_export_csf(ctx, t, mode, fp, fi, vals));
  for (int l = 0; l < t->nd; ++l) nfib[l] = (int64_t)fi[l].size();
  if (fptr && fids) {
    for (int l = 0; l < t->nd; ++l) {
      if (fptr[l]) std::memcpy(fptr[l], fp[l].data(), fp[l].size() * sizeof(int64_t));
      if (fids[l]) std::memcpy(fids[l], fi[l].data(), fi[l].size() * sizeof(int32_t));
    }
  }
  if (values) std::memcpy(values, vals.data(), vals.size() * sizeof(float));
  return TENRSO_OK;
}

int tenrso_sp_hicoo(tenrso_ctx* ctx, const tenrso_sptensor* t, const int64_t* block_shape, int64_t* num_blocks,
                    int64_t* block_coords, int64_t* block_ptrs, int32_t* local, float* values) {
  clear_error();
  if (bad_ctx(ctx) || !t || !block_shape || !num_blocks) return TENRSO_ERR_INVALID_ARGUMENT;
  for (int k = 0; k < t->nd; ++k)
    if (block_shape[k] <= 0) {
      // hicoo.rs:139-143
      set_error("Invalid block shape: block_shape cannot contain zeros");
      return TENRSO_ERR_INVALID_ARGUMENT;
    }
  TENRSO_TRY(set_device(ctx));
  std::vector<int64_t> bc, bp;
  std::vector<int32_t> loc;
  std::vector<float> vals;
  TENRSO_TRY(sp_export_hicoo(ctx, t, block_shape, bc, bp, loc, vals));
  *num_blocks = (int64_t)bp.size() - 1;
  if (block_coords) std::memcpy(block_coords, bc.data(), bc.size() * sizeof(int64_t));
  if (block_ptrs) std::memcpy(block_ptrs, bp.data(), bp.size() * sizeof(int64_t));
  if (local) std::memcpy(local, loc.data(), loc.size() * sizeof(int32_t));
  if (values) std::memcpy(values, vals.data(), vals.size() * sizeof(float));
  return TENRSO_OK;
}

int tenrso_sp_fibres(const tenrso_sptensor* t, int mode, int64_t* nfib) {
  clear_error();
  if (!t || !nfib || mode < 0 || mode >= t->nd) {
    set_error("null tensor or output, or mode out of range");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  *nfib = t->modes[mode].nfib;
  return TENRSO_OK;
}

int tenrso_sp_mttkrp_f32(tenrso_ctx* ctx, const tenrso_sptensor* t, const float* const* factors, int64_t rank,
                         int mode, float* out) {
  clear_error();
  if (bad_ctx(ctx) || !t) return TENRSO_ERR_INVALID_ARGUMENT;
  TENRSO_TRY(check_mttkrp_basic(t->nd, t->shape, rank, mode));
  TENRSO_TRY(set_device(ctx));
  TENRSO_TRY(timed_begin(ctx));
  TENRSO_TRY(sp_mttkrp(ctx, t, factors, rank, mode, out));
  return timed_end(ctx);
}

int tenrso_sp_cp_als_f32(tenrso_ctx* ctx, const tenrso_sptensor* t, float* const* factors,
                         const tenrso_cp_options* opts, tenrso_cp_result* result, double* fit_history) {
  clear_error();
  if (bad_ctx(ctx) || !t || !opts) return TENRSO_ERR_INVALID_ARGUMENT;
  if (opts->precision != TENRSO_PREC_FP32) {
    set_error("sparse CP-ALS runs in FP32");
    return TENRSO_ERR_UNSUPPORTED;
  }
  const i64 R = opts->rank;
  AlsProblem pb;
  pb.nd = t->nd;
  pb.shape = t->shape;
  pb.norm_min = [&](double* tmp, double* scal) { return tensor_norm_min(ctx, t->vals, t->nnz, tmp, scal, SC_COUNT); };
  pb.mttkrp = [&](int mode, const float* const* F, const AlsJob* job, MttkrpResult* mr) -> int {
    if (job) TENRSO_TRY(als_job_launch(ctx, *job));
    i64 maxI = 0;
    for (int k = 0; k < t->nd; ++k) maxI = std::max<i64>(maxI, t->shape[k]);
    void* pm = nullptr;
    TENRSO_TRY(ctx_buffer(ctx, SLOT_SPM, (size_t)maxI * R * sizeof(float), &pm));
    TENRSO_TRY(sp_mttkrp(ctx, t, F, R, mode, (float*)pm));
    mr->partial = (const float*)pm;
    mr->splits = 1;
    return TENRSO_OK;
  };
  return cp_als_run(ctx, pb, factors, opts, result, fit_history);
}

// ---------------------------------------------------------------- reverse mode
int tenrso_mttkrp_vjp_f32(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape,
                          const float* const* factors, int64_t rank, int mode, const float* cotangent,
                          float* const* dfactors, float* dX) {
  clear_error();
  if (bad_ctx(ctx)) return TENRSO_ERR_INVALID_ARGUMENT;
  TENRSO_TRY(check_mttkrp_basic(ndim, shape, rank, mode));
  if (rank == 0) return TENRSO_OK;
  TENRSO_TRY(set_device(ctx));
  return mttkrp_vjp(ctx, X, ndim, shape, factors, rank, mode, cotangent, dfactors, dX);
}

int tenrso_cp_factor_grads_f32(tenrso_ctx* ctx, const float* grad_output, int ndim, const int64_t* shape,
                               const float* const* factors, int64_t rank, const float* weights,
                               float* const* grads) {
  clear_error();
  if (bad_ctx(ctx) || !grads) return TENRSO_ERR_INVALID_ARGUMENT;
  TENRSO_TRY(check_mttkrp_basic(ndim, shape, rank, 0));
  if (rank == 0) return TENRSO_OK;
  TENRSO_TRY(set_device(ctx));
  return cp_factor_grads(ctx, grad_output, ndim, shape, factors, rank, weights, grads);
}

int tenrso_cp_loss_grad_f32(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape,
                            const float* const* factors, int64_t rank, float* const* grads, double* loss) {
  clear_error();
  if (bad_ctx(ctx)) return TENRSO_ERR_INVALID_ARGUMENT;
  TENRSO_TRY(check_mttkrp_basic(ndim, shape, rank, 0));
  if (rank <= 0) {
    set_error("Invalid rank: %lld", (long long)rank);
    return TENRSO_ERR_INVALID_RANK;
  }
  TENRSO_TRY(set_device(ctx));
  return cp_loss_grad(ctx, X, ndim, shape, factors, rank, grads, loss);
}

int tenrso_cp_gd_f32(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape, float* const* factors,
                     int64_t rank, int steps, double lr, double* loss_history) {
  clear_error();
  if (bad_ctx(ctx)) return TENRSO_ERR_INVALID_ARGUMENT;
  TENRSO_TRY(check_mttkrp_basic(ndim, shape, rank, 0));
  if (rank <= 0 || steps < 0) {
    set_error("rank must be positive and steps non-negative");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  TENRSO_TRY(set_device(ctx));
  return cp_gd(ctx, X, ndim, shape, factors, rank, steps, lr, loss_history);
}

}  // extern "C"

