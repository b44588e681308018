/*
 * tenrso_hip.h -- C ABI of the MI355X (gfx950) CP decomposition hot path.
 *
 * Drop-in boundary for tenrso's CP path.  Every entry point below names the
 * reference function it replaces (paths relative to the tenrso repository).
 *
 * Conventions
 *   - Tensors are dense, C-contiguous (row-major, last mode fastest), like
 *     DenseND<T> (tenrso-core/src/dense/types.rs:40-43).  Factor matrices are
 *     row-major I_k x R, like ndarray Array2 (standard layout).
 *   - Functions without the `_host` suffix take DEVICE pointers (hipMalloc'd or
 *     torch.cuda tensors); `shape` and the `factors` pointer arrays themselves are
 *     host arrays.  `_host` variants take host pointers and stage through the
 *     context's device buffers (what a Rust `extern "C"` shim over ArrayView calls).
 *   - Every call returns a status (TENRSO_OK == 0).  On failure the message that
 *     the reference would put in its anyhow/CpError is available from
 *     tenrso_last_error() (thread-local).  Nothing panics or throws across the ABI.
 *   - Work is enqueued on the context's stream.  Device-pointer calls are
 *     asynchronous unless stated; `_host` calls and cp_als synchronise.
 *   - A context serialises its calls; distinct contexts may be used from distinct
 *     threads.
 */
#ifndef TENRSO_HIP_H
#define TENRSO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: CpError (tenrso-decomp/src/cp/types.rs:13-35) and
 *      KernelError / anyhow messages (tenrso-kernels/src/mttkrp.rs:85-122) ---- */
enum tenrso_status {
  TENRSO_OK = 0,
  TENRSO_ERR_INVALID_ARGUMENT = 1,       /* mttkrp/khatri_rao validation (anyhow::bail!) */
  TENRSO_ERR_INVALID_RANK = 2,           /* CpError::InvalidRank */
  TENRSO_ERR_INVALID_TOLERANCE = 3,      /* CpError::InvalidTolerance */
  TENRSO_ERR_SHAPE_MISMATCH = 4,         /* CpError::ShapeMismatch */
  TENRSO_ERR_LINALG = 5,                 /* CpError::LinalgError */
  TENRSO_ERR_NONNEGATIVE = 6,            /* CpError::NonnegativeViolation */
  TENRSO_ERR_INVALID_REGULARIZATION = 7, /* CpError::InvalidRegularization */
  TENRSO_ERR_DEVICE = 8,                 /* HIP runtime failure */
  TENRSO_ERR_UNSUPPORTED = 9
};

/* precision of the MTTKRP inside cp_als */
enum tenrso_precision {
  TENRSO_PREC_FP32 = 0, /* plain FP32 MFMA MTTKRP: the drop-in */
  TENRSO_PREC_BF16 = 1  /* BF16-input / FP32-accumulate MFMA MTTKRP */
};

/* ConvergenceReason (tenrso-decomp/src/cp/types.rs:312-322) */
enum tenrso_reason {
  TENRSO_REASON_FIT_TOLERANCE = 0,
  TENRSO_REASON_MAX_ITERATIONS = 1,
  TENRSO_REASON_OSCILLATION = 2,
  TENRSO_REASON_TIME_LIMIT = 3
};

typedef struct tenrso_ctx tenrso_ctx;

/* Thread-local message of the last failing call on this thread ("" if none). */
const char* tenrso_last_error(void);
/* Library version string. */
const char* tenrso_version(void);

/* Context: owns a HIP stream, or borrows `stream` when borrow_stream != 0
 * (NULL then means the legacy default stream, e.g. torch's default stream), plus
 * device scratch and timing events.  No reference counterpart: the reference is
 * in-process CPU code. */
int tenrso_ctx_create(int device, int borrow_stream, void* stream, tenrso_ctx** out);
int tenrso_ctx_destroy(tenrso_ctx* ctx);
int tenrso_ctx_synchronize(tenrso_ctx* ctx);
/* When timing is enabled, MTTKRP entry points record HIP events around their
 * kernels and synchronise; tenrso_ctx_last_kernel_ms returns the last duration. */
int tenrso_ctx_set_timing(tenrso_ctx* ctx, int enabled);
int tenrso_ctx_last_kernel_ms(tenrso_ctx* ctx, double* ms);

/* ------------------------------------------------------------------------
 * Dense MTTKRP.  Replaces tenrso_kernels::mttkrp (tenrso-kernels/src/mttkrp.rs:74-146)
 * and, with the same contract, mttkrp_blocked / _blocked_parallel / _fused /
 * _fused_parallel (mttkrp.rs:395-498, 528-609, 783-878, 903-992).
 *   out[i, r] = sum_{i_k, k != mode} X[..., i, ...] * prod_{k != mode} F_k[i_k, r]
 * X: prod(shape) floats; factors[k]: shape[k] x rank; out: shape[mode] x rank.
 * Khatri-Rao rows are formed on the fly; no unfolding copy is made.
 * ------------------------------------------------------------------------ */
int tenrso_mttkrp_f32(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape,
                      const float* const* factors, int64_t rank, int mode, float* out);
/* Argument check with the reference's messages and order (mttkrp.rs:85-122,
 * 194-196), for callers whose factors may disagree in shape: factor_rows[k] x
 * factor_cols[k] is the shape of factor k.  No device work. */
int tenrso_mttkrp_check(int ndim, const int64_t* shape, int nfactors, const int64_t* factor_rows,
                        const int64_t* factor_cols, int mode);
int tenrso_mttkrp_f32_host(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape,
                           const float* const* factors, int64_t rank, int mode, float* out);
/* Same contract with FP64 accumulation of the split partial sums into a double
 * output (used by cp_als; exposed for precision studies). */
int tenrso_mttkrp_f32_f64out(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape,
                             const float* const* factors, int64_t rank, int mode, double* out);

/* BF16-input / FP32-accumulate MFMA MTTKRP.  X_bf16 is the tensor rounded to
 * bfloat16 (RNE, tenrso_convert_f32_to_bf16); factors stay FP32 and the
 * Khatri-Rao rows are formed in FP32 and rounded once to bf16.  No reference
 * counterpart (the reference has no reduced-precision MTTKRP). */
int tenrso_mttkrp_bf16(tenrso_ctx* ctx, const uint16_t* X_bf16, int ndim, const int64_t* shape,
                       const float* const* factors, int64_t rank, int mode, float* out);
int tenrso_convert_f32_to_bf16(tenrso_ctx* ctx, const float* src, int64_t n, uint16_t* dst);

/* Profiling aid (no reference counterpart): launches only the main MFMA (or
 * fallback) MTTKRP kernel `reps` times back to back between two HIP events on
 * the context's stream and returns the average device time per launch in *ms.
 * precision selects the fp32 (X: float*) or bf16 (X: uint16_t*) kernel, or, with
 * 2, the three-way bf16 split kernel that the BF16 cp_als runs for its last mode
 * (X: float*, mode = ndim - 1). */
int tenrso_mttkrp_profile(tenrso_ctx* ctx, int precision, const void* X, int ndim, const int64_t* shape,
                          const float* const* factors, int64_t rank, int mode, int reps, double* ms);

/* No-allocation MTTKRP: the contract of tenrso_mttkrp_f32, but its device
 * workspace (split partial sums; Khatri-Rao rows of the slow modes for tensors
 * of order >= 4) lives in a file-scope device array of 16 MiB per GPU, held
 * until the process ends, instead of the context's buffers: the call makes no
 * hipMalloc, hipFree or other allocating call.  No counterpart in the
 * reference; it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops).  The context must have been created
 * beforehand (tenrso_ctx_create sets the workspace up on its device).  A shape
 * whose workspace exceeds 16 MiB even with the fewest split partials is
 * TENRSO_ERR_UNSUPPORTED.  All calls on one GPU share the array: calls that
 * overlap on one GPU race, so callers serialise them. */
int tenrso_mttkrp_f32_noalloc(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape,
                              const float* const* factors, int64_t rank, int mode, float* out);

/* Out-of-core MTTKRP: the streaming MTTKRP that tenrso-ooc's example describes
 * (tenrso-ooc/examples/mttkrp_streaming.rs:66-110: chunk the tensor, a partial
 * MTTKRP per chunk, accumulate; the example itself calls the in-memory mttkrp).
 * X is a HOST tensor (pageable, or pinned with hipHostMalloc / torch pin_memory:
 * then copied without staging), factors and out are host arrays.  The tensor
 * streams in mode-0 slabs through two device buffers sized so that this call's
 * device memory stays within device_budget_bytes (including the split partial
 * sums of the per-slab MFMA MTTKRP), copies overlapped with the previous slab's
 * kernels; mode-n > 0 partials are summed in FP64 in slab order.  stats (may be
 * NULL, 8 doubles): seconds, tensor bytes moved host->device per second (GB/s),
 * slab rows, slabs, input pinned (1/0), host staging seconds.  A budget below
 * two one-row slabs plus the factors and partials is TENRSO_ERR_INVALID_ARGUMENT. */
int tenrso_mttkrp_f32_streaming(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape,
                                const float* const* factors, int64_t rank, int mode, float* out,
                                int64_t device_budget_bytes, double* stats);

/* mttkrp_all_modes (tenrso-kernels/src/utils.rs:282-311): outs[k] for every mode. */
int tenrso_mttkrp_all_modes_f32(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape,
                                const float* const* factors, int64_t rank, float* const* outs);

/* ------------------------------------------------------------------------
 * n-mode products (TTM).  Replaces tenrso_kernels::nmode_product
 * (tenrso-kernels/src/nmode.rs:72-130): for X of shape (I_0..I_{N-1}) and a
 * row-major matrix M of rows x cols with cols == I_mode, Y = X x_mode M has
 * shape (I_0, .., rows, .., I_{N-1}), Y[.., j, ..] = sum_i M[j, i] X[.., i, ..].
 * All pointers are device pointers; the product is a batched FP32 GEMM
 * (rocBLAS) on the context's stream.  The reference's checks and messages:
 * "Mode %d out of bounds for tensor with rank %d", "Matrix columns (%d) must
 * match tensor mode-%d size (%d)" (TENRSO_ERR_INVALID_ARGUMENT).
 * ------------------------------------------------------------------------ */
int tenrso_nmode_product_f32(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape,
                             const float* matrix, int64_t rows, int64_t cols, int mode, float* Y);
/* nmode_products_seq (nmode.rs:241-255) and tucker_operator_ordered
 * (nmode.rs:517-530): the products (matrices[t], modes[t]) applied in order,
 * each checked against the shape the previous ones left; intermediates live in
 * the context's buffers and the last product writes Y, whose shape is returned
 * in out_shape (ndim entries).  Tucker-HOOI's TTM chain
 * (tenrso-decomp/src/tucker.rs:1122-1140) is this call with the transposed
 * factors of every mode but one. */
int tenrso_nmode_products_seq_f32(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape, int nmat,
                                  const float* const* matrices, const int64_t* rows, const int64_t* cols,
                                  const int* modes, float* Y, int64_t* out_shape);

/* ------------------------------------------------------------------------
 * Khatri-Rao product.  Replaces tenrso_kernels::khatri_rao (khatri_rao.rs:56-89)
 * and its _parallel / _blocked / _blocked_parallel variants (117-357), which are
 * bit-identical: C[ia*J + ib, k] = A[ia, k] * B[ib, k].  The reference panics on
 * a column mismatch; here the caller passes one K, and K/I/J <= 0 is
 * TENRSO_ERR_INVALID_ARGUMENT.
 * ------------------------------------------------------------------------ */
int tenrso_khatri_rao_f32(tenrso_ctx* ctx, const float* A, int64_t I, const float* B, int64_t J, int64_t K,
                          float* C);
int tenrso_khatri_rao_f32_host(tenrso_ctx* ctx, const float* A, int64_t I, const float* B, int64_t J,
                               int64_t K, float* C);

/* ------------------------------------------------------------------------
 * CP reconstruction.  Replaces tenrso_kernels::cp_reconstruct (outer.rs:223-274)
 * used by CpDecomp::reconstruct (cp/types.rs:384-415):
 *   out[i_0..i_{N-1}] = sum_r w_r prod_k F_k[i_k, r]   (weights may be NULL = 1).
 * ------------------------------------------------------------------------ */
int tenrso_cp_reconstruct_f32(tenrso_ctx* ctx, int ndim, const int64_t* shape, const float* const* factors,
                              int64_t rank, const float* weights, float* out);
/* CpDecomp::extract_weights (cp/types.rs:421-446), in place on device factors:
 * every column r of every factor with norm n > FLT_EPSILON is divided by n, and
 * weights[r] (device, rank floats) = product of those norms (1 where none is).
 * Norms are summed in FP64 and rounded to fp32 for the division. */
int tenrso_cp_extract_weights_f32(tenrso_ctx* ctx, int ndim, const int64_t* shape, float* const* factors,
                                  int64_t rank, float* weights);

/* ------------------------------------------------------------------------
 * CP-ALS.  Replaces tenrso_decomp::cp_als (cp/core.rs:59-182) and
 * cp_als_constrained (cp/core.rs:225-408) for T = f32, with the initial
 * factors supplied by the caller (the reference draws them from an unseeded
 * thread_rng, cp/helpers.rs:212-231; a shim draws them and passes them here).
 * Per iteration: for each mode, MTTKRP -> Gram-Hadamard (FP64) -> inverse of
 * Gamma (FP64 block Gauss-Jordan, positive-definiteness tested by Sylvester's
 * criterion) -> row solve -> constraints; then the fit (FP64).  A Gamma that is
 * not positive definite is retried once with Gamma += FLT_EPSILON * 10R * I
 * (the ridge retry of cp/helpers.rs:477-490); if that fails too the call returns
 * TENRSO_ERR_LINALG (CpError::LinalgError) and the factor buffers hold the values
 * of an unfinished sweep (undefined).  Convergence, oscillation and time-limit
 * rules are the reference's (core.rs:117-165).
 * ------------------------------------------------------------------------ */
typedef struct tenrso_cp_options {
  int64_t rank;
  int32_t max_iters;
  int32_t precision;             /* enum tenrso_precision */
  double tol;                    /* in [0, 1) */
  double time_limit_s;           /* < 0: None */
  /* CpConstraints (cp/types.rs:198-309), already reduced to effective values */
  int32_t nonnegative;
  int32_t validate_nonneg_input; /* default 1 */
  int32_t orthogonal;
  int32_t tikhonov_order;        /* > 0: apply_tikhonov_to_gram (cp/helpers.rs:630-702) */
  double l2;                     /* effective_l2() */
  double l1;                     /* effective_l1() */
  double tikhonov_lambda;
  int32_t fit_from_mode0;        /* 1: separate mode-0 MTTKRP for <X, Xhat> as in
                                    cp/helpers.rs:558-579; 0 (default): reuse the
                                    last mode's MTTKRP (same value in exact arithmetic) */
  int32_t reserved;
} tenrso_cp_options;

typedef struct tenrso_cp_result {
  int32_t iters;
  int32_t reason;                /* enum tenrso_reason */
  int32_t oscillation_count;
  int32_t oscillated;
  double fit;
  double final_fit_change;
  int32_t solve_path_max;        /* 0 direct solve, 1 ridge retry was needed */
  int32_t reserved;
  double seconds;                /* wall time of the ALS loop */
} tenrso_cp_result;

/* X: device tensor; factors: device pointers, in = initial factors, out = result.
 * fit_history: host buffer of max_iters doubles, or NULL.  Rank <= 128 (ranks
 * 65..128 take the wide-rank inverse); a larger rank returns
 * TENRSO_ERR_UNSUPPORTED before any launch. */
int tenrso_cp_als_f32(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape, float* const* factors,
                      const tenrso_cp_options* opts, tenrso_cp_result* result, double* fit_history);
int tenrso_cp_als_f32_host(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape,
                           float* const* factors, const tenrso_cp_options* opts, tenrso_cp_result* result,
                           double* fit_history);

/* ------------------------------------------------------------------------
 * SVD-based initial factors: InitStrategy::Svd / Nnsvd / LeverageScore of
 * initialize_factors (tenrso-decomp cp/helpers.rs:232-373), from the SVD of every
 * mode-k unfolding, on the device: the left singular vectors and singular values
 * are the eigenpairs of X_(k) X_(k)^T (FP64 Gram, one-workgroup Jacobi), each
 * vector's largest-magnitude entry made positive.  Mode sizes <= 1024, rank 1..64.
 * X: device tensor; factors: device buffers (shape[k] x rank, fp32).  seed drives
 * the N(0, 0.01) draws of the columns past min(rank, I_k, product of the other
 * modes), which the reference takes from an unseeded thread_rng.
 * ------------------------------------------------------------------------ */
enum tenrso_init_kind { TENRSO_INIT_SVD = 2, TENRSO_INIT_NNSVD = 3, TENRSO_INIT_LEVERAGE = 4 };
int tenrso_cp_init_svd(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape, int64_t rank, int kind,
                       uint64_t seed, float* const* factors);

/* ------------------------------------------------------------------------
 * Reverse mode.  Device pointers throughout; rank <= 64 for the FP64 Gram paths.
 * ------------------------------------------------------------------------ */
/* VJP of mttkrp (no reference counterpart; tenrso-ad registers custom rules through
 * VjpOp, tenrso-ad/src/vjp.rs:30-44).  For the cotangent Mbar (shape[mode] x rank):
 *   dfactors[m] = mttkrp(X, factors with factors[mode] := Mbar, m) for m != mode
 *   dX          = cp_reconstruct(factors with factors[mode] := Mbar)
 * dfactors[m] (or the whole array) and dX may be NULL to skip them; dfactors[mode]
 * is not written (the mode's own factor is not an input of the MTTKRP). */
int tenrso_mttkrp_vjp_f32(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape,
                          const float* const* factors, int64_t rank, int mode, const float* cotangent,
                          float* const* dfactors, float* dX);
/* CpReconstructionGrad::compute_factor_gradients (tenrso-ad/src/grad.rs:81-120):
 * grads[n] = mttkrp(grad_output, factors, n) * diag(weights) (weights may be NULL;
 * a NULL grads[n] skips mode n). */
int tenrso_cp_factor_grads_f32(tenrso_ctx* ctx, const float* grad_output, int ndim, const int64_t* shape,
                               const float* const* factors, int64_t rank, const float* weights,
                               float* const* grads);
/* Loss 0.5 ||X - [[A_1..A_N]]||^2 and grads[n] = A_n Gamma_n - mttkrp(X, A, n),
 * Gamma_n = hadamard_{k != n} A_k^T A_k (Gram, Gamma, product and loss in FP64).
 * grads may be NULL (loss only). */
int tenrso_cp_loss_grad_f32(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape,
                            const float* const* factors, int64_t rank, float* const* grads, double* loss);
/* AdOperation-shaped rules (tenrso-ad/src/hooks.rs:44-74: forward, backward, name,
 * num_inputs, num_outputs; a VjpOp, tenrso-ad/src/vjp.rs:30-44, is backward with
 * one output gradient), over BORROWED device inputs that must outlive the op:
 *   TENRSO_AD_MTTKRP          inputs [X, A_0..A_{N-1}], output [M] (shape[mode] x rank);
 *                             backward: dX, dA_m (m != mode) as tenrso_mttkrp_vjp_f32,
 *                             dA_mode = 0
 *   TENRSO_AD_CP_RECONSTRUCT  inputs [A_0..A_{N-1}] (X unused), weights optional,
 *                             output [Xhat] (prod(shape)); backward: dA_n =
 *                             mttkrp(G, A, n) diag(w) (grad.rs:81-120)
 *   TENRSO_AD_CP_LOSS         inputs [X, A_0..A_{N-1}], output [L] (1 float,
 *                             0.5 ||X - Xhat||^2, FP64 inside); backward with the
 *                             device cotangent g: dX = g (X - Xhat), dA_n =
 *                             g (A_n Gamma_n - mttkrp(X, A, n)); rank 1..64
 * forward/backward take arrays of device buffers (outputs[k] of
 * tenrso_ad_op_output_numel(op, k) floats; input_grads[k] of
 * tenrso_ad_op_input_numel(op, k) floats).  A NULL input_grads array, or a NULL
 * entry, skips that gradient.  Calls are asynchronous on the op's context stream. */
enum tenrso_ad_kind { TENRSO_AD_MTTKRP = 0, TENRSO_AD_CP_RECONSTRUCT = 1, TENRSO_AD_CP_LOSS = 2 };
typedef struct tenrso_ad_op tenrso_ad_op;
int tenrso_ad_op_create(tenrso_ctx* ctx, int kind, const float* X, int ndim, const int64_t* shape,
                        const float* const* factors, int64_t rank, int mode, const float* weights,
                        tenrso_ad_op** out);
int tenrso_ad_op_destroy(tenrso_ad_op* op);
const char* tenrso_ad_op_name(const tenrso_ad_op* op);
int tenrso_ad_op_num_inputs(const tenrso_ad_op* op);
int tenrso_ad_op_num_outputs(const tenrso_ad_op* op);
int64_t tenrso_ad_op_input_numel(const tenrso_ad_op* op, int k);
int64_t tenrso_ad_op_output_numel(const tenrso_ad_op* op, int k);
int tenrso_ad_op_forward(tenrso_ad_op* op, float* const* outputs);
int tenrso_ad_op_backward(tenrso_ad_op* op, const float* const* output_grads, float* const* input_grads);

/* Gradient-descent CP fit: `steps` simultaneous updates A_n -= lr * grad_n (lr <= 0:
 * lr = 1 / max_n ||Gamma_n||_F at each step).  loss_history[s] (host, may be NULL)
 * is the loss at the point step s starts from. */
int tenrso_cp_gd_f32(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape, float* const* factors,
                     int64_t rank, int steps, double lr, double* loss_history);

/* ------------------------------------------------------------------------
 * Sparse tensors.  A COO tensor (tenrso-sparse/src/coo.rs:73-348) is uploaded
 * once and sorted on the device, per output mode, in the lexicographic order of
 * CsfTensor::from_coo with mode_order [mode, others ascending] (csf.rs:137-264).
 * Indices are int32 SoA: indices[k * nnz + e] is the mode-k index of nonzero e.
 * Duplicate coordinates are summed (the dense form of CooTensor::deduplicate,
 * coo.rs:320-348).  Validation follows CooTensor::new (coo.rs:101-149).
 * ------------------------------------------------------------------------ */
typedef struct tenrso_sptensor tenrso_sptensor;

int tenrso_sp_create(tenrso_ctx* ctx, int ndim, const int64_t* shape, int64_t nnz, const int32_t* indices,
                     const float* values, int host_pointers, tenrso_sptensor** out);
int tenrso_sp_destroy(tenrso_sptensor* t);
/* Sorted per-mode arrays, copied to host buffers: rowptr (shape[mode]+1),
 * row (nnz), others ((ndim-1) x nnz, modes ascending), values (nnz).  Any
 * pointer may be NULL to skip it. */
int tenrso_sp_mode_arrays(tenrso_ctx* ctx, const tenrso_sptensor* t, int mode, int64_t* rowptr, int32_t* row,
                          int32_t* others, float* values);
/* CSF of mode order [mode, others ascending] with standard fiber pointers
 * (fptr[l][f]..fptr[l][f+1] index level l+1; the leaf level has one entry per
 * nonzero), unlike the nnz-offset pointers of csf.rs:209,246.  Call with
 * fptr/fids NULL to get nfib[l] (fibers per level, ndim entries). */
int tenrso_sp_csf(tenrso_ctx* ctx, const tenrso_sptensor* t, int mode, int64_t* nfib, int64_t* const* fptr,
                  int32_t* const* fids, float* values);
/* HiCOO of the COO as uploaded (HiCooTensor::from_coo, hicoo.rs:121-205): block
 * coordinate = index / block_shape, local coordinate = index % block_shape,
 * nonzeros stably sorted by block coordinate (lexicographic), duplicates kept.
 * Call with the output pointers NULL to get *num_blocks (B).  block_coords is
 * ndim x B (SoA: block_coords[k * B + b]), block_ptrs B + 1 (block b holds
 * nonzeros block_ptrs[b]..block_ptrs[b+1]), local ndim x nnz (SoA), values nnz.
 * A zero in block_shape is TENRSO_ERR_INVALID_ARGUMENT (hicoo.rs:139-143). */
int tenrso_sp_hicoo(tenrso_ctx* ctx, const tenrso_sptensor* t, const int64_t* block_shape, int64_t* num_blocks,
                    int64_t* block_coords, int64_t* block_ptrs, int32_t* local, float* values);
/* Number of compressed fibres of a mode: the nodes one level above the leaves of
 * the CSF of mode order [mode, others ascending] (csf.rs:137-264), i.e. the
 * distinct prefixes (i_mode, i_o1, .., i_o(N-2)) of the nonzeros.  The sparse
 * MTTKRP of 2- to 4-way tensors reads these fibres (start and prefix, 4 (N)
 * bytes each) and the leaves (index and value, 8 bytes per nonzero). */
int tenrso_sp_fibres(const tenrso_sptensor* t, int mode, int64_t* nfib);
/* Sparse MTTKRP (no reference counterpart; contract of the dense mttkrp on the
 * tensor the COO represents).  factors / out are device pointers. */
int tenrso_sp_mttkrp_f32(tenrso_ctx* ctx, const tenrso_sptensor* t, const float* const* factors, int64_t rank,
                         int mode, float* out);
/* CP-ALS on a sparse tensor: the cp_als rules above with the sparse MTTKRP. */
int tenrso_sp_cp_als_f32(tenrso_ctx* ctx, const tenrso_sptensor* t, float* const* factors,
                         const tenrso_cp_options* opts, tenrso_cp_result* result, double* fit_history);

#ifdef __cplusplus
}
#endif

#endif /* TENRSO_HIP_H */
