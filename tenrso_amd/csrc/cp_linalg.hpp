// cp_linalg.hpp -- FP64 normal-equation kernels of the ALS step.
#pragma once
#include "common.hpp"

namespace tenrso {

constexpr int SOLVE_RMAX = 64;  // rank bound of the on-chip (LDS) normal-equation solve
// Rank bound of the wide-rank ALS path (ranks SOLVE_RMAX + 1 .. WIDE_NE): the job
// as its own kernel with the matrix in registers, Winv of stride WIDE_NE.
constexpr int WIDE_NE = 128;

// Regularisation applied to the Gram-Hadamard matrix (cp/core.rs:330-345).
struct GramReg {
  double l2 = 0.0;        // ridge on the diagonal
  int tikhonov_order = 0; // > 0: lambda * D^T D (cp/helpers.rs:630-702)
  double tikhonov_lambda = 0.0;
};

// Epilogue applied to the solved rows (cp/core.rs:351-365).
struct RowEpilogue {
  double l1 = 0.0;        // soft threshold
  int nonneg = 0;         // max(0, x)
};

// Device scalars block used by the ALS loop (doubles).
enum Scalars : int {
  SC_NORM_X = 0,   // ||X||^2
  SC_FIT = 1,      // fit of the last evaluation
  SC_IP = 2,       // <X, Xhat>
  SC_RECON = 3,    // ||Xhat||^2
  SC_STATUS = 4,   // solve path of even sweeps: 0 plain, 1 ridge retry, 2 failure
  SC_MIN_X = 5,    // min(X)
  SC_STATUS_ODD = 6,  // solve path of odd sweeps
  SC_COUNT = 8
};

// ||X||^2 (FP64) and min(X) into scal[SC_NORM_X], scal[SC_MIN_X].  tmp: >= 2*1024 doubles.
int tensor_norm_min(tenrso_ctx* ctx, const float* X, i64 n, double* tmp, double* scal, int clear = 0);
// final reduction of (sum, min) partials [nblk sums, nblk minima] into scal; clears scal[0, clear) first
int norm_min_finish(tenrso_ctx* ctx, const double* part, int nblk, double* scal, int clear);
// G = A^T A (FP64) for an fp32 rows x R factor, on `stream` (null: the context stream).
int gram_f64(tenrso_ctx* ctx, const float* A, i64 rows, i64 R, double* G, hipStream_t stream = nullptr);
// Grams of count factors in one launch, G_k at G + k*R*R
int gram_f64_many(tenrso_ctx* ctx, int count, const float* const* A, const i64* rows, i64 R, double* G);
// A[i,:] = epi(M[i,:] Winv), M = sum of the `splits` fp32 partials [splits][rows][R]
// (or the fp64 M when part is null); ip_part[blk] = sum_i M[i,:].A[i,:] (FP64).
int solve_rows(tenrso_ctx* ctx, const float* part, int splits, const double* M, i64 rows, i64 R,
               const double* Winv, const RowEpilogue& epi, float* A, double* ip_part, int* nblk_out);
// A <- Q of the QR factorisation of the rows x R factor A (Householder, FP64,
// one workgroup), normalised to a positive diagonal of R.
int orthonormalize_rows(tenrso_ctx* ctx, float* A, i64 rows, i64 R);
// ip_part[blk] = sum_{i in blk} M[i,:] . A[i,:]
int inner_rows(tenrso_ctx* ctx, const double* M, const float* A, i64 rows, i64 R, double* ip_part, int* nblk_out);
// x = max(x, 0) on an fp32 array.
int clamp_nonneg(tenrso_ctx* ctx, float* A, i64 n);
// Makes the context stream wait for a wide-rank job still pending on the side
// stream (no-op otherwise).
int job_join(tenrso_ctx* ctx);
// fill an fp64 array
int fill_f64(tenrso_ctx* ctx, double* p, i64 n, double v);

}  // namespace tenrso
