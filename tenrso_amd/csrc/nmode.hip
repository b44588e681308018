// nmode.hip -- n-mode products (TTM) and the sequential TTM chain of Tucker-HOOI.
//
// nmode_product (tenrso-kernels/src/nmode.rs:72-130): for X of shape
// (I_0, .., I_k, .., I_{N-1}) and M of shape (J, I_k), Y = X x_k M has shape
// (I_0, .., J, .., I_{N-1}) with Y[.., j, ..] = sum_i M[j, i] X[.., i, ..].  The
// reference unfolds X, multiplies and folds back; here X is read in place as
// [P, I_k, Q] (P = product of the modes before k, Q after) and Y written in place
// as [P, J, Q], so the product is a plain batched GEMM: for each p,
// Y_p (J x Q) = M (J x I_k) . X_p (I_k x Q).  A plain GEMM goes to rocBLAS
// (FP32 SGEMM): in its column-major terms Y_p^T (Q x J) = X_p^T (Q x I_k) . M^T,
// batched over p with strides I_k Q and J Q; for Q = 1 the P products fuse into
// one GEMM, Y^T (J x P) = M (J x I_k) . X^T (I_k x P).
// nmode_products_seq (nmode.rs:241-255) applies the products in order; the
// intermediate tensors live in the context's buffers.
#include <rocblas/rocblas.h>

#include <vector>

#include "mttkrp.hpp"

namespace tenrso {

namespace {

__global__ __launch_bounds__(256) void zero_fill_kernel(float* __restrict__ p, i64 n) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e < n) p[e] = 0.0f;
}

int blas_handle(tenrso_ctx* ctx, rocblas_handle* h) {
  if (!ctx->blas) {
    rocblas_handle nh = nullptr;
    if (rocblas_create_handle(&nh) != rocblas_status_success) {
      set_error("rocblas_create_handle failed");
      return TENRSO_ERR_DEVICE;
    }
    ctx->blas = nh;
  }
  *h = (rocblas_handle)ctx->blas;
  if (rocblas_set_stream(*h, ctx->stream) != rocblas_status_success) {
    set_error("rocblas_set_stream failed");
    return TENRSO_ERR_DEVICE;
  }
  return TENRSO_OK;
}

}  // namespace

void blas_destroy(tenrso_ctx* ctx) {
  if (ctx->blas) (void)rocblas_destroy_handle((rocblas_handle)ctx->blas);
  ctx->blas = nullptr;
}

int nmode_product_f32(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, const float* M, i64 J, int mode,
                      float* Y) {
  i64 P = 1, Q = 1;
  for (int k = 0; k < mode; ++k) P *= shape[k];
  for (int k = mode + 1; k < nd; ++k) Q *= shape[k];
  const i64 I = shape[mode];
  if (P * Q * J == 0) return TENRSO_OK;
  rocblas_handle h;
  TENRSO_TRY(blas_handle(ctx, &h));
  const float one = 1.0f, zero = 0.0f;
  rocblas_status st;
  if (I == 0) {
    // an empty contraction: Y = 0 (a fill kernel, not a memset: callers may
    // capture the call in a graph)
    const i64 n = P * J * Q;
    hipLaunchKernelGGL(zero_fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, Y, n);
    TENRSO_HIP_TRY(hipGetLastError());
    return TENRSO_OK;
  }
  if (Q == 1) {
    // Y^T (J x P, column-major, ld J) = M (J x I: M^T of the column-major I x J view) . X^T (I x P, ld I)
    st = rocblas_sgemm_64(h, rocblas_operation_transpose, rocblas_operation_none, J, P, I, &one, M, I, X, I, &zero, Y,
                          J);
  } else {
    // per p: Y_p^T (Q x J, ld Q) = X_p^T (Q x I, ld Q) . M^T (I x J: the column-major view of M, ld I)
    st = rocblas_sgemm_strided_batched_64(h, rocblas_operation_none, rocblas_operation_none, Q, J, I, &one, X, Q,
                                          I * Q, M, I, 0, &zero, Y, Q, J * Q, P);
  }
  if (st != rocblas_status_success) {
    set_error("rocBLAS SGEMM failed (status %d)", (int)st);
    return TENRSO_ERR_DEVICE;
  }
  return TENRSO_OK;
}

// Y = X x_{modes[0]} M_0 x_{modes[1]} M_1 ..., products applied in order;
// out_shape receives the final shape.  Intermediates ping-pong between two
// context buffers; the last product writes Y.
int nmode_products_seq_f32(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, int nmat,
                           const float* const* Ms, const i64* rows, const int* modes, float* Y, i64* out_shape) {
  std::vector<i64> cur(shape, shape + nd);
  const float* src = X;
  for (int t = 0; t < nmat; ++t) {
    std::vector<i64> next = cur;
    next[modes[t]] = rows[t];
    float* dst = Y;
    if (t + 1 < nmat) {
      i64 n = 1;
      for (i64 v : next) n *= v;
      void* buf = nullptr;
      TENRSO_TRY(ctx_buffer(ctx, (t & 1) ? SLOT_TMP2 : SLOT_TMP, (size_t)std::max<i64>(1, n) * sizeof(float), &buf));
      dst = (float*)buf;
    }
    TENRSO_TRY(nmode_product_f32(ctx, src, nd, cur.data(), Ms[t], rows[t], modes[t], dst));
    src = dst;
    cur = next;
  }
  if (nmat == 0) {
    i64 n = 1;
    for (i64 v : cur) n *= v;
    if (n > 0 && Y != X)
      TENRSO_HIP_TRY(hipMemcpyAsync(Y, X, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
  }
  for (int k = 0; k < nd; ++k) out_shape[k] = cur[k];
  return TENRSO_OK;
}

}  // namespace tenrso

using namespace tenrso;

namespace {

// nmode.rs:82-98: the reference's checks, order and messages
int check_nmode(int nd, const int64_t* shape, int64_t J, int64_t cols, int mode) {
  if (nd <= 0 || !shape) {
    set_error("tensor must have at least one mode");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  if (mode < 0 || mode >= nd) {
    set_error("Mode %d out of bounds for tensor with rank %d", mode, nd);
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  if (cols != shape[mode]) {
    set_error("Matrix columns (%lld) must match tensor mode-%d size (%lld)", (long long)cols, mode,
              (long long)shape[mode]);
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  if (J < 0) {
    set_error("negative matrix row count %lld", (long long)J);
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  if (nd > MAXD) {
    set_error("tensors with more than %d modes are not supported", MAXD);
    return TENRSO_ERR_UNSUPPORTED;
  }
  return TENRSO_OK;
}

}  // namespace

extern "C" int tenrso_nmode_product_f32(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape,
                                        const float* matrix, int64_t rows, int64_t cols, int mode, float* Y) {
  clear_error();
  if (!ctx) {
    set_error("null context");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  TENRSO_TRY(check_nmode(ndim, shape, rows, cols, mode));
  TENRSO_TRY(set_device(ctx));
  return nmode_product_f32(ctx, X, ndim, shape, matrix, rows, mode, Y);
}

extern "C" int tenrso_nmode_products_seq_f32(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape,
                                             int nmat, const float* const* matrices, const int64_t* rows,
                                             const int64_t* cols, const int* modes, float* Y, int64_t* out_shape) {
  clear_error();
  if (!ctx || !out_shape || nmat < 0 || (nmat > 0 && (!matrices || !rows || !cols || !modes))) {
    set_error("null argument");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  // each product is checked against the shape it meets (the previous products' result)
  std::vector<int64_t> cur(shape, shape + ndim);
  for (int t = 0; t < nmat; ++t) {
    TENRSO_TRY(check_nmode(ndim, cur.data(), rows[t], cols[t], modes[t]));
    cur[modes[t]] = rows[t];
  }
  TENRSO_TRY(set_device(ctx));
  return nmode_products_seq_f32(ctx, X, ndim, shape, nmat, matrices, rows, modes, Y, out_shape);
}
