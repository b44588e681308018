// ad_ops.hip -- AdOperation-shaped reverse-mode rules at the C ABI.
//
// tenrso-ad lets a framework register custom rules through two traits:
//   VjpOp::vjp(output_grad) -> Vec<grad>                  (tenrso-ad/src/vjp.rs:30-44)
//   AdOperation::{forward, backward, name, num_inputs, num_outputs}
//                                                         (tenrso-ad/src/hooks.rs:44-74)
// A tenrso_ad_op is one such rule over borrowed device inputs, so a Rust
// `impl AdOperation<f32>` (or `impl VjpOp<f32>`, which is backward with one
// output gradient) is a thin forwarding wrapper.  Three rules:
//   "mttkrp"          inputs [X, A_0 .. A_{N-1}], output [M = mttkrp(X, A, mode)]
//                     backward(Mbar): dX = cp_reconstruct(A with A_mode := Mbar),
//                     dA_m = mttkrp(X, A with A_mode := Mbar, m) for m != mode,
//                     dA_mode = 0 (A_mode does not enter M)
//   "cp_reconstruct"  inputs [A_0 .. A_{N-1}], output [Xhat = sum_r w_r o_k a_k^r]
//                     backward(G): dA_n = mttkrp(G, A, n) diag(w)
//                     (CpReconstructionGrad, tenrso-ad/src/grad.rs:81-120)
//   "cp_loss"         inputs [X, A_0 .. A_{N-1}], output [L = 0.5 ||X - [[A]]||^2]
//                     (one element), backward(g): dX = g (X - Xhat),
//                     dA_n = g (A_n Gamma_n - mttkrp(X, A, n))
// Gradients are written only where the caller passes a buffer (a null array
// or null entry skips that input).  The cotangent of cp_loss is read on the
// device, so no call synchronises with the host.
#include "ad.hpp"
#include "kr_reconstruct.hpp"
#include "mttkrp.hpp"

struct tenrso_ad_op {
  tenrso_ctx* ctx = nullptr;
  int kind = 0;
  int nd = 0;
  int mode = 0;
  tenrso::i64 R = 0;
  tenrso::i64 shape[tenrso::MAXD] = {};
  const float* X = nullptr;
  const float* F[tenrso::MAXD] = {};
  const float* w = nullptr;
};

namespace tenrso {
namespace {

const char* const OP_NAMES[] = {"mttkrp", "cp_reconstruct", "cp_loss"};

inline unsigned cdiv(i64 a, i64 b) { return (unsigned)((a + b - 1) / b); }

__global__ __launch_bounds__(256) void zero_kernel(float* __restrict__ p, i64 n) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e < n) p[e] = 0.0f;
}

// p *= g[0]
__global__ __launch_bounds__(256) void scale_by_kernel(float* __restrict__ p, i64 n, const float* __restrict__ g) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e < n) p[e] = p[e] * g[0];
}

__global__ void loss_to_f32_kernel(const double* __restrict__ loss, float* __restrict__ out) {
  if (threadIdx.x == 0) out[0] = (float)loss[0];
}

i64 numel(const tenrso_ad_op* op) {
  i64 t = 1;
  for (int k = 0; k < op->nd; ++k) t *= op->shape[k];
  return t;
}

// element count of input k (factor inputs are shape[m] x R)
i64 input_numel(const tenrso_ad_op* op, int k) {
  const bool has_x = op->kind != TENRSO_AD_CP_RECONSTRUCT;
  if (has_x && k == 0) return numel(op);
  const int m = has_x ? k - 1 : k;
  return op->shape[m] * op->R;
}

int launch_zero(tenrso_ctx* ctx, float* p, i64 n) {
  if (n == 0) return TENRSO_OK;
  hipLaunchKernelGGL(zero_kernel, dim3(cdiv(n, 256)), dim3(256), 0, ctx->stream, p, n);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

int launch_scale(tenrso_ctx* ctx, float* p, i64 n, const float* g) {
  if (n == 0) return TENRSO_OK;
  hipLaunchKernelGGL(scale_by_kernel, dim3(cdiv(n, 256)), dim3(256), 0, ctx->stream, p, n, g);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

int backward_mttkrp(tenrso_ad_op* op, const float* mbar, float* const* gin) {
  tenrso_ctx* ctx = op->ctx;
  float* dX = gin ? gin[0] : nullptr;
  float* dF[MAXD] = {};
  if (gin)
    for (int m = 0; m < op->nd; ++m) dF[m] = gin[1 + m];
  float* dmode = dF[op->mode];
  dF[op->mode] = nullptr;
  TENRSO_TRY(mttkrp_vjp(ctx, op->X, op->nd, op->shape, op->F, op->R, op->mode, mbar, dF, dX));
  if (dmode) TENRSO_TRY(launch_zero(ctx, dmode, op->shape[op->mode] * op->R));
  return TENRSO_OK;
}

int backward_reconstruct(tenrso_ad_op* op, const float* gout, float* const* gin) {
  if (!gin) return TENRSO_OK;
  return cp_factor_grads(op->ctx, gout, op->nd, op->shape, op->F, op->R, op->w, gin);
}

int backward_loss(tenrso_ad_op* op, const float* g, float* const* gin) {
  tenrso_ctx* ctx = op->ctx;
  float* dX = gin ? gin[0] : nullptr;
  float* dF[MAXD] = {};
  bool anyF = false;
  if (gin)
    for (int m = 0; m < op->nd; ++m) {
      dF[m] = gin[1 + m];
      anyF = anyF || dF[m];
    }
  if (anyF) {
    TENRSO_TRY(cp_loss_grad_dev(ctx, op->X, op->nd, op->shape, op->F, op->R, dF, nullptr));
    for (int m = 0; m < op->nd; ++m)
      if (dF[m]) TENRSO_TRY(launch_scale(ctx, dF[m], op->shape[m] * op->R, g));
  }
  if (dX) TENRSO_TRY(cp_residual_f32(ctx, op->nd, op->shape, op->F, op->R, nullptr, op->X, g, dX));
  return TENRSO_OK;
}

}  // namespace
}  // namespace tenrso

using namespace tenrso;

extern "C" {

int tenrso_ad_op_create(tenrso_ctx* ctx, int kind, const float* X, int ndim, const int64_t* shape,
                        const float* const* factors, int64_t rank, int mode, const float* weights,
                        tenrso_ad_op** out) {
  clear_error();
  if (!out) {
    set_error("output pointer is null");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  if (!ctx) {
    set_error("null context");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  if (kind < TENRSO_AD_MTTKRP || kind > TENRSO_AD_CP_LOSS) {
    set_error("unknown AD operation kind %d", kind);
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  if (!factors || !shape) {
    set_error("null factors or shape");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  TENRSO_TRY(mttkrp_check_args(ndim, shape, rank, kind == TENRSO_AD_MTTKRP ? mode : 0));
  if (kind != TENRSO_AD_CP_RECONSTRUCT && !X) {
    set_error("operation %s needs the tensor X", OP_NAMES[kind]);
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  if (kind == TENRSO_AD_CP_LOSS && (rank <= 0 || rank > 64)) {
    set_error("Invalid rank: %lld (the loss rule's FP64 Gram bound is 64)", (long long)rank);
    return TENRSO_ERR_INVALID_RANK;
  }
  if (ndim > MAXD_RECON) {
    set_error("tensors with more than %d modes are not supported", MAXD_RECON);
    return TENRSO_ERR_UNSUPPORTED;
  }
  tenrso_ad_op* op = new tenrso_ad_op();
  op->ctx = ctx;
  op->kind = kind;
  op->nd = ndim;
  op->mode = kind == TENRSO_AD_MTTKRP ? mode : 0;
  op->R = rank;
  op->X = kind == TENRSO_AD_CP_RECONSTRUCT ? nullptr : X;
  op->w = kind == TENRSO_AD_CP_RECONSTRUCT ? weights : nullptr;
  for (int k = 0; k < ndim; ++k) {
    op->shape[k] = shape[k];
    op->F[k] = factors[k];
  }
  *out = op;
  return TENRSO_OK;
}

int tenrso_ad_op_destroy(tenrso_ad_op* op) {
  delete op;
  return TENRSO_OK;
}

const char* tenrso_ad_op_name(const tenrso_ad_op* op) { return op ? OP_NAMES[op->kind] : ""; }

int tenrso_ad_op_num_inputs(const tenrso_ad_op* op) {
  if (!op) return -1;
  return op->kind == TENRSO_AD_CP_RECONSTRUCT ? op->nd : op->nd + 1;
}

int tenrso_ad_op_num_outputs(const tenrso_ad_op* op) { return op ? 1 : -1; }

int64_t tenrso_ad_op_input_numel(const tenrso_ad_op* op, int k) {
  if (!op || k < 0 || k >= tenrso_ad_op_num_inputs(op)) return -1;
  return input_numel(op, k);
}

int64_t tenrso_ad_op_output_numel(const tenrso_ad_op* op, int k) {
  if (!op || k != 0) return -1;
  switch (op->kind) {
    case TENRSO_AD_MTTKRP: return op->shape[op->mode] * op->R;
    case TENRSO_AD_CP_RECONSTRUCT: return numel(op);
    default: return 1;
  }
}

int tenrso_ad_op_forward(tenrso_ad_op* op, float* const* outputs) {
  clear_error();
  if (!op || !outputs || !outputs[0]) {
    set_error("null operation or output buffer");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  tenrso_ctx* ctx = op->ctx;
  TENRSO_TRY(set_device(ctx));
  switch (op->kind) {
    case TENRSO_AD_MTTKRP:
      if (op->R == 0 || op->shape[op->mode] == 0) return TENRSO_OK;
      return mttkrp_dense(ctx, TENRSO_PREC_FP32, op->X, nullptr, op->nd, op->shape, op->F, op->R, op->mode,
                          MttkrpOut::F32, outputs[0], nullptr);
    case TENRSO_AD_CP_RECONSTRUCT:
      return cp_reconstruct_f32(ctx, op->nd, op->shape, op->F, op->R, op->w, outputs[0]);
    default: {
      const double* loss = nullptr;
      TENRSO_TRY(cp_loss_grad_dev(ctx, op->X, op->nd, op->shape, op->F, op->R, nullptr, &loss));
      hipLaunchKernelGGL(loss_to_f32_kernel, dim3(1), dim3(64), 0, ctx->stream, loss, outputs[0]);
      TENRSO_HIP_TRY(hipGetLastError());
      return TENRSO_OK;
    }
  }
}

int tenrso_ad_op_backward(tenrso_ad_op* op, const float* const* output_grads, float* const* input_grads) {
  clear_error();
  if (!op || !output_grads || !output_grads[0]) {
    set_error("null operation or output gradient");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  TENRSO_TRY(set_device(op->ctx));
  if (op->R == 0) {
    if (input_grads)
      for (int k = 0; k < tenrso_ad_op_num_inputs(op); ++k)
        if (input_grads[k]) TENRSO_TRY(launch_zero(op->ctx, input_grads[k], input_numel(op, k)));
    return TENRSO_OK;
  }
  switch (op->kind) {
    case TENRSO_AD_MTTKRP: return backward_mttkrp(op, output_grads[0], input_grads);
    case TENRSO_AD_CP_RECONSTRUCT: return backward_reconstruct(op, output_grads[0], input_grads);
    default: return backward_loss(op, output_grads[0], input_grads);
  }
}

}  // extern "C"
