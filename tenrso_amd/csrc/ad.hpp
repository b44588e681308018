// ad.hpp -- VJP / gradient launchers of the CP path.
#pragma once
#include "common.hpp"

namespace tenrso {

int mttkrp_vjp(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, const float* const* F, i64 R, int mode,
               const float* mbar, float* const* dF, float* dX);
int cp_factor_grads(tenrso_ctx* ctx, const float* Gt, int nd, const i64* shape, const float* const* F, i64 R,
                    const float* w, float* const* grads);
int cp_loss_grad(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, const float* const* F, i64 R,
                 float* const* grads, double* loss_host);
// the same with the loss left on the device: *loss_dev points at it (FP64,
// context scratch, valid until the next call on the context)
int cp_loss_grad_dev(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, const float* const* F, i64 R,
                     float* const* grads, const double** loss_dev);
int cp_gd(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, float* const* F, i64 R, int steps, double lr,
          double* loss_hist);

}  // namespace tenrso
