// kr_reconstruct.hpp -- Khatri-Rao product and CP reconstruction launchers.
#pragma once
#include "common.hpp"

namespace tenrso {

constexpr int MAXD_RECON = 8;

int khatri_rao_f32(tenrso_ctx* ctx, const float* A, i64 I, const float* B, i64 J, i64 K, float* C);
int cp_reconstruct_f32(tenrso_ctx* ctx, int nd, const i64* shape, const float* const* F, i64 R, const float* w,
                       float* out);
// out = gscale[0] * (X - reconstruction); X or gscale may be null (then the
// reconstruction itself, unscaled)
int cp_residual_f32(tenrso_ctx* ctx, int nd, const i64* shape, const float* const* F, i64 R, const float* w,
                    const float* X, const float* gscale, float* out);
int cp_extract_weights_f32(tenrso_ctx* ctx, int nd, const i64* shape, float* const* F, i64 R, float* w);

}  // namespace tenrso
