// cp_init.hpp -- SVD-based CP initialisations on the device.
#pragma once
#include "common.hpp"

namespace tenrso {

// factors F[k] (shape[k] x R, fp32, device) from the SVD of each unfolding;
// kind: TENRSO_INIT_SVD / _NNSVD / _LEVERAGE; seed drives the N(0, 0.01) draws
// of the columns past min(R, I_k, J_k); sweeps (optional, per mode): Jacobi sweeps used.
int cp_init_svd(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, i64 R, int kind, uint64_t seed,
                float* const* F, int* sweeps);

}  // namespace tenrso
