// cp_als.hpp -- device CP-ALS driver.
#pragma once
#include "common.hpp"

namespace tenrso {

int cp_als_device(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, float* const* F,
                  const tenrso_cp_options* o, tenrso_cp_result* res, double* hist);

}  // namespace tenrso
