// cp_als.hpp -- device CP-ALS driver.
#pragma once
#include <functional>

#include "common.hpp"
#include "mttkrp.hpp"

namespace tenrso {

struct AlsJob;

// What the ALS loop needs from a tensor format: ||X||^2 and min(X) into the
// scalar block, an optional one-time preparation, and the mode-n MTTKRP as fp32
// split partials [splits][I_n][R] (summed in FP64 by the row solve).  The MTTKRP
// must also run `job` (when not null) before the row solve that follows it: in
// the same launch where it can, else as its own kernel.
struct AlsProblem {
  int nd = 0;
  const i64* shape = nullptr;
  // ||X||^2 and min(X) into scal[SC_NORM_X], scal[SC_MIN_X], the other slots of
  // the scalar block cleared; tmp: scratch for the partials
  std::function<int(double* tmp, double* scal)> norm_min;
  std::function<int()> prepare;
  std::function<int(int mode, const float* const* F, const AlsJob* job, MttkrpResult* out)> mttkrp;
};

int cp_als_run(tenrso_ctx* ctx, AlsProblem& pb, float* const* F, const tenrso_cp_options* o, tenrso_cp_result* res,
               double* hist);
int cp_als_device(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, float* const* F,
                  const tenrso_cp_options* o, tenrso_cp_result* res, double* hist);

}  // namespace tenrso
