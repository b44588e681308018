// common.hpp -- shared host/device helpers of the gfx950 CP library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/tenrso_hip.h"

namespace tenrso {

using i64 = int64_t;

// ---- error reporting ------------------------------------------------------
void set_error(const char* fmt, ...);
void clear_error();

struct Status {
  int code = TENRSO_OK;
  bool ok() const { return code == TENRSO_OK; }
};

#define TENRSO_HIP_TRY(expr)                                                                      \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) {                                                                       \
      ::tenrso::set_error("HIP error %s at %s:%d (%s)", hipGetErrorString(e_), __FILE__, __LINE__, \
                          #expr);                                                                 \
      return TENRSO_ERR_DEVICE;                                                                   \
    }                                                                                             \
  } while (0)

#define TENRSO_TRY(expr)         \
  do {                           \
    int rc_ = (expr);            \
    if (rc_ != TENRSO_OK) return rc_; \
  } while (0)

// ---- device buffer pool ---------------------------------------------------
// Grow-only named buffers owned by a context (partials, Grams, staging).
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace tenrso

struct tenrso_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int num_cus = 256;
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_kernel_ms = 0.0;
  int profile_reps = 0;               // > 0: mttkrp_dense times its main kernel
  int wg_per_cu_f32 = 0;              // MTTKRP workgroups per CU; 0: the kernel's occupancy
  int wg_per_cu_bf16 = 0;             // (TENRSO_WG_PER_CU_F32 / _BF16 override)
  int job_gate = 1;                   // BF16 MTTKRP waits for the job's Gram loads (TENRSO_JOB_GATE=0: off)
  unsigned gate_stamp = 0;            // value the next gated launch waits for
  std::vector<tenrso::DevBuf> bufs;  // indexed by slot id
  void* pinned = nullptr;            // small pinned (coherent) host buffer for scalar readback
  double* pinned_dev = nullptr;      // the same buffer as a device pointer (kernels write the fit there)
  // workgroups per CU of each MTTKRP kernel instance (occupancy query cache)
  const void* occ_kern[32] = {};
  int occ_val[32] = {};
  int occ_n = 0;
  void* blas = nullptr;              // rocBLAS handle of the n-mode products, made on first use
  // side stream of the wide-rank ALS job (made on first use): the job overlaps the
  // MTTKRP launched after it, and the next row solve joins it
  hipStream_t job_stream = nullptr;
  hipEvent_t job_fork = nullptr, job_join = nullptr;
  bool job_join_pending = false;
};

namespace tenrso {

enum BufSlot : int {
  SLOT_PARTIAL = 0,
  SLOT_M64,
  SLOT_GRAMS,
  SLOT_PRE,
  SLOT_SOLVE,
  SLOT_SCALARS,
  SLOT_XBF16,
  SLOT_XBF16_T,
  SLOT_STAGE_X,
  SLOT_STAGE_F,
  SLOT_STAGE_OUT,
  SLOT_TMP,
  SLOT_TMP2,
  SLOT_IP,
  SLOT_SPM,
  SLOT_ORTHO,
  SLOT_FSET,
  SLOT_EIG_A,
  SLOT_EIG_V,
  SLOT_EIG_PART,
  SLOT_EIG_W,
  SLOT_GATE,
  SLOT_GRAM_CHUNKS,
  SLOT_XPAD,
  SLOT_COUNT
};

int ctx_buffer(tenrso_ctx* ctx, int slot, size_t bytes, void** out);

// releases the context's rocBLAS handle (nmode.hip)
void blas_destroy(tenrso_ctx* ctx);

inline int set_device(tenrso_ctx* ctx) {
  TENRSO_HIP_TRY(hipSetDevice(ctx->device));
  return TENRSO_OK;
}

}  // namespace tenrso
