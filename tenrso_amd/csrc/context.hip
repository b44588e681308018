// context.hip -- context lifetime, device buffer pool and error strings.
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "common.hpp"
#include "mttkrp.hpp"

namespace tenrso {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

void clear_error() { g_last_error.clear(); }

int ctx_buffer(tenrso_ctx* ctx, int slot, size_t bytes, void** out) {
  if ((int)ctx->bufs.size() < SLOT_COUNT) ctx->bufs.resize(SLOT_COUNT);
  DevBuf& b = ctx->bufs[slot];
  if (bytes == 0) bytes = 256;
  if (b.bytes < bytes) {
    if (b.ptr) {
      // the stream may still use the old buffer
      TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
      TENRSO_HIP_TRY(hipFree(b.ptr));
      b.ptr = nullptr;
      b.bytes = 0;
    }
    const size_t want = (bytes + 4095) & ~(size_t)4095;
    hipError_t e = hipMalloc(&b.ptr, want);
    if (e != hipSuccess) {
      set_error("device allocation of %zu bytes failed: %s", want, hipGetErrorString(e));
      b.ptr = nullptr;
      return TENRSO_ERR_DEVICE;
    }
    b.bytes = want;
  }
  *out = b.ptr;
  return TENRSO_OK;
}

}  // namespace tenrso

using namespace tenrso;

extern "C" {

const char* tenrso_last_error(void) { return g_last_error.c_str(); }

const char* tenrso_version(void) { return "tenrso-hip 0.1.0 (gfx950)"; }

int tenrso_ctx_create(int device, int borrow_stream, void* stream, tenrso_ctx** out) {
  clear_error();
  if (!out) {
    set_error("output pointer is null");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  int ndev = 0;
  TENRSO_HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) {
    set_error("device %d out of range (%d devices)", device, ndev);
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  TENRSO_HIP_TRY(hipSetDevice(device));
  tenrso_ctx* c = new tenrso_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    c->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
      set_error("device %d is %s; this library is built for gfx950 (MI355X)", device, prop.gcnArchName);
      delete c;
      return TENRSO_ERR_UNSUPPORTED;
    }
  }
  {
    // the no-allocation entry point's workspace address on this device
    const int rc = noalloc_scratch_init(device);
    if (rc != TENRSO_OK) {
      delete c;
      return rc;
    }
  }
  if (borrow_stream) {
    c->stream = (hipStream_t)stream;
    c->own_stream = false;
  } else {
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      set_error("hipStreamCreate failed: %s", hipGetErrorString(e));
      delete c;
      return TENRSO_ERR_DEVICE;
    }
    c->own_stream = true;
  }
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      hipHostMalloc(&c->pinned, 4096, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void**)&c->pinned_dev, c->pinned, 0) != hipSuccess) {
    set_error("context setup failed");
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return TENRSO_ERR_DEVICE;
  }
  c->bufs.resize(SLOT_COUNT);
  if (const char* e = std::getenv("TENRSO_WG_PER_CU_F32")) c->wg_per_cu_f32 = std::max(1, std::atoi(e));
  if (const char* e = std::getenv("TENRSO_WG_PER_CU_BF16")) c->wg_per_cu_bf16 = std::max(1, std::atoi(e));
  if (const char* e = std::getenv("TENRSO_JOB_GATE")) c->job_gate = std::atoi(e) != 0;
  *out = c;
  return TENRSO_OK;
}

int tenrso_ctx_destroy(tenrso_ctx* c) {
  if (!c) return TENRSO_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (c->job_stream) (void)hipStreamSynchronize(c->job_stream);  // before its buffers go
  tenrso::blas_destroy(c);
  for (auto& b : c->bufs)
    if (b.ptr) (void)hipFree(b.ptr);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->job_stream) (void)hipStreamDestroy(c->job_stream);
  if (c->job_fork) (void)hipEventDestroy(c->job_fork);
  if (c->job_join) (void)hipEventDestroy(c->job_join);
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->own_stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return TENRSO_OK;
}

int tenrso_ctx_synchronize(tenrso_ctx* c) {
  if (!c) {
    set_error("null context");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  TENRSO_HIP_TRY(hipStreamSynchronize(c->stream));
  return TENRSO_OK;
}

int tenrso_ctx_set_timing(tenrso_ctx* c, int enabled) {
  if (!c) {
    set_error("null context");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  c->timing = enabled != 0;
  return TENRSO_OK;
}

int tenrso_ctx_last_kernel_ms(tenrso_ctx* c, double* ms) {
  if (!c || !ms) {
    set_error("null argument");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  *ms = c->last_kernel_ms;
  return TENRSO_OK;
}

}  // extern "C"
