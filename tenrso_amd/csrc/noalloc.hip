// noalloc.hip -- the MTTKRP entry point that allocates nothing.
//
// tenrso_mttkrp_f32_noalloc has the contract of tenrso_mttkrp_f32 (the
// reference's mttkrp, tenrso-kernels/src/mttkrp.rs:74-146) but takes its device
// workspace -- the split partial sums and, for tensors of order >= 4, the
// Khatri-Rao rows of the slow modes -- from `scratch`, a file-scope device array
// of NOALLOC_SCRATCH_BYTES, instead of the context's grow-only buffers.  It is
// for callers that cannot call hipMalloc (code under graph capture, tight
// serving loops).
//
// `scratch` is a device symbol: in host code its name is a host-side stand-in,
// not a device address.  Its address on each GPU (every GPU holds its own copy)
// is resolved once per device with hipGetSymbolAddress, with that device
// current, when the first context of that device is created
// (noalloc_scratch_init, called by tenrso_ctx_create); only that address is
// ever passed to kernels.
//
// Every call on a GPU shares that GPU's array, and the array keeps the last
// call's contents.  Nothing in it needs to start at zero: the MTTKRP kernels
// write every partial sum they later read.  Calls that overlap on one GPU (two
// streams, two threads) race on the array; callers serialise them.
#include <mutex>

#include "mttkrp.hpp"

namespace tenrso {

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};
}  // namespace

// Called with `device` current (tenrso_ctx_create); resolves the device address
// of `scratch` on that device once.
int noalloc_scratch_init(int device) {
  if (device < 0 || device >= MAX_DEVICES) {
    set_error("device %d out of range for the no-allocation workspace", device);
    return TENRSO_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch_ptr[device]) return TENRSO_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    set_error("hipGetSymbolAddress of the no-allocation workspace failed on device %d: %s", device,
              hipGetErrorString(e));
    return TENRSO_ERR_DEVICE;
  }
  g_scratch_ptr[device] = p;
  return TENRSO_OK;
}

}  // namespace tenrso

using namespace tenrso;

extern "C" int tenrso_mttkrp_f32_noalloc(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape,
                                         const float* const* factors, int64_t rank, int mode, float* out) {
  clear_error();
  if (!ctx) {
    set_error("null context");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  TENRSO_TRY(mttkrp_check_args(ndim, shape, rank, mode));
  if (rank == 0 || shape[mode] == 0) return TENRSO_OK;
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (ctx->device >= 0 && ctx->device < MAX_DEVICES) base = g_scratch_ptr[ctx->device];
  }
  if (!base) {
    set_error("the no-allocation workspace of device %d was not set up (create a context on it first)",
              ctx->device);
    return TENRSO_ERR_DEVICE;
  }
  TENRSO_HIP_TRY(hipSetDevice(ctx->device));
  MttkrpScratch ws;
  ws.base = base;
  ws.bytes = NOALLOC_SCRATCH_BYTES;
  return mttkrp_dense(ctx, TENRSO_PREC_FP32, X, nullptr, ndim, shape, factors, rank, mode, MttkrpOut::F32, out,
                      nullptr, nullptr, false, &ws);
}
