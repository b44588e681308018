// ooc.hip -- out-of-core (streaming) dense MTTKRP.
//
// The reference's streaming MTTKRP (tenrso-ooc/examples/mttkrp_streaming.rs:66-110)
// is a stub that calls the in-memory mttkrp; its comments state the intended
// algorithm: chunk the tensor into slices, compute a partial MTTKRP per chunk,
// accumulate.  Here that algorithm runs for a tensor held in HOST memory that
// need not fit in a given device-memory budget:
//   * the tensor is cut along mode 0 into slabs X[i0 .. i0 + s, ...] (contiguous
//     in a C-order tensor), s chosen so that two slab buffers, the factors and the
//     partial sums fit in the budget;
//   * slabs are copied host -> device on a copy stream into two device buffers
//     in turn (double buffering) while the MFMA MTTKRP of the previous slab runs
//     on the context stream; events order reuse of a buffer after its kernels;
//   * a pageable host tensor is first copied, slab by slab, into two pinned
//     staging buffers by the host thread (DMA needs pinned memory to overlap);
//     a tensor the caller allocated pinned (hipHostMalloc) is copied directly;
//   * mode 0: each slab yields the output rows i0 .. i0 + s; mode n > 0: each
//     slab yields a full I_n x R partial (the slab's rows of A_0 only), summed
//     over slabs in FP64 in slab order (deterministic), rounded to fp32 at the
//     end.
// The per-slab kernel is the in-memory MFMA MTTKRP (mttkrp_dense), so ragged
// slab sizes and every mode take the same code as the in-core path.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "mttkrp.hpp"

namespace tenrso {

namespace {

inline unsigned cdiv(i64 a, i64 b) { return (unsigned)((a + b - 1) / b); }

__global__ __launch_bounds__(256) void add_f64_kernel(double* __restrict__ acc, const double* __restrict__ x, i64 n) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e < n) acc[e] += x[e];
}

__global__ __launch_bounds__(256) void f64_to_f32_kernel(const double* __restrict__ x, float* __restrict__ y, i64 n) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e < n) y[e] = (float)x[e];
}

__global__ __launch_bounds__(256) void fill_f64_zero_kernel(double* __restrict__ p, i64 n) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e < n) p[e] = 0.0;
}

// Every allocation of one streaming call, released on every path out of it.
struct OocResources {
  hipStream_t main = nullptr;  // the context stream, drained before anything is freed
  std::vector<void*> dev;
  std::vector<void*> pinned;
  hipStream_t copy = nullptr;
  hipEvent_t ev[8] = {};
  int nev = 0;
  int alloc(void** p, size_t bytes) {
    *p = nullptr;
    TENRSO_HIP_TRY(hipMalloc(p, std::max<size_t>(bytes, 256)));
    dev.push_back(*p);
    return TENRSO_OK;
  }
  int alloc_pinned(void** p, size_t bytes) {
    *p = nullptr;
    TENRSO_HIP_TRY(hipHostMalloc(p, std::max<size_t>(bytes, 256), hipHostMallocDefault));
    pinned.push_back(*p);
    return TENRSO_OK;
  }
  int event(hipEvent_t* e) {
    TENRSO_HIP_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming));
    ev[nev++] = *e;
    return TENRSO_OK;
  }
  ~OocResources() {
    if (main) (void)hipStreamSynchronize(main);
    if (copy) (void)hipStreamSynchronize(copy);
    for (int i = 0; i < nev; ++i) (void)hipEventDestroy(ev[i]);
    if (copy) (void)hipStreamDestroy(copy);
    for (void* p : dev) (void)hipFree(p);
    for (void* p : pinned) (void)hipHostFree(p);
  }
};

}  // namespace

int mttkrp_streaming_f32(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, const float* const* F, i64 R,
                         int mode, float* out, i64 budget, double* stats) {
  const i64 I0 = shape[0];
  i64 slab_elems_per_row = 1;
  for (int k = 1; k < nd; ++k) slab_elems_per_row *= shape[k];
  const i64 row_bytes = slab_elems_per_row * (i64)sizeof(float);
  const i64 In = shape[mode];
  // device memory besides the slab buffers: factors, the FP64 accumulator and
  // one slab's FP64 partial, and the MTTKRP's own split partials: its grid of at
  // most 512 workgroups gives each 32-row tile at most 512 / tiles splits, so
  // splits x I_n x R fp32 stays below 512 x 32 x R x 4 bytes (the generic kernel's
  // below 512 x 256 x 4), plus a margin
  i64 ftot = 0;
  for (int k = 0; k < nd; ++k) ftot += shape[k] * R;
  const i64 fixed = ftot * 4 + 2 * In * R * 8 + std::max<i64>(512 * 32 * 4 * R, 512 * 256 * 4) + ((i64)256 << 10);
  if (budget <= fixed + 2 * row_bytes) {
    set_error("device budget of %lld bytes cannot hold two slabs of one mode-0 row (%lld bytes) besides %lld bytes of "
              "factors and partial sums",
              (long long)budget, (long long)row_bytes, (long long)fixed);
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  const i64 s = std::max<i64>(1, std::min<i64>(I0, (budget - fixed) / 2 / row_bytes));
  const i64 nslab = (I0 + s - 1) / s;
  const size_t slab_bytes = (size_t)(s * row_bytes);

  OocResources rs;
  rs.main = ctx->stream;
  hipPointerAttribute_t attr{};
  bool host_pinned = false;
  if (hipPointerGetAttributes(&attr, X) == hipSuccess)
    host_pinned = attr.type == hipMemoryTypeHost;
  (void)hipGetLastError();  // a pageable pointer is not an error

  void *dslab[2], *dF, *dacc, *dtmp;
  TENRSO_TRY(rs.alloc(&dslab[0], slab_bytes));
  TENRSO_TRY(rs.alloc(&dslab[1], slab_bytes));
  TENRSO_TRY(rs.alloc(&dF, (size_t)ftot * 4));
  TENRSO_TRY(rs.alloc(&dacc, (size_t)In * R * 8));
  TENRSO_TRY(rs.alloc(&dtmp, (size_t)In * R * 8));
  void* stage[2] = {nullptr, nullptr};
  if (!host_pinned) {
    TENRSO_TRY(rs.alloc_pinned(&stage[0], slab_bytes));
    TENRSO_TRY(rs.alloc_pinned(&stage[1], slab_bytes));
  }
  TENRSO_HIP_TRY(hipStreamCreateWithFlags(&rs.copy, hipStreamNonBlocking));
  hipEvent_t copied[2], consumed[2], staged_free[2];
  for (int b = 0; b < 2; ++b) {
    TENRSO_TRY(rs.event(&copied[b]));
    TENRSO_TRY(rs.event(&consumed[b]));
    TENRSO_TRY(rs.event(&staged_free[b]));
  }
  // factors to the device
  std::vector<const float*> dFk(nd);
  {
    i64 off = 0;
    for (int k = 0; k < nd; ++k) {
      float* p = (float*)dF + off;
      TENRSO_HIP_TRY(hipMemcpyAsync(p, F[k], (size_t)shape[k] * R * 4, hipMemcpyHostToDevice, ctx->stream));
      dFk[k] = p;
      off += shape[k] * R;
    }
  }
  if (mode != 0)
    hipLaunchKernelGGL(fill_f64_zero_kernel, dim3(cdiv(In * R, 256)), dim3(256), 0, ctx->stream, (double*)dacc,
                       In * R);
  TENRSO_HIP_TRY(hipGetLastError());
  TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));

  auto t0 = std::chrono::steady_clock::now();
  double copy_s = 0.0;
  // queue the copy of slab k into device buffer k % 2 (after the kernels that
  // read that buffer last, slab k - 2, are done)
  auto issue_copy = [&](i64 k) -> int {
    const int b = (int)(k & 1);
    const i64 r0 = k * s, rows = std::min<i64>(s, I0 - r0);
    const size_t bytes = (size_t)(rows * row_bytes);
    const char* src = reinterpret_cast<const char*>(X) + (size_t)(r0 * row_bytes);
    if (k >= 2) TENRSO_HIP_TRY(hipStreamWaitEvent(rs.copy, consumed[b], 0));
    if (!host_pinned) {
      // the DMA from this staging buffer (slab k - 2) must be done before the
      // host overwrites it
      if (k >= 2) TENRSO_HIP_TRY(hipEventSynchronize(staged_free[b]));
      auto c0 = std::chrono::steady_clock::now();
      std::memcpy(stage[b], src, bytes);
      copy_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
      src = reinterpret_cast<const char*>(stage[b]);
    }
    TENRSO_HIP_TRY(hipMemcpyAsync(dslab[b], src, bytes, hipMemcpyHostToDevice, rs.copy));
    if (!host_pinned) TENRSO_HIP_TRY(hipEventRecord(staged_free[b], rs.copy));
    TENRSO_HIP_TRY(hipEventRecord(copied[b], rs.copy));
    return TENRSO_OK;
  };

  // slab k's kernels are queued before the host stages slab k + 1, so the host
  // copy into pinned memory overlaps the DMA and the kernels of slab k
  TENRSO_TRY(issue_copy(0));
  for (i64 k = 0; k < nslab; ++k) {
    const int b = (int)(k & 1);
    const i64 r0 = k * s, rows = std::min<i64>(s, I0 - r0);
    TENRSO_HIP_TRY(hipStreamWaitEvent(ctx->stream, copied[b], 0));
    i64 sshape[MAXD];
    for (int q = 0; q < nd; ++q) sshape[q] = shape[q];
    sshape[0] = rows;
    std::vector<const float*> Fs(dFk);
    Fs[0] = dFk[0] + r0 * R;  // the slab's rows of A_0
    // the per-slab output: rows r0 .. r0 + rows of a device fp32 result (mode 0)
    // or an FP64 partial added into the accumulator (other modes)
    if (mode == 0) {
      TENRSO_TRY(mttkrp_dense(ctx, TENRSO_PREC_FP32, dslab[b], nullptr, nd, sshape, Fs.data(), R, 0, MttkrpOut::F64,
                              (double*)dacc + r0 * R, nullptr));
    } else {
      TENRSO_TRY(mttkrp_dense(ctx, TENRSO_PREC_FP32, dslab[b], nullptr, nd, sshape, Fs.data(), R, mode,
                              MttkrpOut::F64, dtmp, nullptr));
      hipLaunchKernelGGL(add_f64_kernel, dim3(cdiv(In * R, 256)), dim3(256), 0, ctx->stream, (double*)dacc,
                         (const double*)dtmp, In * R);
      TENRSO_HIP_TRY(hipGetLastError());
    }
    TENRSO_HIP_TRY(hipEventRecord(consumed[b], ctx->stream));
    if (k + 1 < nslab) TENRSO_TRY(issue_copy(k + 1));
  }
  // fp32 result to the host (through the second slab buffer, which is free now)
  TENRSO_HIP_TRY(hipStreamWaitEvent(ctx->stream, consumed[(nslab - 1) & 1], 0));
  float* d32 = (float*)dtmp;  // I_n x R fp32 fits in the FP64 temporary
  hipLaunchKernelGGL(f64_to_f32_kernel, dim3(cdiv(In * R, 256)), dim3(256), 0, ctx->stream, (const double*)dacc, d32,
                     In * R);
  TENRSO_HIP_TRY(hipGetLastError());
  TENRSO_HIP_TRY(hipMemcpyAsync(out, d32, (size_t)In * R * 4, hipMemcpyDeviceToHost, ctx->stream));
  TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (stats) {
    stats[0] = secs;
    stats[1] = (double)I0 * (double)row_bytes / secs / 1e9;  // tensor bytes moved host -> device per second
    stats[2] = (double)s;
    stats[3] = (double)nslab;
    stats[4] = host_pinned ? 1.0 : 0.0;
    stats[5] = copy_s;
  }
  return TENRSO_OK;
}

}  // namespace tenrso

using namespace tenrso;

extern "C" int tenrso_mttkrp_f32_streaming(tenrso_ctx* ctx, const float* X, int ndim, const int64_t* shape,
                                           const float* const* factors, int64_t rank, int mode, float* out,
                                           int64_t device_budget_bytes, double* stats) {
  clear_error();
  if (!ctx) {
    set_error("null context");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  TENRSO_TRY(mttkrp_check_args(ndim, shape, rank, mode));
  if (!X || !factors || !out) {
    set_error("null tensor, factors or output");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  if (rank == 0 || shape[mode] == 0) return TENRSO_OK;
  for (int k = 0; k < ndim; ++k)
    if (shape[k] == 0) {
      std::memset(out, 0, (size_t)shape[mode] * rank * sizeof(float));
      return TENRSO_OK;
    }
  TENRSO_HIP_TRY(hipSetDevice(ctx->device));
  return mttkrp_streaming_f32(ctx, X, ndim, shape, factors, rank, mode, out, device_budget_bytes, stats);
}
