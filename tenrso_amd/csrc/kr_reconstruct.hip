// kr_reconstruct.hip -- Khatri-Rao product and CP reconstruction kernels.
//
// khatri_rao  (tenrso-kernels/src/khatri_rao.rs:56-89): C[ia*J + ib, k] = A[ia,k] * B[ib,k]
//   One rounding per element, so the result is bit-identical to the reference's
//   serial / parallel / blocked variants.  Write-bound: each thread produces 4
//   consecutive k of one output row (16-byte stores when K % 4 == 0).
// cp_reconstruct (tenrso-kernels/src/outer.rs:223-274):
//   out[i_0..i_{N-1}] = sum_r (prod_k F_k[i_k, r]) * w_r, r ascending, product left to right.
//   Write-bound; each thread owns one output element.
#include "kr_reconstruct.hpp"

namespace tenrso {

static inline unsigned cdiv(i64 a, i64 b) { return (unsigned)((a + b - 1) / b); }

// vec: K % 4 == 0 and A, B, C 16-byte aligned (checked on the host): 16-byte
// loads and stores, else scalar
__global__ __launch_bounds__(256) void khatri_rao_kernel(const float* __restrict__ A, const float* __restrict__ B,
                                                         i64 I, i64 J, i64 K, float* __restrict__ C, int vec) {
  const i64 t = (i64)blockIdx.x * 256 + threadIdx.x;
  const i64 kq = (K + 3) / 4;
  if (t >= I * J * kq) return;
  const i64 row = t / kq, k0 = (t - row * kq) * 4;
  const i64 ia = row / J, ib = row - ia * J;
  const float* a = A + ia * K;
  const float* b = B + ib * K;
  float* c = C + row * K;
  if (vec && k0 + 3 < K) {
    const float4 av = *reinterpret_cast<const float4*>(a + k0);
    const float4 bv = *reinterpret_cast<const float4*>(b + k0);
    *reinterpret_cast<float4*>(c + k0) = make_float4(av.x * bv.x, av.y * bv.y, av.z * bv.z, av.w * bv.w);
  } else {
    for (i64 k = k0; k < K && k < k0 + 4; ++k) c[k] = a[k] * b[k];
  }
}

int khatri_rao_f32(tenrso_ctx* ctx, const float* A, i64 I, const float* B, i64 J, i64 K, float* C) {
  const i64 n = I * J * ((K + 3) / 4);
  if (n == 0) return TENRSO_OK;
  const int vec = K % 4 == 0 && (uintptr_t)A % 16 == 0 && (uintptr_t)B % 16 == 0 && (uintptr_t)C % 16 == 0;
  hipLaunchKernelGGL(khatri_rao_kernel, dim3(cdiv(n, 256)), dim3(256), 0, ctx->stream, A, B, I, J, K, C, vec);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

struct ReconArgs {
  int nd;
  i64 R;
  i64 shape[MAXD_RECON];
  const float* F[MAXD_RECON];
  const float* w;
  i64 total;
};

__global__ __launch_bounds__(256) void reconstruct_kernel(ReconArgs a, float* __restrict__ out) {
  const i64 x = (i64)blockIdx.x * 256 + threadIdx.x;
  if (x >= a.total) return;
  const float* rows[MAXD_RECON];
  i64 rem = x;
  for (int k = a.nd - 1; k >= 0; --k) {
    const i64 ix = rem % a.shape[k];
    rem /= a.shape[k];
    rows[k] = a.F[k] + ix * a.R;
  }
  float acc = 0.0f;
  for (i64 r = 0; r < a.R; ++r) {
    float p = 1.0f;
    for (int k = 0; k < a.nd; ++k) p = p * rows[k][r];
    if (a.w) p = p * a.w[r];
    acc = acc + p;
  }
  out[x] = acc;
}

int cp_reconstruct_f32(tenrso_ctx* ctx, int nd, const i64* shape, const float* const* F, i64 R, const float* w,
                       float* out) {
  ReconArgs a{};
  a.nd = nd;
  a.R = R;
  a.total = 1;
  for (int k = 0; k < nd; ++k) {
    a.shape[k] = shape[k];
    a.F[k] = F[k];
    a.total *= shape[k];
  }
  a.w = w;
  if (a.total == 0) return TENRSO_OK;
  hipLaunchKernelGGL(reconstruct_kernel, dim3(cdiv(a.total, 256)), dim3(256), 0, ctx->stream, a, out);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

}  // namespace tenrso
