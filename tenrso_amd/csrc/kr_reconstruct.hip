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
  const float* xsub;    // not null: out = xsub - reconstruction (a residual)
  const float* gscale;  // not null: out *= gscale[0] (a device scalar)
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
  if (a.xsub) acc = a.xsub[x] - acc;
  if (a.gscale) acc = acc * a.gscale[0];
  out[x] = acc;
}

int cp_reconstruct_f32(tenrso_ctx* ctx, int nd, const i64* shape, const float* const* F, i64 R, const float* w,
                       float* out) {
  return cp_residual_f32(ctx, nd, shape, F, R, w, nullptr, nullptr, out);
}

int cp_residual_f32(tenrso_ctx* ctx, int nd, const i64* shape, const float* const* F, i64 R, const float* w,
                    const float* X, const float* gscale, float* out) {
  ReconArgs a{};
  a.xsub = X;
  a.gscale = gscale;
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

// ---- extract_weights (cp/types.rs:421-446) ----------------------------------
// For every factor k and column r: n = ||F_k[:, r]||; where n > f32 epsilon the
// column is divided by n and w_r is multiplied by n (w starts at ones).  The
// norms are formed in FP64 (squares of the fp32 entries, fixed order), rounded to
// fp32 for the division, as the reference's f32 instantiation divides by its f32
// norm.  One workgroup per (factor, 64-column block): 16 row lanes x 64 columns,
// each row lane sums rows lane, lane + 16, ...; the 16 sums are added in lane
// order in LDS.  A second launch multiplies the weights in factor order.
struct NormArgs {
  float* F[MAXD_RECON];
  i64 rows[MAXD_RECON];
};

constexpr int EW_COLS = 64, EW_LANES = 16;

__global__ __launch_bounds__(EW_COLS* EW_LANES) void column_normalize_kernel(NormArgs a, i64 R, int cblocks,
                                                                              double* __restrict__ norms) {
  __shared__ double part[EW_LANES][EW_COLS];
  __shared__ float scale[EW_COLS];
  const int k = blockIdx.x / cblocks;
  const i64 c = (i64)(blockIdx.x - k * cblocks) * EW_COLS + (threadIdx.x % EW_COLS);
  const int lane = threadIdx.x / EW_COLS;
  float* __restrict__ F = a.F[k];
  const i64 I = a.rows[k];
  double s = 0.0;
  if (c < R)
    for (i64 i = lane; i < I; i += EW_LANES) {
      const double v = F[i * R + c];
      s += v * v;
    }
  part[lane][threadIdx.x % EW_COLS] = s;
  __syncthreads();
  if (lane == 0) {
    double t = 0.0;
    for (int l = 0; l < EW_LANES; ++l) t += part[l][threadIdx.x];
    const double n = sqrt(t);
    const bool keep = n > (double)__FLT_EPSILON__;
    scale[threadIdx.x] = keep ? (float)n : 0.0f;
    if (c < R) norms[(i64)k * R + c] = keep ? n : 1.0;
  }
  __syncthreads();
  const float n = scale[threadIdx.x % EW_COLS];
  if (c < R && n != 0.0f)
    for (i64 i = lane; i < I; i += EW_LANES) F[i * R + c] = F[i * R + c] / n;
}

__global__ __launch_bounds__(256) void weights_from_norms_kernel(const double* __restrict__ norms, int nd, i64 R,
                                                                 float* __restrict__ w) {
  const i64 r = (i64)blockIdx.x * 256 + threadIdx.x;
  if (r >= R) return;
  double p = 1.0;
  for (int k = 0; k < nd; ++k) p *= norms[(i64)k * R + r];
  w[r] = (float)p;
}

int cp_extract_weights_f32(tenrso_ctx* ctx, int nd, const i64* shape, float* const* F, i64 R, float* w) {
  if (R == 0) return TENRSO_OK;
  NormArgs a{};
  for (int k = 0; k < nd; ++k) {
    a.F[k] = F[k];
    a.rows[k] = shape[k];
  }
  void* norms;
  TENRSO_TRY(ctx_buffer(ctx, SLOT_TMP2, (size_t)nd * R * sizeof(double), &norms));
  const int cblocks = (int)cdiv(R, EW_COLS);
  hipLaunchKernelGGL(column_normalize_kernel, dim3(nd * cblocks), dim3(EW_COLS * EW_LANES), 0, ctx->stream, a, R,
                     cblocks, (double*)norms);
  TENRSO_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(weights_from_norms_kernel, dim3(cdiv(R, 256)), dim3(256), 0, ctx->stream,
                     (const double*)norms, nd, R, w);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

}  // namespace tenrso
