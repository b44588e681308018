// mttkrp_dense.hip -- dense mode-n MTTKRP for gfx950 (MI355X).
//
// Contract of tenrso_kernels::mttkrp (tenrso-kernels/src/mttkrp.rs:74-146):
//   V[i, r] = sum_j X_(n)[i, j] * KR[j, r]
// where X_(n) is the forward-order unfolding (mttkrp.rs:149-177) and KR the
// forward-order Khatri-Rao product of the other factors (mttkrp.rs:182-240).
//
// Design (no unfolding copy, no materialised KR):
//   * X is viewed in place as [P, I, Q] (P = prod of modes before n, Q = after).
//     The contraction index j = (p, q) runs over the other modes in forward order,
//     so j = p*Q + q and X[p, i, q] sits at (p*I + i)*Q + q.
//   * j is cut into blocks of KB consecutive indices of the fastest other mode.
//     Within a block KR[j, r] = PRE[j / F_fast, r] * FAST[j % F_fast, r], where
//     FAST is the factor of the fastest other mode and PRE the forward KR of the
//     remaining other modes (a factor itself for 3-way tensors).  Each workgroup
//     writes the KR block, transposed to [r][j], into LDS (double buffered).
//   * Rows i are the MFMA M dimension, rank columns r the N dimension, j the K
//     dimension.  FP32 uses v_mfma_f32_16x16x4_f32; BF16 uses
//     v_mfma_f32_16x16x32_bf16.  For q-contiguous X (mode < N-1, "case A") each
//     lane loads 4 (fp32) or 8 (bf16) consecutive j straight from HBM; for the
//     last mode (Q == 1, "case B", fp32 only) lanes run along i, which is
//     contiguous, and each lane loads one j per MFMA k-step.
//   * Workgroups split j; each writes an [I_tile x R] fp32 partial, summed in a
//     fixed order (FP64) by reduce_partials or by the ALS solve kernel.  No float
//     atomics: results are run-to-run deterministic.
#include <algorithm>

#include "mttkrp.hpp"

namespace tenrso {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

struct KernArgs {
  i64 I, Q;          // X viewed as [P, I, Q]
  i64 R;             // factor row stride (rank)
  i64 fast_size;     // size of the fastest other mode
  const float* pre;  // PRE (rows = J / fast_size, stride R) or null (= 1)
  const float* fast; // FAST factor (fast_size x R)
  i64 nblocks;       // J / KB
  int splits;
  int r0;            // first rank column of this launch (grid.z chunk * 16*NT)
  int rcount;        // rank columns handled by the launch (<= 16*NT)
};

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// KR block (rows j0 .. j0+KB-1, columns r0 .. r0+RT-1) into LDS, transposed [r][jj].
template <int RT, int KB, int PITCH, int NTH, typename ST>
__device__ __forceinline__ void gen_kr_block(ST* __restrict__ s, const KernArgs& a, i64 j0, int tid) {
  const i64 slow = j0 / a.fast_size;
  const i64 fast0 = j0 - slow * a.fast_size;
  static_assert(NTH % RT == 0, "thread count must be a multiple of the rank tile");
  const int r = tid % RT;
  const int jstart = tid / RT;
  const bool valid = r < a.rcount;
  const int rc = valid ? a.r0 + r : a.r0;
  const float pre = (valid && a.pre) ? a.pre[slow * a.R + rc] : 1.0f;
  const float* fp = a.fast + fast0 * a.R + rc;
#pragma unroll
  for (int jj = jstart; jj < KB; jj += NTH / RT) {
    float v = valid ? pre * fp[(i64)jj * a.R] : 0.0f;
    if constexpr (sizeof(ST) == 4) {
      s[r * PITCH + jj] = v;
    } else {
      s[r * PITCH + jj] = f32_to_bf16_rne(v);
    }
  }
}

// ---------------------------------------------------------------------------
// FP32 MFMA kernel.  NT: rank tiles of 16 (RT = 16*NT <= 64).  WAVES: waves per
// workgroup, each owning MT m-tiles of 16 rows.  KB = 32 contraction indices per
// block: 2 halves (h) x 4 lane groups (g) x 4 consecutive (s).
// ---------------------------------------------------------------------------
template <int NT, int WAVES, int MT, bool CASE_B>
__global__ __launch_bounds__(WAVES * 64) void mttkrp_f32_mfma(const float* __restrict__ X, KernArgs a,
                                                              float* __restrict__ partial) {
  constexpr int KB = 32, PITCH = KB + 4, RT = 16 * NT, NTH = WAVES * 64, ROWS = WAVES * 16 * MT;
  __shared__ __attribute__((aligned(16))) float sKR[2][RT * PITCH];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int g = lane >> 4, c = lane & 15;
  const int split = blockIdx.y;
  const i64 b_begin = (i64)split * a.nblocks / a.splits;
  const i64 b_end = (i64)(split + 1) * a.nblocks / a.splits;
  const i64 row_base = (i64)blockIdx.x * ROWS + wave * 16 * MT;

  i64 rowc[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    i64 rr = row_base + mt * 16 + c;
    rowc[mt] = rr < a.I ? rr : a.I - 1;  // clamp: rows past I load valid data, are never stored
  }

  floatx4 acc[MT][NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = floatx4{0.f, 0.f, 0.f, 0.f};

  // xa[mt][h] holds j = j0 + 16h + 4g + s, s = 0..3
  auto loadx = [&](i64 b, floatx4 (&xa)[MT][2]) {
    const i64 j0 = b * KB;
    if constexpr (!CASE_B) {
      const i64 p = j0 / a.Q, q0 = j0 - p * a.Q;
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        const float* base = X + (p * a.I + rowc[mt]) * a.Q + q0 + 4 * g;
        xa[mt][0] = *reinterpret_cast<const floatx4*>(base);
        xa[mt][1] = *reinterpret_cast<const floatx4*>(base + 16);
      }
    } else {
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        const float* base = X + (j0 + 4 * g) * a.I + rowc[mt];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          floatx4 v;
          v[0] = base[(16 * h + 0) * a.I];
          v[1] = base[(16 * h + 1) * a.I];
          v[2] = base[(16 * h + 2) * a.I];
          v[3] = base[(16 * h + 3) * a.I];
          xa[mt][h] = v;
        }
      }
    }
  };

  floatx4 xa[MT][2], xn[MT][2];
  if (b_begin < b_end) {
    gen_kr_block<RT, KB, PITCH, NTH>(sKR[0], a, b_begin * KB, tid);
    loadx(b_begin, xa);
  }
  __syncthreads();
  for (i64 b = b_begin; b < b_end; ++b) {
    const int buf = (int)((b - b_begin) & 1);
    if (b + 1 < b_end) {
      loadx(b + 1, xn);
      gen_kr_block<RT, KB, PITCH, NTH>(sKR[buf ^ 1], a, (b + 1) * KB, tid);
    }
    const float* s = sKR[buf];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const floatx4 bv = *reinterpret_cast<const floatx4*>(&s[(16 * nt + c) * PITCH + 16 * h + 4 * g]);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[mt][h][0], bv[0], acc[mt][nt], 0, 0, 0);
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[mt][h][1], bv[1], acc[mt][nt], 0, 0, 0);
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[mt][h][2], bv[2], acc[mt][nt], 0, 0, 0);
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[mt][h][3], bv[3], acc[mt][nt], 0, 0, 0);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      xa[mt][0] = xn[mt][0];
      xa[mt][1] = xn[mt][1];
    }
  }

  // D layout (16x16): lane (g, c), register v -> row 4g + v, column c
  float* out = partial + (i64)split * a.I * a.R;
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int col = a.r0 + 16 * nt + c;
      if (16 * nt + c >= a.rcount) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const i64 row = row_base + mt * 16 + 4 * g + v;
        if (row < a.I) out[row * a.R + col] = acc[mt][nt][v];
      }
    }
}

// ---------------------------------------------------------------------------
// BF16 MFMA kernel (case A only: q-contiguous X).  KB = 32*KS contraction indices
// per block; each lane holds 8 consecutive j per k-step (16-byte loads).
// ---------------------------------------------------------------------------
template <int NT, int WAVES, int MT, int KS>
__global__ __launch_bounds__(WAVES * 64) void mttkrp_bf16_mfma(const uint16_t* __restrict__ X, KernArgs a,
                                                               float* __restrict__ partial) {
  constexpr int KB = 32 * KS, PITCH = KB + 8, RT = 16 * NT, NTH = WAVES * 64, ROWS = WAVES * 16 * MT;
  __shared__ __attribute__((aligned(16))) uint16_t sKR[2][RT * PITCH];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int g = lane >> 4, c = lane & 15;
  const int split = blockIdx.y;
  const i64 b_begin = (i64)split * a.nblocks / a.splits;
  const i64 b_end = (i64)(split + 1) * a.nblocks / a.splits;
  const i64 row_base = (i64)blockIdx.x * ROWS + wave * 16 * MT;

  i64 rowc[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    i64 rr = row_base + mt * 16 + c;
    rowc[mt] = rr < a.I ? rr : a.I - 1;
  }
  floatx4 acc[MT][NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = floatx4{0.f, 0.f, 0.f, 0.f};

  auto loadx = [&](i64 b, bf16x8 (&xa)[MT][KS]) {
    const i64 j0 = b * KB;
    const i64 p = j0 / a.Q, q0 = j0 - p * a.Q;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const uint16_t* base = X + (p * a.I + rowc[mt]) * a.Q + q0 + 8 * g;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) xa[mt][ks] = *reinterpret_cast<const bf16x8*>(base + 32 * ks);
    }
  };

  bf16x8 xa[MT][KS], xn[MT][KS];
  if (b_begin < b_end) {
    gen_kr_block<RT, KB, PITCH, NTH>(sKR[0], a, b_begin * KB, tid);
    loadx(b_begin, xa);
  }
  __syncthreads();
  for (i64 b = b_begin; b < b_end; ++b) {
    const int buf = (int)((b - b_begin) & 1);
    if (b + 1 < b_end) {
      loadx(b + 1, xn);
      gen_kr_block<RT, KB, PITCH, NTH>(sKR[buf ^ 1], a, (b + 1) * KB, tid);
    }
    const uint16_t* s = sKR[buf];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const bf16x8 bv = *reinterpret_cast<const bf16x8*>(&s[(16 * nt + c) * PITCH + 32 * ks + 8 * g]);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa[mt][ks], bv, acc[mt][nt], 0, 0, 0);
      }
    __syncthreads();
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) xa[mt][ks] = xn[mt][ks];
  }

  float* out = partial + (i64)split * a.I * a.R;
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int col = a.r0 + 16 * nt + c;
      if (16 * nt + c >= a.rcount) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const i64 row = row_base + mt * 16 + 4 * g + v;
        if (row < a.I) out[row * a.R + col] = acc[mt][nt][v];
      }
    }
}

// ---------------------------------------------------------------------------
// Generic fallback for shapes the MFMA tiling does not cover (fastest other mode
// not a multiple of the block, rank > 64 corner cases are handled by chunks).
// One thread per (i, r); grid.y splits j.  FP32 accumulate like the MFMA path.
// ---------------------------------------------------------------------------
struct GenericArgs {
  int nd, mode, n_oth;
  i64 I, Q, R, J;
  i64 oth_size[MAXD];
  const float* oth_fac[MAXD];
  int splits;
  int bf16;
};

__global__ __launch_bounds__(256) void mttkrp_generic(const void* __restrict__ Xv, GenericArgs a,
                                                      float* __restrict__ partial) {
  const i64 idx = (i64)blockIdx.x * 256 + threadIdx.x;
  if (idx >= a.I * a.R) return;
  const i64 i = idx / a.R, r = idx - (idx / a.R) * a.R;
  const int split = blockIdx.y;
  const i64 j_begin = (i64)split * a.J / a.splits, j_end = (i64)(split + 1) * a.J / a.splits;
  float acc = 0.f;
  for (i64 j = j_begin; j < j_end; ++j) {
    i64 rem = j;
    float kr = 1.f;
    for (int m = a.n_oth - 1; m >= 0; --m) {
      const i64 sz = a.oth_size[m];
      const i64 ix = rem % sz;
      rem /= sz;
      kr *= a.oth_fac[m][ix * a.R + r];
    }
    const i64 p = j / a.Q, q = j - p * a.Q;
    const i64 off = (p * a.I + i) * a.Q + q;
    float x;
    if (a.bf16) {
      const uint16_t h = reinterpret_cast<const uint16_t*>(Xv)[off];
      x = __uint_as_float((uint32_t)h << 16);
    } else {
      x = reinterpret_cast<const float*>(Xv)[off];
    }
    acc = fmaf(x, kr, acc);
  }
  partial[(i64)split * a.I * a.R + idx] = acc;
}

// ---------------------------------------------------------------------------
// Reduction of split partials (fixed order, FP64) and conversions.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void reduce_partials_kernel(const float* __restrict__ part, int splits, i64 n,
                                                              float* __restrict__ o32, double* __restrict__ o64) {
  const i64 idx = (i64)blockIdx.x * 256 + threadIdx.x;
  if (idx >= n) return;
  double s = 0.0;
  for (int sp = 0; sp < splits; ++sp) s += (double)part[(i64)sp * n + idx];
  if (o32) o32[idx] = (float)s;
  if (o64) o64[idx] = s;
}

__global__ __launch_bounds__(256) void f32_to_bf16_kernel(const float* __restrict__ src, i64 n,
                                                          uint16_t* __restrict__ dst) {
  const i64 i4 = ((i64)blockIdx.x * 256 + threadIdx.x) * 4;
  if (i4 + 3 < n) {
    const floatx4 v = *reinterpret_cast<const floatx4*>(src + i4);
    ushort4 o;
    o.x = f32_to_bf16_rne(v[0]);
    o.y = f32_to_bf16_rne(v[1]);
    o.z = f32_to_bf16_rne(v[2]);
    o.w = f32_to_bf16_rne(v[3]);
    *reinterpret_cast<ushort4*>(dst + i4) = o;
  } else {
    for (i64 i = i4; i < n; ++i) dst[i] = f32_to_bf16_rne(src[i]);
  }
}

// [rows, K] -> [K, rows] transpose through a 32x32 LDS tile, writing bf16
// (from fp32 with RNE rounding, or a plain copy of bf16 input).
__device__ __forceinline__ uint16_t to_bf16(float v) { return f32_to_bf16_rne(v); }
__device__ __forceinline__ uint16_t to_bf16(uint16_t v) { return v; }

template <typename TIN>
__global__ __launch_bounds__(256) void transpose_bf16_kernel(const TIN* __restrict__ src, i64 rows, i64 K,
                                                             uint16_t* __restrict__ dst) {
  __shared__ TIN tile[32][33];
  const i64 r0 = (i64)blockIdx.y * 32, k0 = (i64)blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int yy = ty; yy < 32; yy += 8) {
    const i64 r = r0 + yy, k = k0 + tx;
    tile[yy][tx] = (r < rows && k < K) ? src[r * K + k] : (TIN)0;
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    const i64 k = k0 + yy, r = r0 + tx;
    if (k < K && r < rows) dst[k * rows + r] = to_bf16(tile[tx][yy]);
  }
}

// Forward-order Khatri-Rao of n factors: row = (i_0, ..., i_{n-1}) last fastest.
struct KrManyArgs {
  int n;
  i64 R, rows;
  i64 sizes[MAXD];
  const float* F[MAXD];
};
__global__ __launch_bounds__(256) void kr_many_kernel(KrManyArgs a, float* __restrict__ dst) {
  const i64 idx = (i64)blockIdx.x * 256 + threadIdx.x;
  if (idx >= a.rows * a.R) return;
  const i64 row = idx / a.R, r = idx - (idx / a.R) * a.R;
  i64 rem = row;
  float v = 1.f;
  float vals[MAXD];
  for (int m = a.n - 1; m >= 0; --m) {
    const i64 ix = rem % a.sizes[m];
    rem /= a.sizes[m];
    vals[m] = a.F[m][ix * a.R + r];
  }
  for (int m = 0; m < a.n; ++m) v = v * vals[m];  // left-to-right, as khatri_rao_except_mode
  dst[idx] = v;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static inline unsigned cdiv(i64 a, i64 b) { return (unsigned)((a + b - 1) / b); }

int reduce_partials(tenrso_ctx* ctx, const float* partial, int splits, i64 n, float* out32, double* out64) {
  if (n == 0) return TENRSO_OK;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(cdiv(n, 256)), dim3(256), 0, ctx->stream, partial, splits, n,
                     out32, out64);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

int khatri_rao_many(tenrso_ctx* ctx, int n, const float* const* F, const i64* sizes, i64 R, float* dst) {
  KrManyArgs a{};
  a.n = n;
  a.R = R;
  a.rows = 1;
  for (int m = 0; m < n; ++m) {
    a.sizes[m] = sizes[m];
    a.F[m] = F[m];
    a.rows *= sizes[m];
  }
  hipLaunchKernelGGL(kr_many_kernel, dim3(cdiv(a.rows * R, 256)), dim3(256), 0, ctx->stream, a, dst);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

int convert_bf16(tenrso_ctx* ctx, const float* src, int nd, const i64* shape, bool last_to_front, uint16_t* dst) {
  i64 total = 1;
  for (int k = 0; k < nd; ++k) total *= shape[k];
  if (total == 0) return TENRSO_OK;
  if (!last_to_front || nd < 2) {
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(cdiv(cdiv(total, 4), 256)), dim3(256), 0, ctx->stream, src, total,
                       dst);
  } else {
    const i64 K = shape[nd - 1], rows = total / K;
    hipLaunchKernelGGL(transpose_bf16_kernel<float>, dim3(cdiv(K, 32), cdiv(rows, 32)), dim3(256), 0, ctx->stream,
                       src, rows, K, dst);
  }
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

int transpose_last_to_front_u16(tenrso_ctx* ctx, const uint16_t* src, int nd, const i64* shape, uint16_t* dst) {
  i64 total = 1;
  for (int k = 0; k < nd; ++k) total *= shape[k];
  if (total == 0) return TENRSO_OK;
  const i64 K = shape[nd - 1], rows = total / K;
  hipLaunchKernelGGL(transpose_bf16_kernel<uint16_t>, dim3(cdiv(K, 32), cdiv(rows, 32)), dim3(256), 0, ctx->stream,
                     src, rows, K, dst);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

namespace {

struct Plan {
  i64 P = 1, I = 1, Q = 1, J = 1;
  int n_oth = 0;
  int oth[MAXD];
  i64 oth_size[MAXD];
  i64 fast_size = 1;
};

Plan make_plan(int nd, const i64* shape, int mode) {
  Plan p;
  for (int k = 0; k < mode; ++k) p.P *= shape[k];
  for (int k = mode + 1; k < nd; ++k) p.Q *= shape[k];
  p.I = shape[mode];
  p.J = p.P * p.Q;
  for (int k = 0; k < nd; ++k)
    if (k != mode) {
      p.oth[p.n_oth] = k;
      p.oth_size[p.n_oth] = shape[k];
      ++p.n_oth;
    }
  p.fast_size = p.n_oth ? p.oth_size[p.n_oth - 1] : 1;
  return p;
}

}  // namespace

// Splits so that the grid fills the chip while the partial volume stays small:
// about 2 workgroups per CU.
static int choose_splits(tenrso_ctx* ctx, i64 tiles, i64 nblocks) {
  const i64 target = 2LL * ctx->num_cus;
  i64 s = std::max<i64>(1, target / std::max<i64>(1, tiles));
  s = std::min<i64>(s, nblocks);
  s = std::min<i64>(s, 4096);
  return (int)std::max<i64>(1, s);
}

int mttkrp_dense(tenrso_ctx* ctx, int prec, const void* X, const void* X_t, int nd, const i64* shape,
                 const float* const* F, i64 R, int mode, MttkrpOut outk, void* out, MttkrpResult* res) {
  Plan pl = make_plan(nd, shape, mode);
  const bool bf16 = prec == TENRSO_PREC_BF16;
  const float* Fv[MAXD];
  for (int k = 0; k < nd; ++k) Fv[k] = F[k];
  const void* Xuse = X;
  int mode_use = mode;
  i64 shape_use[MAXD];
  for (int k = 0; k < nd; ++k) shape_use[k] = shape[k];

  // BF16 last mode: run mode 0 of the last-to-front copy (same j order).
  if (bf16 && mode == nd - 1 && nd >= 2) {
    if (!X_t) {
      set_error("internal: bf16 last-mode MTTKRP needs the last-to-front copy");
      return TENRSO_ERR_INVALID_ARGUMENT;
    }
    Xuse = X_t;
    mode_use = 0;
    shape_use[0] = shape[nd - 1];
    Fv[0] = F[nd - 1];
    for (int k = 1; k < nd; ++k) {
      shape_use[k] = shape[k - 1];
      Fv[k] = F[k - 1];
    }
    pl = make_plan(nd, shape_use, 0);
  }

  const i64 I = pl.I;
  const i64 nout = I * R;
  const int RT = R <= 16 ? 16 : (R <= 32 ? 32 : 64);
  const int NT = RT / 16;
  const int rchunks = (int)((R + RT - 1) / RT);

  // PRE: product of the other modes except the fastest (forward order).
  const float* pre = nullptr;
  if (pl.n_oth >= 3) {
    const float* pf[MAXD];
    i64 ps[MAXD];
    i64 prows = 1;
    for (int m = 0; m < pl.n_oth - 1; ++m) {
      pf[m] = Fv[pl.oth[m]];
      ps[m] = pl.oth_size[m];
      prows *= ps[m];
    }
    void* buf = nullptr;
    TENRSO_TRY(ctx_buffer(ctx, SLOT_TMP, (size_t)(prows * R) * sizeof(float), &buf));
    TENRSO_TRY(khatri_rao_many(ctx, pl.n_oth - 1, pf, ps, R, (float*)buf));
    pre = (const float*)buf;
  } else if (pl.n_oth == 2) {
    pre = Fv[pl.oth[0]];
  }
  const float* fastf = pl.n_oth ? Fv[pl.oth[pl.n_oth - 1]] : nullptr;

  KernArgs a{};
  a.I = I;
  a.Q = pl.Q;
  a.R = R;
  a.fast_size = pl.fast_size;
  a.pre = pre;
  a.fast = fastf;

  const bool caseB = pl.Q == 1;
  int KB = bf16 ? 64 : 32;
  if (bf16 && pl.fast_size % 64 != 0 && pl.fast_size % 32 == 0) KB = 32;
  const bool mfma_ok = pl.n_oth >= 1 && pl.J > 0 && (pl.fast_size % KB == 0) && !(bf16 && caseB) && I > 0;

  float* partial = nullptr;
  int splits = 1;
  if (mfma_ok) {
    constexpr int WAVES = 4, MT = 1;
    const i64 rows_per_wg = WAVES * 16 * MT;
    const i64 tiles = (i64)cdiv(I, rows_per_wg) * rchunks;
    a.nblocks = pl.J / KB;
    splits = choose_splits(ctx, tiles, a.nblocks);
    a.splits = splits;
    void* pbuf = nullptr;
    TENRSO_TRY(ctx_buffer(ctx, SLOT_PARTIAL, (size_t)splits * nout * sizeof(float), &pbuf));
    partial = (float*)pbuf;
    const dim3 block(WAVES * 64);
    for (int rc = 0; rc < rchunks; ++rc) {
      a.r0 = rc * RT;
      a.rcount = (int)std::min<i64>(RT, R - a.r0);
      const dim3 grid(cdiv(I, rows_per_wg), splits, 1);
      if (!bf16) {
        const float* Xf = (const float*)Xuse;
        if (!caseB) {
          if (NT == 1) hipLaunchKernelGGL((mttkrp_f32_mfma<1, WAVES, MT, false>), grid, block, 0, ctx->stream, Xf, a, partial);
          else if (NT == 2) hipLaunchKernelGGL((mttkrp_f32_mfma<2, WAVES, MT, false>), grid, block, 0, ctx->stream, Xf, a, partial);
          else hipLaunchKernelGGL((mttkrp_f32_mfma<4, WAVES, MT, false>), grid, block, 0, ctx->stream, Xf, a, partial);
        } else {
          if (NT == 1) hipLaunchKernelGGL((mttkrp_f32_mfma<1, WAVES, MT, true>), grid, block, 0, ctx->stream, Xf, a, partial);
          else if (NT == 2) hipLaunchKernelGGL((mttkrp_f32_mfma<2, WAVES, MT, true>), grid, block, 0, ctx->stream, Xf, a, partial);
          else hipLaunchKernelGGL((mttkrp_f32_mfma<4, WAVES, MT, true>), grid, block, 0, ctx->stream, Xf, a, partial);
        }
      } else {
        const uint16_t* Xb = (const uint16_t*)Xuse;
        if (KB == 64) {
          if (NT == 1) hipLaunchKernelGGL((mttkrp_bf16_mfma<1, WAVES, MT, 2>), grid, block, 0, ctx->stream, Xb, a, partial);
          else if (NT == 2) hipLaunchKernelGGL((mttkrp_bf16_mfma<2, WAVES, MT, 2>), grid, block, 0, ctx->stream, Xb, a, partial);
          else hipLaunchKernelGGL((mttkrp_bf16_mfma<4, WAVES, MT, 2>), grid, block, 0, ctx->stream, Xb, a, partial);
        } else {
          if (NT == 1) hipLaunchKernelGGL((mttkrp_bf16_mfma<1, WAVES, MT, 1>), grid, block, 0, ctx->stream, Xb, a, partial);
          else if (NT == 2) hipLaunchKernelGGL((mttkrp_bf16_mfma<2, WAVES, MT, 1>), grid, block, 0, ctx->stream, Xb, a, partial);
          else hipLaunchKernelGGL((mttkrp_bf16_mfma<4, WAVES, MT, 1>), grid, block, 0, ctx->stream, Xb, a, partial);
        }
      }
      TENRSO_HIP_TRY(hipGetLastError());
    }
  } else {
    GenericArgs ga{};
    ga.nd = nd;
    ga.mode = mode_use;
    ga.n_oth = pl.n_oth;
    ga.I = I;
    ga.Q = pl.Q;
    ga.R = R;
    ga.J = pl.J;
    for (int m = 0; m < pl.n_oth; ++m) {
      ga.oth_size[m] = pl.oth_size[m];
      ga.oth_fac[m] = Fv[pl.oth[m]];
    }
    ga.bf16 = bf16 ? 1 : 0;
    const i64 tiles = cdiv(nout, 256);
    splits = (int)std::max<i64>(1, std::min<i64>(pl.J, (2LL * ctx->num_cus + tiles - 1) / std::max<i64>(1, tiles)));
    splits = std::min(splits, 1024);
    ga.splits = splits;
    void* pbuf = nullptr;
    TENRSO_TRY(ctx_buffer(ctx, SLOT_PARTIAL, (size_t)splits * nout * sizeof(float), &pbuf));
    partial = (float*)pbuf;
    if (nout > 0) {
      hipLaunchKernelGGL(mttkrp_generic, dim3((unsigned)tiles, splits), dim3(256), 0, ctx->stream, Xuse, ga,
                         partial);
      TENRSO_HIP_TRY(hipGetLastError());
    }
  }

  if (res) {
    res->partial = partial;
    res->splits = splits;
  }
  if (outk == MttkrpOut::F32) return reduce_partials(ctx, partial, splits, nout, (float*)out, nullptr);
  if (outk == MttkrpOut::F64) return reduce_partials(ctx, partial, splits, nout, nullptr, (double*)out);
  return TENRSO_OK;
}

}  // namespace tenrso
