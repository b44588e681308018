// sparse.hip -- sparse COO tensors and sparse MTTKRP on gfx950.
//
// The reference has COO / CSF / HiCOO formats (tenrso-sparse/src/coo.rs:73-348,
// csf.rs:137-264, hicoo.rs:121-205) but no sparse MTTKRP.  The contract here is
// the dense MTTKRP (tenrso-kernels/src/mttkrp.rs:74-146) of the tensor the COO
// represents, with duplicate coordinates summed (the dense form of
// CooTensor::deduplicate, coo.rs:320-348):
//   M[i_n, r] = sum_{nonzeros e} x_e * prod_{k != n} A_k[i_k(e), r].
//
// Layout (per output mode n, built once on the device):
//   the nonzeros sorted lexicographically by (i_n, other modes ascending) -- the
//   order CsfTensor::from_coo produces for mode_order [n, others...] -- stored as
//   SoA: row (i_n, int32), the other modes' indices (int32) and values (f32),
//   plus rowptr (int64, I_n + 1) and the CSF fiber structure (standard fptr
//   semantics: fptr[l][f] indexes level l+1, unlike csf.rs:209,246).
//
// MTTKRP kernel: a worker of LPN lanes (one lane per rank column) owns a
// contiguous chunk of SP_CH sorted nonzeros and accumulates row by row in
// registers.  Rows strictly inside the chunk belong to the worker alone and are
// stored directly; the chunk's first and last rows may be shared with
// neighbouring chunks and go to two carry slots.  The carry slots (sorted by
// row) are folded by the same tile-wise segmented reduction, level by level,
// until one tile remains.  No atomics: the result is deterministic.
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <vector>

#include "sparse.hpp"

namespace tenrso {

typedef float floatx4 __attribute__((ext_vector_type(4)));

static inline unsigned cdiv(i64 a, i64 b) { return (unsigned)((a + b - 1) / b); }

// ---------------------------------------------------------------- build
struct KeyArgs {
  int nd;
  int mode;
  i64 nnz;
  i64 shape[MAXD_SP];
  const int32_t* idx;  // SoA: idx[k * nnz + e]
};

// key = (i_n, i_o1, i_o2, ...) linearised, others in ascending mode order
__global__ __launch_bounds__(256) void make_keys_kernel(KeyArgs a, unsigned long long* __restrict__ keys,
                                                        int32_t* __restrict__ pos) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e >= a.nnz) return;
  unsigned long long k = (unsigned long long)a.idx[(i64)a.mode * a.nnz + e];
  for (int m = 0; m < a.nd; ++m) {
    if (m == a.mode) continue;
    k = k * (unsigned long long)a.shape[m] + (unsigned long long)a.idx[(i64)m * a.nnz + e];
  }
  keys[e] = k;
  pos[e] = (int32_t)e;
}

struct GatherArgs {
  int nd, mode;
  i64 nnz;
  const int32_t* idx;
  const float* vals;
  const int32_t* perm;
  int32_t* row;
  int32_t* oth;  // (nd-1) x nnz
  float* vout;
};

__global__ __launch_bounds__(256) void gather_sorted_kernel(GatherArgs a) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e >= a.nnz) return;
  const i64 src = a.perm[e];
  a.row[e] = a.idx[(i64)a.mode * a.nnz + src];
  int o = 0;
  for (int m = 0; m < a.nd; ++m) {
    if (m == a.mode) continue;
    a.oth[(i64)o * a.nnz + e] = a.idx[(i64)m * a.nnz + src];
    ++o;
  }
  a.vout[e] = a.vals[src];
}

// rowptr[i] = first position with row >= i (binary search over the sorted rows)
__global__ __launch_bounds__(256) void rowptr_kernel(const int32_t* __restrict__ row, i64 nnz, i64 rows,
                                                     int64_t* __restrict__ rowptr) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i > rows) return;
  i64 lo = 0, hi = nnz;
  while (lo < hi) {
    const i64 mid = (lo + hi) >> 1;
    if ((i64)row[mid] < i) lo = mid + 1;
    else hi = mid;
  }
  rowptr[i] = lo;
}

__global__ __launch_bounds__(256) void bounds_kernel(KeyArgs a, int* __restrict__ bad) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e >= a.nnz) return;
  for (int m = 0; m < a.nd; ++m) {
    const int32_t v = a.idx[(i64)m * a.nnz + e];
    if (v < 0 || (i64)v >= a.shape[m]) bad[0] = 1;
  }
}

// ---- compressed fibres (see SpMode) ----------------------------------------
// flag[e] = 1 where the prefix (row, oth[0 .. D-2]) of sorted nonzero e differs
// from that of e - 1 (e = 0 always starts a fibre)
__global__ __launch_bounds__(256) void fib_flag_kernel(const int32_t* __restrict__ row,
                                                       const int32_t* __restrict__ oth, i64 nnz, int D,
                                                       int32_t* __restrict__ flag) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e >= nnz) return;
  int f = e == 0 || row[e] != row[e - 1];
  for (int l = 0; l + 1 < D && !f; ++l) f = oth[(i64)l * nnz + e] != oth[(i64)l * nnz + e - 1];
  flag[e] = f;
}

// pos = exclusive scan of flag: fibre f = pos[e] starts at e where flag[e].
// Scatters the prefixes and writes the leaves with their fibre-end bit.
__global__ __launch_bounds__(256) void fib_scatter_kernel(const int32_t* __restrict__ flag,
                                                          const int32_t* __restrict__ pos,
                                                          const int32_t* __restrict__ row,
                                                          const int32_t* __restrict__ oth, i64 nnz, int D, i64 nfib,
                                                          int32_t* __restrict__ leafe, int32_t* __restrict__ fpre) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e >= nnz) return;
  if (flag[e]) {
    const i64 f = pos[e];
    fpre[f] = row[e];
    for (int l = 1; l < D; ++l) fpre[(i64)l * nfib + f] = oth[(i64)(l - 1) * nnz + e];
  }
  const bool last = e == nnz - 1 || flag[e + 1];
  leafe[e] = (int32_t)((uint32_t)oth[(i64)(D - 1) * nnz + e] | (last ? 0x80000000u : 0u));
}

// chunk_fib[c] = the fibre holding nonzero c * W (inclusive scan - 1)
__global__ __launch_bounds__(256) void chunk_fib_kernel(const int32_t* __restrict__ flag,
                                                        const int32_t* __restrict__ pos, i64 nchunks, int W,
                                                        int32_t* __restrict__ cf) {
  const i64 c = (i64)blockIdx.x * 256 + threadIdx.x;
  if (c >= nchunks) return;
  const i64 e = c * W;
  cf[c] = pos[e] + flag[e] - 1;
}

// Device temporaries of one build step, freed on every path out of it.
struct DevTemps {
  std::vector<void*> ptrs;
  int alloc(void** p, size_t bytes) {
    *p = nullptr;
    TENRSO_HIP_TRY(hipMalloc(p, bytes));
    ptrs.push_back(*p);
    return TENRSO_OK;
  }
  ~DevTemps() {
    for (void* p : ptrs) (void)hipFree(p);
  }
};

// The compressed fibres of one mode from its sorted nonzeros (flags, an
// exclusive scan, a scatter of the fibre starts and prefixes, the chunk table).
static int sp_build_fibres(tenrso_ctx* ctx, tenrso_sptensor* t, int mode) {
  SpMode& md = t->modes[mode];
  const i64 nnz = t->nnz;
  const int D = t->nd - 1;
  DevTemps tmps;
  int32_t *flag, *pos;
  TENRSO_TRY(tmps.alloc((void**)&flag, (size_t)nnz * 4));
  TENRSO_TRY(tmps.alloc((void**)&pos, (size_t)nnz * 4));
  hipLaunchKernelGGL(fib_flag_kernel, dim3(cdiv(nnz, 256)), dim3(256), 0, ctx->stream, md.row, md.oth, nnz, D, flag);
  TENRSO_HIP_TRY(hipGetLastError());
  size_t tmp_bytes = 0;
  TENRSO_HIP_TRY(rocprim::exclusive_scan(nullptr, tmp_bytes, flag, pos, 0, (size_t)nnz, rocprim::plus<int32_t>(),
                                         ctx->stream));
  void* tmp = nullptr;
  TENRSO_TRY(tmps.alloc(&tmp, std::max<size_t>(tmp_bytes, 16)));
  TENRSO_HIP_TRY(rocprim::exclusive_scan(tmp, tmp_bytes, flag, pos, 0, (size_t)nnz, rocprim::plus<int32_t>(),
                                         ctx->stream));
  int32_t last[2] = {0, 0};
  TENRSO_HIP_TRY(hipMemcpyAsync(&last[0], pos + nnz - 1, 4, hipMemcpyDeviceToHost, ctx->stream));
  TENRSO_HIP_TRY(hipMemcpyAsync(&last[1], flag + nnz - 1, 4, hipMemcpyDeviceToHost, ctx->stream));
  TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
  md.nfib = (i64)last[0] + last[1];
  md.nchunks = (nnz + SP_CSF_W - 1) / SP_CSF_W;
  TENRSO_HIP_TRY(hipMalloc(&md.leafe, (size_t)nnz * 4));
  TENRSO_HIP_TRY(hipMalloc(&md.fpre, (size_t)std::max<i64>(1, md.nfib * D) * 4));
  TENRSO_HIP_TRY(hipMalloc(&md.chunk_fib, (size_t)md.nchunks * 4));
  hipLaunchKernelGGL(fib_scatter_kernel, dim3(cdiv(nnz, 256)), dim3(256), 0, ctx->stream, flag, pos, md.row, md.oth,
                     nnz, D, md.nfib, md.leafe, md.fpre);
  hipLaunchKernelGGL(chunk_fib_kernel, dim3(cdiv(md.nchunks, 256)), dim3(256), 0, ctx->stream, flag, pos, md.nchunks,
                     SP_CSF_W, md.chunk_fib);
  TENRSO_HIP_TRY(hipGetLastError());
  // the temporaries are freed below: the kernels reading them must be done
  TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return TENRSO_OK;
}

int sp_build_mode(tenrso_ctx* ctx, tenrso_sptensor* t, int mode) {
  SpMode& md = t->modes[mode];
  const i64 nnz = t->nnz;
  md.rows = t->shape[mode];
  TENRSO_HIP_TRY(hipMalloc(&md.rowptr, (size_t)(md.rows + 1) * sizeof(int64_t)));
  TENRSO_HIP_TRY(hipMalloc(&md.row, (size_t)std::max<i64>(1, nnz) * sizeof(int32_t)));
  TENRSO_HIP_TRY(hipMalloc(&md.oth, (size_t)std::max<i64>(1, nnz * (t->nd - 1)) * sizeof(int32_t)));
  TENRSO_HIP_TRY(hipMalloc(&md.vals, (size_t)std::max<i64>(1, nnz) * sizeof(float)));
  if (nnz > 0) {
    DevTemps tmps;  // freed when this block ends, also on an early error return
    unsigned long long *kin, *kout;
    int32_t *pin, *pout;
    TENRSO_TRY(tmps.alloc((void**)&kin, (size_t)nnz * 8));
    TENRSO_TRY(tmps.alloc((void**)&kout, (size_t)nnz * 8));
    TENRSO_TRY(tmps.alloc((void**)&pin, (size_t)nnz * 4));
    TENRSO_TRY(tmps.alloc((void**)&pout, (size_t)nnz * 4));
    KeyArgs ka{};
    ka.nd = t->nd;
    ka.mode = mode;
    ka.nnz = nnz;
    for (int k = 0; k < t->nd; ++k) ka.shape[k] = t->shape[k];
    ka.idx = t->idx;
    hipLaunchKernelGGL(make_keys_kernel, dim3(cdiv(nnz, 256)), dim3(256), 0, ctx->stream, ka, kin, pin);
    TENRSO_HIP_TRY(hipGetLastError());
    int bits = 0;
    {
      unsigned long long maxkey = 1;
      for (int k = 0; k < t->nd; ++k) maxkey *= (unsigned long long)t->shape[k];
      while (bits < 64 && (maxkey >> bits) != 0) ++bits;
    }
    size_t tmp_bytes = 0;
    TENRSO_HIP_TRY(rocprim::radix_sort_pairs(nullptr, tmp_bytes, kin, kout, pin, pout, (size_t)nnz, 0, bits,
                                             ctx->stream));
    void* tmp = nullptr;
    TENRSO_TRY(tmps.alloc(&tmp, std::max<size_t>(tmp_bytes, 16)));
    TENRSO_HIP_TRY(rocprim::radix_sort_pairs(tmp, tmp_bytes, kin, kout, pin, pout, (size_t)nnz, 0, bits,
                                             ctx->stream));
    GatherArgs ga{t->nd, mode, nnz, t->idx, t->vals, pout, md.row, md.oth, md.vals};
    hipLaunchKernelGGL(gather_sorted_kernel, dim3(cdiv(nnz, 256)), dim3(256), 0, ctx->stream, ga);
    TENRSO_HIP_TRY(hipGetLastError());
    // the temporaries are freed below: the kernels reading them must be done
    TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  hipLaunchKernelGGL(rowptr_kernel, dim3(cdiv(md.rows + 1, 256)), dim3(256), 0, ctx->stream, md.row, nnz, md.rows,
                     md.rowptr);
  TENRSO_HIP_TRY(hipGetLastError());
  if (nnz > 0 && t->nd >= 2) TENRSO_TRY(sp_build_fibres(ctx, t, mode));
  return TENRSO_OK;
}

// Index bounds of CooTensor::new (coo.rs:127-139) on the device.  Fails closed:
// any HIP error is reported as a device error, never as "in bounds".  The flag is
// cleared by a fill kernel, not a memset.
__global__ void clear_flag_kernel(int* __restrict__ f) { f[0] = 0; }

static int check_index_bounds(tenrso_ctx* ctx, int nd, const i64* shape, i64 nnz, const int32_t* idx) {
  DevTemps tmps;
  int* bad = nullptr;
  TENRSO_TRY(tmps.alloc((void**)&bad, sizeof(int)));
  hipLaunchKernelGGL(clear_flag_kernel, dim3(1), dim3(1), 0, ctx->stream, bad);
  TENRSO_HIP_TRY(hipGetLastError());
  KeyArgs ka{};
  ka.nd = nd;
  ka.nnz = nnz;
  for (int k = 0; k < nd; ++k) ka.shape[k] = shape[k];
  ka.idx = idx;
  hipLaunchKernelGGL(bounds_kernel, dim3(cdiv(nnz, 256)), dim3(256), 0, ctx->stream, ka, bad);
  TENRSO_HIP_TRY(hipGetLastError());
  int hbad = 1;
  TENRSO_HIP_TRY(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (hbad) {
    set_error("Index out of bounds");
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  return TENRSO_OK;
}

int sp_create(tenrso_ctx* ctx, int nd, const i64* shape, i64 nnz, const int32_t* idx, const float* vals,
              bool host, tenrso_sptensor** out) {
  *out = nullptr;
  if (nd < 1 || nd > MAXD_SP) {
    set_error("sparse tensors need 1..%d modes", MAXD_SP);
    return TENRSO_ERR_INVALID_ARGUMENT;
  }
  // CooTensor::new validation (coo.rs:101-149)
  for (int k = 0; k < nd; ++k)
    if (shape[k] <= 0) {
      set_error("Invalid shape: Shape cannot contain zeros");
      return TENRSO_ERR_INVALID_ARGUMENT;
    }
  if (nnz < 0 || nnz > (i64)INT32_MAX) {
    set_error("nnz %lld out of the supported range", (long long)nnz);
    return TENRSO_ERR_UNSUPPORTED;
  }
  {
    long double prod = 1.0L;
    for (int k = 0; k < nd; ++k) prod *= (long double)shape[k];
    if (prod >= 18446744073709551615.0L) {
      set_error("tensor too large for 64-bit sort keys");
      return TENRSO_ERR_UNSUPPORTED;
    }
    for (int k = 0; k < nd; ++k)
      if (shape[k] > (i64)INT32_MAX) {
        set_error("mode size exceeds the int32 index range");
        return TENRSO_ERR_UNSUPPORTED;
      }
  }
  auto* t = new tenrso_sptensor();
  t->nd = nd;
  t->nnz = nnz;
  for (int k = 0; k < nd; ++k) t->shape[k] = shape[k];
  int rc = TENRSO_OK;
  auto fail = [&](int code) {
    sp_destroy(t);
    return code;
  };
  if (hipMalloc(&t->idx, (size_t)std::max<i64>(1, nnz * nd) * 4) != hipSuccess ||
      hipMalloc(&t->vals, (size_t)std::max<i64>(1, nnz) * 4) != hipSuccess) {
    set_error("device allocation failed for the sparse tensor");
    return fail(TENRSO_ERR_DEVICE);
  }
  const hipMemcpyKind kind = host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  if (nnz > 0) {
    if (hipMemcpyAsync(t->idx, idx, (size_t)nnz * nd * 4, kind, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(t->vals, vals, (size_t)nnz * 4, kind, ctx->stream) != hipSuccess) {
      set_error("copy of the COO arrays failed");
      return fail(TENRSO_ERR_DEVICE);
    }
    // index bounds (CooTensor::new, coo.rs:127-139)
    rc = check_index_bounds(ctx, nd, shape, nnz, t->idx);
    if (rc != TENRSO_OK) return fail(rc);
  }
  for (int m = 0; m < nd; ++m) {
    rc = sp_build_mode(ctx, t, m);
    if (rc != TENRSO_OK) return fail(rc);
  }
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
    set_error("sparse build failed");
    return fail(TENRSO_ERR_DEVICE);
  }
  *out = t;
  return TENRSO_OK;
}

void sp_destroy(tenrso_sptensor* t) {
  if (!t) return;
  for (int m = 0; m < MAXD_SP; ++m) {
    SpMode& md = t->modes[m];
    if (md.rowptr) (void)hipFree(md.rowptr);
    if (md.row) (void)hipFree(md.row);
    if (md.oth) (void)hipFree(md.oth);
    if (md.vals) (void)hipFree(md.vals);
    if (md.leafe) (void)hipFree(md.leafe);
    if (md.fpre) (void)hipFree(md.fpre);
    if (md.chunk_fib) (void)hipFree(md.chunk_fib);
  }
  if (t->idx) (void)hipFree(t->idx);
  if (t->vals) (void)hipFree(t->vals);
  delete t;
}

// ---------------------------------------------------------------- MTTKRP
constexpr int SP_CH = 1024;  // nonzeros per worker chunk
constexpr int SP_U = 4;      // nonzeros in flight per worker step

struct SpArgs {
  int nd_oth;
  i64 nnz, rows, R;
  int r0, rcount;
  const int32_t* row;
  const int32_t* oth;
  const float* vals;
  const float* F[MAXD_SP];  // factors of the other modes, ascending mode order
  i64 nworkers;
};

template <int LPN>
__global__ __launch_bounds__(256) void sp_mttkrp_kernel(SpArgs a, float* __restrict__ out,
                                                        float* __restrict__ cval, int32_t* __restrict__ crow) {
  const i64 gt = (i64)blockIdx.x * 256 + threadIdx.x;
  const i64 w = gt / LPN;
  const int lane = (int)(gt - w * LPN);
  if (w >= a.nworkers) return;
  const bool live = lane < a.rcount;
  const int col = a.r0 + (live ? lane : 0);
  const i64 e0 = w * SP_CH;
  const i64 e1 = (e0 + SP_CH) < a.nnz ? e0 + SP_CH : a.nnz;
  float* c0 = cval + (2 * w) * a.R;
  float* c1 = cval + (2 * w + 1) * a.R;
  const int first = a.row[e0];
  int cur = first;
  float acc = 0.f;
  for (i64 e = e0; e < e1; e += SP_U) {
    int rr[SP_U];
    float v[SP_U];
#pragma unroll
    for (int u = 0; u < SP_U; ++u) {
      const i64 ee = e + u;
      const bool ok = ee < e1;
      const i64 es = ok ? ee : e0;
      rr[u] = ok ? a.row[es] : cur;
      float p = ok ? a.vals[es] : 0.f;
      for (int m = 0; m < a.nd_oth; ++m) {
        const i64 ix = a.oth[(i64)m * a.nnz + es];
        p *= a.F[m][ix * a.R + col];
      }
      v[u] = p;
    }
#pragma unroll
    for (int u = 0; u < SP_U; ++u) {
      if (e + u >= e1) break;  // past the chunk's end: must not switch rows
      if (rr[u] != cur) {
        // flush row `cur`: the chunk's first row goes to carry 0, later rows are
        // interior (the row changed again inside this chunk) and stored directly
        if (cur == first) {
          if (live) c0[col] = acc;
        } else if (live) {
          out[(i64)cur * a.R + col] = acc;
        }
        cur = rr[u];
        acc = 0.f;
      }
      acc += v[u];
    }
  }
  // The chunk's last row is a carry.  A single-row chunk emits (first, acc) and
  // (first, 0), so the carry keys stay sorted with no gaps.
  if (cur == first) {
    if (live) {
      c0[col] = acc;
      c1[col] = 0.f;
    }
  } else if (live) {
    c1[col] = acc;
  }
  if (lane == 0) {
    crow[2 * w] = first;
    crow[2 * w + 1] = cur;
  }
}

// Vector form for R % 4 == 0: a worker is G lanes (lane q owns rank columns
// r0 + 4q .. 4q+3 and gathers factor rows as 16-byte vectors), a wave holds
// 64 / G workers, and each worker runs SP_W consecutive sorted nonzeros.  The
// worker's row / value / index runs are staged SP_SUB at a time into
// wave-private LDS (dynamic, (nd_oth + 2) arrays of SP_SUB words per worker)
// with vector loads; the nonzeros are then processed SP_VU at a time with
// every factor gather issued before any is used.  Same carry contract as
// sp_mttkrp_kernel: the worker's first row goes to carry slot 2w, its last row
// to slot 2w + 1, rows strictly inside the run are stored to `out` directly.
constexpr int SP_W = 128;    // nonzeros per vector worker
constexpr int SP_SUB = 32;   // nonzeros staged per pass
constexpr int SP_VU = 4;     // nonzeros in flight per worker (a multiple of 4: staged entries are read as 16-byte vectors)

typedef int32_t intx4 __attribute__((ext_vector_type(4)));
typedef int32_t intx2 __attribute__((ext_vector_type(2)));

template <int G>
static constexpr size_t sp_vec_lds_bytes(int nd_oth) {
  return (size_t)4 * (64 / G) * (nd_oth + 2) * SP_SUB * sizeof(int32_t);
}

// Copy entries [j0, j0 + PER) of a staged run: vector loads when the run is
// full and 16-byte aligned (an index array of a mode > 0 starts at m * nnz
// words), element-wise (zero past n) otherwise.
template <int PER>
__device__ __forceinline__ void sp_stage_part(const int32_t* __restrict__ src, int32_t* dst, int j0, int n) {
  if (n == SP_SUB && (reinterpret_cast<uintptr_t>(src) & 15) == 0) {
    if constexpr (PER == 2) {
      *reinterpret_cast<intx2*>(dst + j0) = *reinterpret_cast<const intx2*>(src + j0);
    } else {
#pragma unroll
      for (int k = 0; k < PER; k += 4)
        *reinterpret_cast<intx4*>(dst + j0 + k) = *reinterpret_cast<const intx4*>(src + j0 + k);
    }
  } else {
#pragma unroll
    for (int k = 0; k < PER; ++k) dst[j0 + k] = (j0 + k < n) ? src[j0 + k] : 0;
  }
}

template <int G>
__global__ __launch_bounds__(256) void sp_mttkrp_vec_kernel(SpArgs a, float* __restrict__ out,
                                                            float* __restrict__ cval, int32_t* __restrict__ crow) {
  constexpr int NPW = 64 / G;      // workers per wave
  constexpr int PER = SP_SUB / G;  // staged entries per lane and array
  static_assert(PER == 2 || PER % 4 == 0, "staging width");
  extern __shared__ int32_t sp_stage[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane / G, q = lane % G;
  const int na = a.nd_oth + 2;  // row, value, the other modes' indices
  int32_t* st = sp_stage + ((wave * NPW + g) * na) * SP_SUB;
  const i64 w = ((i64)blockIdx.x * 4 + wave) * NPW + g;
  const bool active = w < a.nworkers;
  const i64 e0 = w * SP_W;
  const i64 e1 = active ? ((e0 + SP_W) < a.nnz ? e0 + SP_W : a.nnz) : e0;
  const bool live = 4 * q < a.rcount;
  const int col = a.r0 + (live ? 4 * q : 0);
  const int first = active ? a.row[e0] : 0;
  int cur = first;
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  auto flush = [&](int row, const floatx4& v) {
    if (!live) return;
    float* dst = (row == first) ? cval + (2 * w) * a.R + col : out + (i64)row * a.R + col;
    *reinterpret_cast<floatx4*>(dst) = v;
  };
  for (i64 base = e0; base < e1; base += SP_SUB) {
    const int n = (int)((e1 - base) < SP_SUB ? (e1 - base) : SP_SUB);
    sp_stage_part<PER>(a.row + base, st, q * PER, n);
    sp_stage_part<PER>(reinterpret_cast<const int32_t*>(a.vals) + base, st + SP_SUB, q * PER, n);
    for (int m = 0; m < a.nd_oth; ++m) sp_stage_part<PER>(a.oth + (i64)m * a.nnz + base, st + (2 + m) * SP_SUB, q * PER, n);
    // LDS traffic of one wave is processed in order: the group's reads below
    // see its writes above; the barrier only keeps the compiler from moving them
    __builtin_amdgcn_wave_barrier();
    for (int i0 = 0; i0 < n; i0 += SP_VU) {
      int rr[SP_VU];
      floatx4 pv[SP_VU];
#pragma unroll
      for (int u = 0; u < SP_VU; u += 4) {
        const intx4 r4 = *reinterpret_cast<const intx4*>(st + i0 + u);
        const intx4 v4 = *reinterpret_cast<const intx4*>(st + SP_SUB + i0 + u);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          rr[u + k] = r4[k];
          const float v = __int_as_float(v4[k]);
          pv[u + k] = floatx4{v, v, v, v};
        }
      }
      for (int m = 0; m < a.nd_oth; ++m) {
        const float* Fm = a.F[m] + col;
        floatx4 fv[SP_VU];
#pragma unroll
        for (int u = 0; u < SP_VU; u += 4) {
          const intx4 x4 = *reinterpret_cast<const intx4*>(st + (2 + m) * SP_SUB + i0 + u);
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            // staged entries past n are zero (row 0 is a valid factor row)
            fv[u + k] = *reinterpret_cast<const floatx4*>(Fm + (i64)x4[k] * a.R);
          }
        }
#pragma unroll
        for (int u = 0; u < SP_VU; ++u) pv[u] *= fv[u];
      }
#pragma unroll
      for (int u = 0; u < SP_VU; ++u) {
        if (i0 + u >= n) break;  // past the run's end: must not switch rows
        if (rr[u] != cur) {
          flush(cur, acc);
          cur = rr[u];
          acc = floatx4{0.f, 0.f, 0.f, 0.f};
        }
        acc += pv[u];
      }
    }
    __builtin_amdgcn_wave_barrier();  // staged entries read before the next pass overwrites them
  }
  if (!active) return;
  // the worker's last row is a carry; a single-row worker emits (first, acc) and (first, 0)
  if (live) {
    float* c0 = cval + (2 * w) * a.R + col;
    float* c1 = cval + (2 * w + 1) * a.R + col;
    if (cur == first) {
      *reinterpret_cast<floatx4*>(c0) = acc;
      *reinterpret_cast<floatx4*>(c1) = floatx4{0.f, 0.f, 0.f, 0.f};
    } else {
      *reinterpret_cast<floatx4*>(c1) = acc;
    }
  }
  if (q == 0) {
    crow[2 * w] = first;
    crow[2 * w + 1] = cur;
  }
}

// Fibre form (R % 4 == 0, 16-byte aligned factors, 2 <= nd <= 4): the same
// worker / carry contract as sp_mttkrp_vec_kernel over the compressed fibres.
// Per nonzero: the leaf (last other mode's index with the fibre-end bit) and
// value, one factor-row gather, and v * F_last[leaf] added to the fibre's sum;
// per fibre, at its last nonzero (or the run's last): the sum times the factor
// rows of the prefix's middle indices (D - 1 gathers) added to the row.  So the
// middle factors are gathered once per fibre instead of once per nonzero, and
// the row / middle indices are read once per fibre.  A worker's fibre prefixes
// are staged SP_FWIN at a time into LDS when the batch's fibres pass the window;
// the fibre of each nonzero is the chunk's first fibre plus the ends before it.
// Window: 64 fibres per worker, 32 where a wave holds 16 workers (LDS).
template <int G>
static constexpr int sp_fwin() {
  return G >= 8 ? 64 : 32;
}

template <int G, int D>
static constexpr size_t sp_csf_lds_bytes() {
  return (size_t)4 * (64 / G) * (2 * SP_SUB + D * sp_fwin<G>()) * sizeof(int32_t);
}

struct SpCsfArgs {
  i64 nnz, rows, R, nfib;
  int r0, rcount;
  const int32_t* leafe;    // leaf index | fibre-end bit, per sorted nonzero
  const float* vals;
  const int32_t* fpre;     // D x nfib
  const int32_t* chunk_fib;
  const float* F[MAXD_SP]; // factors of the other modes, ascending mode order
  i64 nworkers;
};

template <int G, int D>
__global__ __launch_bounds__(256) void sp_mttkrp_csf_kernel(SpCsfArgs a, float* __restrict__ out,
                                                            float* __restrict__ cval, int32_t* __restrict__ crow) {
  constexpr int NPW = 64 / G;      // workers per wave
  constexpr int PER = SP_SUB / G;  // staged leaf entries per lane and array
  constexpr int SP_FWIN = sp_fwin<G>();
  constexpr int WSTRIDE = 2 * SP_SUB + D * SP_FWIN;
  static_assert(PER == 2 || PER % 4 == 0, "staging width");
  static_assert(SP_FWIN >= 2 * SP_VU, "fibre window");
  extern __shared__ int32_t sp_stage[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane / G, q = lane % G;
  int32_t* sLeaf = sp_stage + (wave * NPW + g) * WSTRIDE;
  int32_t* sVal = sLeaf + SP_SUB;
  int32_t* sPre = sVal + SP_SUB;  // D x SP_FWIN prefixes (row first)
  const i64 w = ((i64)blockIdx.x * 4 + wave) * NPW + g;
  const bool active = w < a.nworkers;
  const i64 e0 = w * SP_CSF_W;
  const i64 e1 = active ? ((e0 + SP_CSF_W) < a.nnz ? e0 + SP_CSF_W : a.nnz) : e0;
  const bool live = 4 * q < a.rcount;
  const int col = a.r0 + (live ? 4 * q : 0);
  i64 fb = active ? a.chunk_fib[w] : 0;  // fibre of the next nonzero
  i64 wlo = fb;                           // first fibre of the staged window
  const int first = active ? a.fpre[fb] : 0;
  int cur = first;
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};   // the current row's sum
  floatx4 facc = {0.f, 0.f, 0.f, 0.f};  // the current fibre's leaf sum
  auto flush = [&](int row, const floatx4& v) {
    if (!live) return;
    float* dst = (row == first) ? cval + (2 * w) * a.R + col : out + (i64)row * a.R + col;
    *reinterpret_cast<floatx4*>(dst) = v;
  };
  auto stage_window = [&]() {
    for (int j = q; j < SP_FWIN; j += G) {
      const i64 f = wlo + j < a.nfib ? wlo + j : a.nfib - 1;  // past the last fibre: never used
#pragma unroll
      for (int l = 0; l < D; ++l) sPre[l * SP_FWIN + j] = a.fpre[(i64)l * a.nfib + f];
    }
  };
  if (active) stage_window();
  for (i64 base = e0; base < e1; base += SP_SUB) {
    const int n = (int)((e1 - base) < SP_SUB ? (e1 - base) : SP_SUB);
    sp_stage_part<PER>(a.leafe + base, sLeaf, q * PER, n);
    sp_stage_part<PER>(reinterpret_cast<const int32_t*>(a.vals) + base, sVal, q * PER, n);
    __builtin_amdgcn_wave_barrier();
    for (int i0 = 0; i0 < n; i0 += SP_VU) {
      // a batch touches fibres fb .. fb + SP_VU - 1: keep them in the window
      if (fb + SP_VU > wlo + SP_FWIN) {
        wlo = fb;
        __builtin_amdgcn_wave_barrier();  // earlier window reads issued before it is overwritten
        stage_window();
        __builtin_amdgcn_wave_barrier();
      }
      int le[SP_VU];
      float v[SP_VU];
#pragma unroll
      for (int u = 0; u < SP_VU; u += 4) {
        // staged entries past n are zero: leaf 0, no fibre end, value 0
        const intx4 l4 = *reinterpret_cast<const intx4*>(sLeaf + i0 + u);
        const intx4 v4 = *reinterpret_cast<const intx4*>(sVal + i0 + u);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          le[u + k] = l4[k];
          v[u + k] = __int_as_float(v4[k]);
        }
      }
      int fw[SP_VU];     // window index of the nonzero's fibre
      bool done[SP_VU];  // the nonzero closes its fibre's contribution
      int ends = 0;
#pragma unroll
      for (int u = 0; u < SP_VU; ++u) {
        fw[u] = (int)(fb - wlo) + ends;
        done[u] = i0 + u < n && (le[u] < 0 || base + i0 + u == e1 - 1);
        ends += le[u] < 0 ? 1 : 0;
      }
      // the fibres' rows and middle indices, all LDS reads issued together
      int rr[SP_VU];
      int mi[D > 1 ? D - 1 : 1][SP_VU];
#pragma unroll
      for (int u = 0; u < SP_VU; ++u) {
        rr[u] = sPre[fw[u]];
#pragma unroll
        for (int l = 1; l < D; ++l) mi[l - 1][u] = sPre[l * SP_FWIN + fw[u]];
      }
      floatx4 fl[SP_VU];
#pragma unroll
      for (int u = 0; u < SP_VU; ++u)
        fl[u] = *reinterpret_cast<const floatx4*>(a.F[D - 1] + col + (i64)(le[u] & 0x7fffffff) * a.R);
      floatx4 mid[D > 1 ? D - 1 : 1][SP_VU];
      if constexpr (D > 1) {
#pragma unroll
        for (int l = 1; l < D; ++l)
#pragma unroll
          for (int u = 0; u < SP_VU; ++u)
            mid[l - 1][u] = done[u] ? *reinterpret_cast<const floatx4*>(a.F[l - 1] + col + (i64)mi[l - 1][u] * a.R)
                                    : floatx4{0.f, 0.f, 0.f, 0.f};
      }
#pragma unroll
      for (int u = 0; u < SP_VU; ++u) {
        if (i0 + u >= n) break;  // past the run's end
        facc += v[u] * fl[u];
        if (done[u]) {
          floatx4 t = facc;
          if constexpr (D > 1) {
#pragma unroll
            for (int l = 0; l < D - 1; ++l) t *= mid[l][u];
          }
          if (rr[u] != cur) {
            flush(cur, acc);
            cur = rr[u];
            acc = floatx4{0.f, 0.f, 0.f, 0.f};
          }
          acc += t;
          facc = floatx4{0.f, 0.f, 0.f, 0.f};
        }
      }
      fb += ends;
    }
    __builtin_amdgcn_wave_barrier();  // staged entries read before the next pass overwrites them
  }
  if (!active) return;
  if (live) {
    float* c0 = cval + (2 * w) * a.R + col;
    float* c1 = cval + (2 * w + 1) * a.R + col;
    if (cur == first) {
      *reinterpret_cast<floatx4*>(c0) = acc;
      *reinterpret_cast<floatx4*>(c1) = floatx4{0.f, 0.f, 0.f, 0.f};
    } else {
      *reinterpret_cast<floatx4*>(c1) = acc;
    }
  }
  if (q == 0) {
    crow[2 * w] = first;
    crow[2 * w + 1] = cur;
  }
}

// Segmented reduction of sorted (key, R-vector) slots, one LPN-lane worker per
// tile of SP_TILE slots: keys strictly inside a tile are complete and stored to
// `out`; the tile's first and last keys go to the next level's slots (2 per
// tile).  With `final` set a single worker stores every key.  Fixed order, so
// the result is deterministic.
constexpr int SP_TILE = 32;
constexpr int SP_SEG = 8;  // slots loaded together

template <int LPN>
__global__ __launch_bounds__(256) void sp_segreduce_kernel(const float* __restrict__ ival, const int32_t* __restrict__ ikey,
                                                           i64 S, i64 R, int r0, int rcount, int final_pass,
                                                           float* __restrict__ out, float* __restrict__ oval,
                                                           int32_t* __restrict__ okey) {
  const i64 gt = (i64)blockIdx.x * 256 + threadIdx.x;
  const i64 w = gt / LPN;
  const int lane = (int)(gt - w * LPN);
  const i64 tile = final_pass ? S : SP_TILE;
  const i64 nw = final_pass ? 1 : (S + SP_TILE - 1) / SP_TILE;
  if (w >= nw) return;
  const bool live = lane < rcount;
  const int col = r0 + (live ? lane : 0);
  const i64 k0 = w * tile, k1 = (k0 + tile) < S ? k0 + tile : S;
  const int first = ikey[k0];
  int cur = first;
  double acc = 0.0;
  // SP_SEG slots per round: their keys and values are loaded before any is used
  for (i64 kb = k0; kb < k1; kb += SP_SEG) {
    int kk[SP_SEG];
    float vv[SP_SEG];
#pragma unroll
    for (int u = 0; u < SP_SEG; ++u) {
      const bool ok = kb + u < k1;
      kk[u] = ok ? ikey[kb + u] : cur;
      vv[u] = ok ? ival[(kb + u) * R + col] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < SP_SEG; ++u) {
      if (kb + u >= k1) break;
      if (kk[u] != cur) {
        if (cur == first && !final_pass) {
          if (live) oval[(2 * w) * R + col] = (float)acc;
        } else if (live) {
          out[(i64)cur * R + col] = (float)acc;
        }
        cur = kk[u];
        acc = 0.0;
      }
      acc += (double)vv[u];
    }
  }
  if (final_pass) {
    if (live) out[(i64)cur * R + col] = (float)acc;
    return;
  }
  if (cur == first) {
    if (live) {
      oval[(2 * w) * R + col] = (float)acc;
      oval[(2 * w + 1) * R + col] = 0.f;
    }
  } else if (live) {
    oval[(2 * w + 1) * R + col] = (float)acc;
  }
  if (lane == 0) {
    okey[2 * w] = first;
    okey[2 * w + 1] = cur;
  }
}

__global__ __launch_bounds__(256) void zero_f32_kernel(float* __restrict__ p, i64 n) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

template <int LPN>
static int sp_mttkrp_launch(tenrso_ctx* ctx, SpArgs& a, float* out, float* cv0, int32_t* ck0, float* cv1,
                            int32_t* ck1, bool vec, const SpCsfArgs* csf) {
  const i64 R = a.R;
  if (csf) {
    constexpr int G = LPN / 4;
    constexpr int NPW = 64 / G;
    const int D = a.nd_oth;
    SpCsfArgs ca = *csf;
    ca.r0 = a.r0;
    ca.rcount = a.rcount;
    const dim3 grid(cdiv(ca.nworkers, 4 * NPW)), block(256);
    constexpr size_t lds1 = sp_csf_lds_bytes<G, 1>(), lds2 = sp_csf_lds_bytes<G, 2>(),
                     lds3 = sp_csf_lds_bytes<G, 3>();
    if (D == 1)
      hipLaunchKernelGGL((sp_mttkrp_csf_kernel<G, 1>), grid, block, lds1, ctx->stream, ca, out, cv0, ck0);
    else if (D == 2)
      hipLaunchKernelGGL((sp_mttkrp_csf_kernel<G, 2>), grid, block, lds2, ctx->stream, ca, out, cv0, ck0);
    else
      hipLaunchKernelGGL((sp_mttkrp_csf_kernel<G, 3>), grid, block, lds3, ctx->stream, ca, out, cv0, ck0);
  } else if (vec) {
    constexpr int G = LPN / 4;
    constexpr int NPW = 64 / G;
    hipLaunchKernelGGL(sp_mttkrp_vec_kernel<G>, dim3(cdiv(a.nworkers, 4 * NPW)), dim3(256),
                       sp_vec_lds_bytes<G>(a.nd_oth), ctx->stream, a, out, cv0, ck0);
  } else {
    hipLaunchKernelGGL(sp_mttkrp_kernel<LPN>, dim3(cdiv(a.nworkers * LPN, 256)), dim3(256), 0, ctx->stream, a, out,
                       cv0, ck0);
  }
  i64 S = 2 * a.nworkers;
  float* iv = cv0;
  int32_t* ik = ck0;
  float* ov = cv1;
  int32_t* ok = ck1;
  while (S > SP_TILE) {
    const i64 nw = (S + SP_TILE - 1) / SP_TILE;
    hipLaunchKernelGGL(sp_segreduce_kernel<LPN>, dim3(cdiv(nw * LPN, 256)), dim3(256), 0, ctx->stream,
                       (const float*)iv, (const int32_t*)ik, S, R, a.r0, a.rcount, 0, out, ov, ok);
    S = 2 * nw;
    std::swap(iv, ov);
    std::swap(ik, ok);
  }
  hipLaunchKernelGGL(sp_segreduce_kernel<LPN>, dim3(cdiv(LPN, 256)), dim3(256), 0, ctx->stream, (const float*)iv,
                     (const int32_t*)ik, S, R, a.r0, a.rcount, 1, out, ov, ok);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

int sp_mttkrp(tenrso_ctx* ctx, const tenrso_sptensor* t, const float* const* F, i64 R, int mode, float* out) {
  const SpMode& md = t->modes[mode];
  const i64 I = md.rows;
  if (I * R > 0)
    hipLaunchKernelGGL(zero_f32_kernel, dim3(cdiv(I * R, 256)), dim3(256), 0, ctx->stream, out, I * R);
  TENRSO_HIP_TRY(hipGetLastError());
  if (t->nnz == 0 || R == 0) return TENRSO_OK;
  SpArgs a{};
  a.nd_oth = t->nd - 1;
  a.nnz = t->nnz;
  a.rows = I;
  a.R = R;
  a.row = md.row;
  a.oth = md.oth;
  a.vals = md.vals;
  int o = 0;
  for (int k = 0; k < t->nd; ++k)
    if (k != mode) a.F[o++] = F[k];
  // 16-byte factor gathers and stores need rank columns in groups of 4 and
  // 16-byte aligned factor / output bases
  bool vec = (R % 4) == 0 && (reinterpret_cast<uintptr_t>(out) & 15) == 0;
  for (int k = 0; k < a.nd_oth; ++k) vec = vec && (reinterpret_cast<uintptr_t>(a.F[k]) & 15) == 0;
  // the fibre form for 2- to 4-way tensors (vector gathers, fibres built)
  const bool use_csf = vec && a.nd_oth >= 1 && a.nd_oth <= 3 && md.leafe != nullptr && md.nfib > 0;
  SpCsfArgs ca{};
  if (use_csf) {
    ca.nnz = t->nnz;
    ca.rows = I;
    ca.R = R;
    ca.nfib = md.nfib;
    ca.leafe = md.leafe;
    ca.vals = md.vals;
    ca.fpre = md.fpre;
    ca.chunk_fib = md.chunk_fib;
    for (int k = 0; k < a.nd_oth; ++k) ca.F[k] = a.F[k];
    ca.nworkers = md.nchunks;
  }
  a.nworkers = use_csf ? md.nchunks : (vec ? (t->nnz + SP_W - 1) / SP_W : (t->nnz + SP_CH - 1) / SP_CH);
  const i64 nslots = 2 * a.nworkers;
  void *pv, *pk;
  TENRSO_TRY(ctx_buffer(ctx, SLOT_TMP, (size_t)2 * nslots * R * sizeof(float), &pv));
  TENRSO_TRY(ctx_buffer(ctx, SLOT_TMP2, (size_t)2 * nslots * sizeof(int32_t), &pk));
  float* cv0 = (float*)pv;
  float* cv1 = cv0 + nslots * R;
  int32_t* ck0 = (int32_t*)pk;
  int32_t* ck1 = ck0 + nslots;
  const int LPN = R <= 16 ? 16 : (R <= 32 ? 32 : 64);
  const int rchunks = (int)((R + LPN - 1) / LPN);
  for (int rc = 0; rc < rchunks; ++rc) {
    a.r0 = rc * LPN;
    a.rcount = (int)std::min<i64>(LPN, R - a.r0);
    const SpCsfArgs* cp = use_csf ? &ca : nullptr;
    if (LPN == 16) TENRSO_TRY(sp_mttkrp_launch<16>(ctx, a, out, cv0, ck0, cv1, ck1, vec, cp));
    else if (LPN == 32) TENRSO_TRY(sp_mttkrp_launch<32>(ctx, a, out, cv0, ck0, cv1, ck1, vec, cp));
    else TENRSO_TRY(sp_mttkrp_launch<64>(ctx, a, out, cv0, ck0, cv1, ck1, vec, cp));
  }
  return TENRSO_OK;
}

// ---------------------------------------------------------------- CSF export
// Standard CSF (SPLATT semantics) of mode order [mode, others ascending] from the
// sorted nonzeros: level l fibers are the runs of equal (i_0..i_l) prefixes;
// fptr[l][f] .. fptr[l][f+1] index the children at level l+1 (the leaf level
// indexes nonzeros).  Built on the host from the device-sorted arrays (a format
// export for callers and tests, not on the MTTKRP path).
int sp_export_csf(tenrso_ctx* ctx, const tenrso_sptensor* t, int mode, std::vector<std::vector<int64_t>>& fptr,
                  std::vector<std::vector<int32_t>>& fids, std::vector<float>& vals) {
  const SpMode& md = t->modes[mode];
  const i64 nnz = t->nnz;
  const int nd = t->nd;
  std::vector<std::vector<int32_t>> lvl(nd, std::vector<int32_t>((size_t)nnz));
  vals.assign((size_t)nnz, 0.f);
  if (nnz > 0) {
    TENRSO_HIP_TRY(hipMemcpyAsync(lvl[0].data(), md.row, (size_t)nnz * 4, hipMemcpyDeviceToHost, ctx->stream));
    for (int l = 1; l < nd; ++l)
      TENRSO_HIP_TRY(hipMemcpyAsync(lvl[l].data(), md.oth + (i64)(l - 1) * nnz, (size_t)nnz * 4,
                                    hipMemcpyDeviceToHost, ctx->stream));
    TENRSO_HIP_TRY(hipMemcpyAsync(vals.data(), md.vals, (size_t)nnz * 4, hipMemcpyDeviceToHost, ctx->stream));
  }
  TENRSO_HIP_TRY(hipStreamSynchronize(ctx->stream));
  fptr.assign(nd, {});
  fids.assign(nd, {});
  // level starts: positions where the prefix (levels 0..l) changes
  std::vector<std::vector<int64_t>> starts(nd);
  for (int l = 0; l < nd; ++l) {
    for (i64 e = 0; e < nnz; ++e) {
      bool change = (e == 0);
      for (int q = 0; q <= l && !change; ++q) change = lvl[q][(size_t)e] != lvl[q][(size_t)e - 1];
      if (change || l == nd - 1) starts[l].push_back(e);
    }
    for (int64_t s : starts[l]) fids[l].push_back(lvl[l][(size_t)s]);
  }
  for (int l = 0; l < nd; ++l) {
    if (l == nd - 1) {
      fptr[l].push_back(0);  // leaf level: one entry per nonzero, no children
      continue;
    }
    // children of fiber f at level l are the level-(l+1) fibers starting inside it
    const auto& cs = starts[l + 1];
    size_t ci = 0;
    for (size_t f = 0; f < starts[l].size(); ++f) {
      fptr[l].push_back((int64_t)ci);
      const int64_t end = (f + 1 < starts[l].size()) ? starts[l][f + 1] : nnz;
      while (ci < cs.size() && cs[ci] < end) ++ci;
    }
    fptr[l].push_back((int64_t)ci);
  }
  return TENRSO_OK;
}


// ---------------------------------------------------------------- HiCOO
struct HiCooArgs {
  int nd;
  i64 nnz;
  i64 bs[MAXD_SP];   // block shape
  i64 nbk[MAXD_SP];  // blocks per mode, ceil(shape / bs)
  const int32_t* idx;
};

// key = block coordinates (index / bs) linearised, mode 0 most significant, so
// the key order is the lexicographic order of the block coordinates
__global__ __launch_bounds__(256) void hicoo_keys_kernel(HiCooArgs a, unsigned long long* __restrict__ keys,
                                                         int32_t* __restrict__ pos) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e >= a.nnz) return;
  unsigned long long k = 0;
  for (int m = 0; m < a.nd; ++m)
    k = k * (unsigned long long)a.nbk[m] + (unsigned long long)(a.idx[(i64)m * a.nnz + e] / a.bs[m]);
  keys[e] = k;
  pos[e] = (int32_t)e;
}

// local coordinates (index % bs) and values in the sorted order
__global__ __launch_bounds__(256) void hicoo_gather_kernel(HiCooArgs a, const int32_t* __restrict__ perm,
                                                           const float* __restrict__ vals, int32_t* __restrict__ local,
                                                           float* __restrict__ vout) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e >= a.nnz) return;
  const i64 src = perm[e];
  for (int m = 0; m < a.nd; ++m) local[(i64)m * a.nnz + e] = (int32_t)(a.idx[(i64)m * a.nnz + src] % a.bs[m]);
  vout[e] = vals[src];
}

int sp_export_hicoo(tenrso_ctx* ctx, const tenrso_sptensor* t, const i64* bs, std::vector<int64_t>& block_coords,
                    std::vector<int64_t>& block_ptrs, std::vector<int32_t>& local, std::vector<float>& vals) {
  const int nd = t->nd;
  const i64 nnz = t->nnz;
  block_coords.clear();
  block_ptrs.assign(1, 0);
  local.assign((size_t)nd * nnz, 0);
  vals.assign((size_t)nnz, 0.f);
  if (nnz == 0) return TENRSO_OK;
  HiCooArgs ha{};
  ha.nd = nd;
  ha.nnz = nnz;
  ha.idx = t->idx;
  int bits = 0;
  {
    unsigned long long maxkey = 1;
    for (int k = 0; k < nd; ++k) {
      ha.bs[k] = bs[k];
      ha.nbk[k] = (t->shape[k] + bs[k] - 1) / bs[k];
      if (maxkey > ~0ull / (unsigned long long)ha.nbk[k]) {
        set_error("HiCOO: the block grid has more than 2^64 blocks");
        return TENRSO_ERR_UNSUPPORTED;
      }
      maxkey *= (unsigned long long)ha.nbk[k];
    }
    while (bits < 64 && (maxkey >> bits) != 0) ++bits;
  }
  unsigned long long *kin = nullptr, *kout = nullptr;
  int32_t *pin = nullptr, *pout = nullptr, *dloc = nullptr;
  float* dval = nullptr;
  void* tmp = nullptr;
  int rc = TENRSO_OK;
  auto fail = [&](hipError_t e) {
    set_error("HIP error in the HiCOO export: %s", hipGetErrorString(e));
    rc = TENRSO_ERR_DEVICE;
  };
  hipError_t e = hipMalloc(&kin, (size_t)nnz * 8);
  if (e == hipSuccess) e = hipMalloc(&kout, (size_t)nnz * 8);
  if (e == hipSuccess) e = hipMalloc(&pin, (size_t)nnz * 4);
  if (e == hipSuccess) e = hipMalloc(&pout, (size_t)nnz * 4);
  if (e == hipSuccess) e = hipMalloc(&dloc, (size_t)nd * nnz * 4);
  if (e == hipSuccess) e = hipMalloc(&dval, (size_t)nnz * 4);
  size_t tmp_bytes = 0;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(hicoo_keys_kernel, dim3(cdiv(nnz, 256)), dim3(256), 0, ctx->stream, ha, kin, pin);
    e = hipGetLastError();
  }
  // LSD radix sort: stable, so nonzeros of one block keep their COO order (the
  // reference's sort_by is stable too)
  if (e == hipSuccess)
    e = rocprim::radix_sort_pairs(nullptr, tmp_bytes, kin, kout, pin, pout, (size_t)nnz, 0, bits, ctx->stream);
  if (e == hipSuccess) e = hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 16));
  if (e == hipSuccess)
    e = rocprim::radix_sort_pairs(tmp, tmp_bytes, kin, kout, pin, pout, (size_t)nnz, 0, bits, ctx->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(hicoo_gather_kernel, dim3(cdiv(nnz, 256)), dim3(256), 0, ctx->stream, ha, pout, t->vals, dloc,
                       dval);
    e = hipGetLastError();
  }
  std::vector<unsigned long long> keys;
  if (e == hipSuccess) {
    keys.resize((size_t)nnz);
    e = hipMemcpyAsync(keys.data(), kout, (size_t)nnz * 8, hipMemcpyDeviceToHost, ctx->stream);
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(local.data(), dloc, (size_t)nd * nnz * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(vals.data(), dval, (size_t)nnz * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) fail(e);
  for (void* p : {(void*)kin, (void*)kout, (void*)pin, (void*)pout, (void*)dloc, (void*)dval, tmp})
    if (p) (void)hipFree(p);
  if (rc != TENRSO_OK) return rc;
  // blocks: runs of equal keys; block coordinates decoded from the key
  std::vector<unsigned long long> bkeys;
  for (i64 i = 0; i < nnz; ++i)
    if (i == 0 || keys[(size_t)i] != keys[(size_t)i - 1]) {
      if (i > 0) block_ptrs.push_back(i);
      bkeys.push_back(keys[(size_t)i]);
    }
  block_ptrs.push_back(nnz);
  const size_t nb = bkeys.size();
  block_coords.assign((size_t)nd * nb, 0);
  for (size_t b = 0; b < nb; ++b) {
    unsigned long long k = bkeys[b];
    for (int m = nd - 1; m >= 0; --m) {
      block_coords[(size_t)m * nb + b] = (int64_t)(k % (unsigned long long)ha.nbk[m]);
      k /= (unsigned long long)ha.nbk[m];
    }
  }
  return TENRSO_OK;
}

}  // namespace tenrso
