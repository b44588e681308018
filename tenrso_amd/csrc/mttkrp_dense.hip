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
//   * j = (slow, k): k runs over the fastest other mode, slow over the rest, and
//     KR[j, r] = PRE[slow, r] * FAST[k, r], where FAST is the factor of the
//     fastest other mode and PRE the forward KR of the remaining other modes (a
//     factor itself for 3-way tensors).  A workgroup stages a window of FAST,
//     transposed to [r][k], in LDS once.  FP32 forms the Khatri-Rao operand in
//     registers (FAST from LDS times PRE[slow, r]); BF16 contracts
//     T[i, r] = sum_k X[.., i, .., k] FAST[k, r] per slow index and adds
//     PRE[slow, r] T[i, r] into its accumulator.  No Khatri-Rao row is ever
//     written to memory.
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
#include <tuple>
#include <type_traits>

#include "als_job.hpp"
#include "mttkrp.hpp"

namespace tenrso {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));

struct KernArgs {
  i64 I, Q;           // X viewed as [P, I, Q]
  i64 R;              // factor row stride (rank)
  i64 fast_size;      // size of the fastest other mode
  i64 fast_stride;    // its extent in memory: fast_size, or more in a copy whose last
                      // mode is padded (Q is then the padded extent too)
  i64 qs;             // Q / fast_stride: slow indices per p (case A)
  const float* pre;   // PRE (rows = J / fast_size, stride R) or null (= 1)
  const float* fast;  // FAST factor (fast_size x R)
  int kw;             // window of the fastest mode held in LDS (multiple of KB, divides fast_size)
  int nwin;           // fast_size / kw
  i64 n_slow;         // J / fast_size
  int slow_splits;    // splits of the slow range; grid.y = nwin * slow_splits
  i64 slow_q, slow_rem;  // n_slow = slow_q * slow_splits + slow_rem
  int r0;             // first rank column of this launch
  int rcount;         // rank columns handled by the launch (<= 16*NT)
  int rtiles;         // row tiles; workgroup tile -> (tile % rtiles, split = tile / rtiles)
  int win_vec;        // FAST 16-byte aligned and R % 4 == 0: 16-byte window loads
};

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// Stage FAST[w*kw .. w*kw+kw) (columns r0 .. r0+RT) into LDS transposed: s[r][kk].
// Work unit u of the workgroup covers ranks r4 .. r4+3 (r4 = 4*(u % (RT/4))) of
// window rows kk = 2*(u / (RT/4)) and kk + 1, so it leaves LDS as one 8-byte
// (fp32) or two 4-byte (bf16 hi, lo) pairs per rank.  With a 16-byte aligned
// FAST and R % 4 == 0 (win_vec) a unit is two 16-byte loads, else eight 4-byte
// loads (one kernel instance per case, so that every instance has a single
// load sequence and the waits stay exact).  Ranks past rcount and rows past kw load column r0 of window row 0
// instead (always inside FAST): those rank columns are never stored and those
// rows never read, so no value select is needed and no load waits early.
// Split in two so the window loads are issued before the first X blocks: the
// LDS writes then wait for the window only (vector loads retire in issue order).
template <int RT, int NTH, int KWMAX>
struct FastWindow {
  static constexpr int UNITS = RT / 4 * (KWMAX / 2);
  static constexpr int PERU = (UNITS + NTH - 1) / NTH;
  // every unit index below PERU*NTH maps in bounds: u >= UNITS gives kk >= KWMAX
  static __device__ __forceinline__ bool live(int u) { return UNITS % NTH == 0 || u < UNITS; }
  floatx4 v[PERU][2];
};

// kv: rows of the window that exist (kw, or fewer in the last window of a
// fastest mode that kw does not divide); a row past kv is read as row 0 (always
// inside FAST) and stored as zero by stage_window_store*.
template <bool WV, int RT, int NTH, int KWMAX>
__device__ __forceinline__ void stage_window_load(FastWindow<RT, NTH, KWMAX>& fw, const KernArgs& a, int w, int kv) {
  const float* base = a.fast + (i64)w * a.kw * a.R + a.r0;
  if constexpr (WV) {
#pragma unroll
    for (int e = 0; e < FastWindow<RT, NTH, KWMAX>::PERU; ++e) {
      const int u = threadIdx.x + e * NTH;
      const int r4 = 4 * (u % (RT / 4)), kk = 2 * (u / (RT / 4));
      const bool ok = kk < kv && r4 < a.rcount;
      const bool ok1 = ok && kk + 1 < kv;
      const float* p = base + (ok ? kk * a.R + r4 : 0);
      fw.v[e][0] = *reinterpret_cast<const floatx4*>(p);
      fw.v[e][1] = *reinterpret_cast<const floatx4*>(p + (ok1 ? a.R : 0));
    }
  } else {
#pragma unroll
    for (int e = 0; e < FastWindow<RT, NTH, KWMAX>::PERU; ++e) {
      const int u = threadIdx.x + e * NTH;
      const int r4 = 4 * (u % (RT / 4)), kk = 2 * (u / (RT / 4));
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const bool ok = kk < kv && r4 + q < a.rcount;
        const bool ok1 = ok && kk + 1 < kv;
        const float* p = base + (ok ? kk * a.R + r4 + q : 0);
        fw.v[e][0][q] = p[0];
        fw.v[e][1][q] = p[ok1 ? a.R : 0];
      }
    }
  }
}

// rows kk >= kv of a staged pair are zero (they meet masked, zero X elements)
template <int RT, int NTH, int KWMAX>
__device__ __forceinline__ void window_zero_tail(FastWindow<RT, NTH, KWMAX>& fw, int kv) {
#pragma unroll
  for (int e = 0; e < FastWindow<RT, NTH, KWMAX>::PERU; ++e) {
    const int u = threadIdx.x + e * NTH;
    const int kk = 2 * (u / (RT / 4));
    if (kk >= kv) fw.v[e][0] = floatx4{0.f, 0.f, 0.f, 0.f};
    if (kk + 1 >= kv) fw.v[e][1] = floatx4{0.f, 0.f, 0.f, 0.f};
  }
}

template <int RT, int PITCH, int NTH, int KWMAX>
__device__ __forceinline__ void stage_window_store(float* __restrict__ s, const FastWindow<RT, NTH, KWMAX>& fw) {
  static_assert(PITCH % 2 == 0, "pairs of window rows are stored as 8-byte words");
#pragma unroll
  for (int e = 0; e < FastWindow<RT, NTH, KWMAX>::PERU; ++e) {
    const int u = threadIdx.x + e * NTH;
    const int r4 = 4 * (u % (RT / 4)), kk = 2 * (u / (RT / 4));
    if (!FastWindow<RT, NTH, KWMAX>::live(u)) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      *reinterpret_cast<float2*>(&s[(r4 + q) * PITCH + kk]) = float2{fw.v[e][0][q], fw.v[e][1][q]};
  }
}

// Same window as an exact bf16 pair: FAST = hi + lo with hi = bf16(FAST) and
// lo = bf16(FAST - hi), so hi + lo carries 16 significant bits.
template <int RT, int PITCH, int NTH, int KWMAX>
__device__ __forceinline__ void stage_window_store_bf16x2(uint16_t* __restrict__ shi, uint16_t* __restrict__ slo,
                                                          const FastWindow<RT, NTH, KWMAX>& fw) {
  static_assert(PITCH % 2 == 0, "pairs of window rows are stored as 4-byte words");
#pragma unroll
  for (int e = 0; e < FastWindow<RT, NTH, KWMAX>::PERU; ++e) {
    const int u = threadIdx.x + e * NTH;
    const int r4 = 4 * (u % (RT / 4)), kk = 2 * (u / (RT / 4));
    if (!FastWindow<RT, NTH, KWMAX>::live(u)) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      // v_cvt_pk_bf16_f32 rounds to nearest even, as f32_to_bf16_rne does for
      // every non-NaN value; element 0 (row kk) is the low half
      const floatx2 f = {fw.v[e][0][q], fw.v[e][1][q]};
      const uint32_t hw = __builtin_bit_cast(uint32_t, __builtin_convertvector(f, bf16x2));
      const floatx2 r = {f[0] - __uint_as_float(hw << 16), f[1] - __uint_as_float(hw & 0xffff0000u)};
      const uint32_t lw = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bf16x2));
      const int o = (r4 + q) * PITCH + kk;
      *reinterpret_cast<uint32_t*>(&shi[o]) = hw;
      *reinterpret_cast<uint32_t*>(&slo[o]) = lw;
    }
  }
}

// Same window as an exact bf16 triple: FAST = hi + lo + lo2 (hi = bf16(FAST),
// lo = bf16(FAST - hi), lo2 = bf16(FAST - hi - lo)); each residual is exact in
// fp32 and the last one has at most 8 significant bits, so for every normal
// fp32 value hi + lo + lo2 equals FAST exactly.
template <int RT, int PITCH, int NTH, int KWMAX>
__device__ __forceinline__ void stage_window_store_bf16x3(uint16_t* __restrict__ shi, uint16_t* __restrict__ slo,
                                                          uint16_t* __restrict__ slo2,
                                                          const FastWindow<RT, NTH, KWMAX>& fw) {
  static_assert(PITCH % 2 == 0, "pairs of window rows are stored as 4-byte words");
#pragma unroll
  for (int e = 0; e < FastWindow<RT, NTH, KWMAX>::PERU; ++e) {
    const int u = threadIdx.x + e * NTH;
    const int r4 = 4 * (u % (RT / 4)), kk = 2 * (u / (RT / 4));
    if (!FastWindow<RT, NTH, KWMAX>::live(u)) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const floatx2 f = {fw.v[e][0][q], fw.v[e][1][q]};
      const uint32_t hw = __builtin_bit_cast(uint32_t, __builtin_convertvector(f, bf16x2));
      const floatx2 r = {f[0] - __uint_as_float(hw << 16), f[1] - __uint_as_float(hw & 0xffff0000u)};
      const uint32_t lw = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bf16x2));
      const floatx2 r2 = {r[0] - __uint_as_float(lw << 16), r[1] - __uint_as_float(lw & 0xffff0000u)};
      const uint32_t l2w = __builtin_bit_cast(uint32_t, __builtin_convertvector(r2, bf16x2));
      const int o = (r4 + q) * PITCH + kk;
      *reinterpret_cast<uint32_t*>(&shi[o]) = hw;
      *reinterpret_cast<uint32_t*>(&slo[o]) = lw;
      *reinterpret_cast<uint32_t*>(&slo2[o]) = l2w;
    }
  }
}

// Exact three-way bf16 split of 8 fp32 values (see stage_window_store_bf16x3).
__device__ __forceinline__ void split_bf16x3(const floatx8& x, bf16x8& h, bf16x8& l, bf16x8& l2) {
  h = __builtin_convertvector(x, bf16x8);
  const floatx8 r = x - __builtin_convertvector(h, floatx8);
  l = __builtin_convertvector(r, bf16x8);
  const floatx8 r2 = r - __builtin_convertvector(l, floatx8);
  l2 = __builtin_convertvector(r2, bf16x8);
}

// A value the compiler cannot prove uniform but that is: move it to scalar registers.
__device__ __forceinline__ i64 uniform_i64(i64 v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (i64)(((uint64_t)hi << 32) | lo);
}

// PRE row used when the MTTKRP has no PRE factor (a matrix): KR = FAST.
#define TENRSO_ONES8 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f
__device__ float kPreOnes[64] = {TENRSO_ONES8, TENRSO_ONES8, TENRSO_ONES8, TENRSO_ONES8,
                                       TENRSO_ONES8, TENRSO_ONES8, TENRSO_ONES8, TENRSO_ONES8};
#undef TENRSO_ONES8

// Block schedule of one wave.  Block t of the wave is block kb = t % nkb of the
// slow index s_lo + wave + WAVES*(t / nkb).  Loads run in t order, so the cursor
// only steps forward, and it stops on the wave's last block (loads past the end
// are redirected by the kernels to one cache line).  Every X and PRE load is thus unconditional
// and the wait before a block's MFMAs counts exactly the loads issued after it,
// which keeps the prefetch in flight.  A wave with no block reads the block of
// slow index min(s_lo + wave, n_slow - 1), always inside X, and computes nothing.
// Case A (X as [P, I, Q], slow = p*qs + sq): the block starts at element
// p*I*Q + sq*fast_size + w*kw + kb*KB.  Case B: the block covers j = slow*fast_size
// + w*kw + kb*KB onwards, one row of I elements per j.  All fields are uniform.
template <int STRIDE, int KB, bool CASE_B>
struct BlockCursor {
  int kb, nkb;
  i64 slow, sq, base, wrap;

  // starts at block kb0 of slow index slow0; each slow index has nkb blocks, and
  // after its last one the cursor moves STRIDE slow indices on
  __device__ __forceinline__ BlockCursor(const KernArgs& a, i64 slow0, int w, int nkb_, int kb0) {
    kb = kb0;
    nkb = nkb_;
    slow = slow0 < a.n_slow ? slow0 : a.n_slow - 1;
    if constexpr (!CASE_B) {
      // 32-bit division whenever the slow range fits (it nearly always does)
      const i64 p = (a.n_slow >> 32) ? slow / a.qs : (i64)((uint32_t)slow / (uint32_t)a.qs);
      sq = slow - p * a.qs;
      base = p * a.I * a.Q + sq * a.fast_stride + (i64)w * a.kw;
      wrap = a.I * a.Q - a.qs * a.fast_stride;  // from (p, qs) to (p + 1, 0)
    } else {
      sq = 0;
      wrap = 0;
      base = slow * a.fast_size + (i64)w * a.kw;
    }
  }
  // element offset in X of the block's first element (lane offsets are added on top)
  __device__ __forceinline__ i64 offset(const KernArgs& a) const {
    if constexpr (!CASE_B) return base + (i64)kb * KB;
    else return (base + (i64)kb * KB) * a.I;
  }
  __device__ __forceinline__ void step(const KernArgs& a) {
    if (++kb == nkb) {
      kb = 0;
      slow += STRIDE;
      base += (i64)STRIDE * a.fast_stride;
      if constexpr (!CASE_B) {
        sq += STRIDE;
        while (sq >= a.qs) {
          sq -= a.qs;
          base += wrap;
        }
      }
    }
  }
};

// Fixed-order reduction of the WAVES per-wave accumulators of a workgroup, spread
// over the waves: tile q = mt*NT + nt belongs to wave q % WAVES, which adds the
// WAVES copies of it in wave order 0, 1, .. (one LDS round trip, 16 bytes per
// lane and tile) and stores it; every wave stores its own tiles, so the partial
// leaves the workgroup from all waves at once.  LDS use: MT*NT*WAVES*64 floatx4.
template <int MT, int NT, int WAVES>
__device__ __forceinline__ void wg_reduce_store(floatx4 (&acc)[MT][NT], float* __restrict__ red, int wave, int lane,
                                                float* __restrict__ out, i64 rows, i64 R, int r0, int rcount,
                                                i64 row_base, bool case_b) {
  floatx4* r4 = reinterpret_cast<floatx4*>(red);
  const int g = lane >> 4, c = lane & 15;
  __syncthreads();  // LDS window no longer read
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int q = mt * NT + nt;
      if (q % WAVES != wave) r4[(q * WAVES + wave) * 64 + lane] = acc[mt][nt];
    }
  __syncthreads();
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int q = mt * NT + nt;
      if (q % WAVES != wave) continue;
      floatx4 sum = wave == 0 ? acc[mt][nt] : r4[(q * WAVES + 0) * 64 + lane];
#pragma unroll
      for (int wv = 1; wv < WAVES; ++wv) sum += wv == wave ? acc[mt][nt] : r4[(q * WAVES + wv) * 64 + lane];
      if (16 * nt + c >= rcount) continue;
      const int col = r0 + 16 * nt + c;
      // D layout (16x16): lane (g, c), register v -> tile row 4g + v, column c
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const i64 row = case_b ? row_base + MT * (4 * g + v) + mt : row_base + 16 * mt + 4 * g + v;
        if (row < rows) out[row * R + col] = sum[v];
      }
    }
}

// Fold the per-slow accumulator T into acc with the PRE row of that slow index:
// acc[:, r] += PRE[slow, r] * T[:, r]  (lane (g, c) holds column 16nt + c).
template <int MT, int NT>
__device__ __forceinline__ void fold_slow(floatx4 (&acc)[MT][NT], floatx4 (&T)[MT][NT], const float (&pre)[NT]) {
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      acc[mt][nt] += T[mt][nt] * pre[nt];
      T[mt][nt] = floatx4{0.f, 0.f, 0.f, 0.f};
    }
}

// ---------------------------------------------------------------------------
// FP32 MFMA kernel (v_mfma_f32_16x16x4f32).
//   A workgroup owns ROWS = 16*MT rows and one window w of the fastest mode; its
//   WAVES waves share the LDS window and split the workgroup's blocks (slow
//   index, block of the window within it) into contiguous balanced ranges;
//   their accumulators are summed in LDS in fixed order at the end.
//   The B operand of the MFMAs is the Khatri-Rao row itself, formed on the fly:
//   the FAST window is read as 4 consecutive k from LDS and scaled by
//   PRE[slow, r] of the lane's column, KR[j, r] = PRE[slow, r] * FAST[k, r]
//   (one fp32 rounding, as the reference's materialised KR).  The accumulators
//   stay in the matrix-core registers for the whole loop.
//   One block = KB = 32 consecutive j: 2 halves h x 4 lane groups g x 4 s, so
//   MFMA step (h, s) of lane group g contracts j = j0 + 16h + 4g + s.
//   Case A (q contiguous): lane (g, c) of m-tile mt holds row row_base+16mt+c and
//     loads 4 consecutive q with one 16-byte load per h.
//   Case B (i contiguous, last mode): rows are interleaved, lane c of m-tile mt
//     holds row row_base + MT*c + mt, so one MT-wide vector load per j covers
//     16*MT consecutive rows.
//   X blocks are prefetched 3 blocks ahead in registers.  The FAST window loads
//   are issued first, then the first three X blocks, so the window reaches LDS
//   while X is still arriving.
// ---------------------------------------------------------------------------
template <int NT, int WAVES, int MT, bool CASE_B, int KWMAX, bool WV, int XL>
__global__ __launch_bounds__(WAVES * 64) void mttkrp_f32_mfma(const float* __restrict__ X, KernArgs a,
                                                              float* __restrict__ partial, AlsJob job, int has_job) {
  constexpr int KB = 32, PITCH = KWMAX + 4, RT = 16 * NT, NTH = WAVES * 64, ROWS = 16 * MT;
  constexpr int LDS_WIN = RT * PITCH, LDS_RED = MT * NT * WAVES * 64 * 4;
  constexpr int LDS_F = LDS_WIN > LDS_RED ? LDS_WIN : LDS_RED;
  constexpr int LDS_FLOATS = LDS_F * 4 > JOB_LDS_BYTES ? LDS_F : JOB_LDS_BYTES / 4;
  __shared__ __attribute__((aligned(16))) float smem[LDS_FLOATS];
  static_assert(NTH == JOB_THREADS, "the ALS job needs a 256-thread workgroup");
  // workgroup 0 of a launch with a job runs the ALS linear algebra of the mode
  if (has_job && blockIdx.x == 0) {
    als_job_run(job, reinterpret_cast<double*>(smem));
    return;
  }
  const int tile = (int)blockIdx.x - (has_job ? 1 : 0);
  const int xt = tile % a.rtiles;
  float* sF = smem;
  const int tid = threadIdx.x, lane = tid & 63;
  // the wave index is uniform: everything derived from it (the block schedule,
  // the X block addresses, the loop bounds) lives in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int g = lane >> 4, c = lane & 15;
  const int split = tile / a.rtiles;
  const int w = split % a.nwin;
  const int ss = split / a.nwin;
  // balanced split of the slow range: the first slow_rem splits take one more
  const i64 s_lo = uniform_i64((i64)ss * a.slow_q + (ss < a.slow_rem ? ss : a.slow_rem));
  const i64 s_hi = s_lo + a.slow_q + (ss < a.slow_rem ? 1 : 0);
  // rows of this window that exist: kw, or fewer in the last window of a fastest
  // mode that kw does not divide (XL > 0: its loads are masked past kv)
  const int kv = XL == 0 ? a.kw : (int)min((i64)a.kw, a.fast_size - (i64)w * a.kw);
  const int nkb = XL == 0 ? a.kw / KB : (kv + KB - 1) / KB;
  // the workgroup's blocks (slow index, then block within it) in WAVES contiguous,
  // balanced ranges: with 9 slow indices of 4 blocks, 9 blocks per wave rather
  // than 12 for one wave and 8 for the others
  const i64 nblk_wg = (s_hi - s_lo) * nkb;
  const i64 wb0 = uniform_i64(wave * nblk_wg / WAVES), wb1 = (wave + 1) * nblk_wg / WAVES;
  const int T = __builtin_amdgcn_readfirstlane((int)(wb1 - wb0));  // blocks of this wave
  const i64 row_base = (i64)xt * ROWS;

  // window loads first (see stage_window_load)
  FastWindow<RT, NTH, KWMAX> fw;
  stage_window_load<WV>(fw, a, w, kv);

  i64 rowc[MT];
  i64 rowv = 0;  // case B: first row of this lane's MT-vector
  if constexpr (!CASE_B) {
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const i64 rr = row_base + 16 * mt + c;
      rowc[mt] = rr < a.I ? rr : a.I - 1;
    }
  } else {
    const i64 rr = row_base + MT * c;
    rowv = rr <= a.I - MT ? rr : a.I - MT;
  }

  floatx4 acc[MT][NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = floatx4{0.f, 0.f, 0.f, 0.f};

  struct Buf {
    floatx4 x[MT][2];
    float pre[NT];
    int kb;
    uint32_t okm;  // XL > 0: bit 4h + s set where k = kb*KB + 16h + 4g + s < kv
  };

  // per-lane element offsets, fixed for the whole loop; the per-block part of
  // every address is wave-uniform
  i64 loff[MT];
  if constexpr (!CASE_B) {
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) loff[mt] = rowc[mt] * a.Q + 4 * g;
  } else {
    loff[0] = (i64)(4 * g) * a.I + rowv;
  }
  // PRE column of each lane: columns past rcount read column r0 (those
  // accumulator columns are never stored)
  const bool has_pre = a.pre != nullptr;
  const float* pre_base = has_pre ? a.pre : kPreOnes;
  const i64 pre_stride = has_pre ? a.R : 0;
  int poff[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) poff[nt] = (has_pre ? a.r0 : 0) + (16 * nt + c < a.rcount ? 16 * nt + c : 0);

  BlockCursor<1, KB, CASE_B> cur(a, s_lo + wb0 / nkb, w, nkb, (int)(wb0 % nkb));

  auto load = [&](int t, Buf& b) {
    b.kb = cur.kb;
    const float* prow = pre_base + cur.slow * pre_stride;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) b.pre[nt] = prow[poff[nt]];
    // past the last block every lane reads the first elements of X (one cache
    // line, inside X for every shape this kernel takes): the load stays
    // unconditional and costs the memory pipeline almost nothing
    const bool live = t < T;
    const float* ubase = X + (live ? cur.offset(a) : 0);
    // XL > 0: elements at k >= kv are not loaded (their address becomes X) and
    // are zeroed in compute(), where the value is used anyway (a select here
    // would wait for the load and drain the prefetch)
    // XL 4 (a zero-padded copy): no masks, the pads are zero
    uint32_t okm = 0xffu;
    if constexpr (XL > 0 && XL != 4) {
      const int k0 = b.kb * KB + 4 * g;
      okm = 0;
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int sidx = 0; sidx < 4; ++sidx)
          okm |= (live && k0 + 16 * h + sidx < kv) ? 1u << (4 * h + sidx) : 0u;
    }
    b.okm = okm;
    if constexpr (!CASE_B) {
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        const float* base = ubase + (live ? loff[mt] : 0);
        if constexpr (XL == 0 || XL == 4) {
          b.x[mt][0] = *reinterpret_cast<const floatx4*>(base);
          b.x[mt][1] = *reinterpret_cast<const floatx4*>(base + 16);
        } else if constexpr (XL == 1) {
          // kv and the lane's k are multiples of 4: a vector is all in or all out
#pragma unroll
          for (int h = 0; h < 2; ++h)
            b.x[mt][h] = *reinterpret_cast<const floatx4*>(((okm >> (4 * h)) & 1u) ? base + 16 * h : X);
        } else if constexpr (XL == 3) {
          // an even fastest mode (and row length) on an 8-byte aligned X.  A block
          // wholly below kv (a wave-uniform test): one 16-byte vector per h, as a
          // copy of 8-byte aligned memory (the compiler emits the 16-byte load,
          // which this hardware serves at any 4-byte alignment).  The window's
          // last, partial block: pairs, all in or all out, so that no load
          // reaches past the end of a row (or of X).
          if (b.kb * KB + KB <= kv) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              floatx4 v;
              __builtin_memcpy(&v, __builtin_assume_aligned(base + 16 * h, 8), sizeof(v));
              b.x[mt][h] = v;
            }
          } else {
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
              for (int pr = 0; pr < 2; ++pr) {
                const float2 v =
                    *reinterpret_cast<const float2*>(((okm >> (4 * h + 2 * pr)) & 1u) ? base + 16 * h + 2 * pr : X);
                b.x[mt][h][2 * pr] = v.x;
                b.x[mt][h][2 * pr + 1] = v.y;
              }
          }
        } else {
#pragma unroll
          for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int sidx = 0; sidx < 4; ++sidx)
              b.x[mt][h][sidx] = *(((okm >> (4 * h + sidx)) & 1u) ? base + 16 * h + sidx : X);
        }
      }
    } else {
      const float* base = ubase + (live ? loff[0] : 0);
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int sidx = 0; sidx < 4; ++sidx) {
          const float* p = XL == 0 ? base + (i64)(16 * h + sidx) * a.I
                                   : (((okm >> (4 * h + sidx)) & 1u) ? base + (i64)(16 * h + sidx) * a.I : X);
          if constexpr (MT == 1) {
            b.x[0][h][sidx] = p[0];
          } else if constexpr (MT == 2) {
            const float2 v = *reinterpret_cast<const float2*>(p);
            b.x[0][h][sidx] = v.x;
            b.x[1][h][sidx] = v.y;
          } else {
            const floatx4 v = *reinterpret_cast<const floatx4*>(p);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) b.x[mt][h][sidx] = v[mt];
          }
        }
    }
    if (t + 1 < T) cur.step(a);
  };

  auto compute = [&](int t, const Buf& b) {
    if (t >= T) return;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      floatx4 xm[MT];
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        xm[mt] = b.x[mt][h];
        if constexpr (XL > 0 && XL != 4) {
#pragma unroll
          for (int sidx = 0; sidx < 4; ++sidx)
            if (!((b.okm >> (4 * h + sidx)) & 1u)) xm[mt][sidx] = 0.f;
        }
      }
      floatx4 bv[NT];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        bv[nt] = *reinterpret_cast<const floatx4*>(&sF[(16 * nt + c) * PITCH + b.kb * KB + 16 * h + 4 * g]) *
                 b.pre[nt];
#pragma unroll
      for (int sidx = 0; sidx < 4; ++sidx)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
          for (int mt = 0; mt < MT; ++mt)
            acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(xm[mt][sidx], bv[nt][sidx], acc[mt][nt], 0, 0, 0);
    }
  };

  Buf b0, b1, b2, b3;
  load(0, b0);
  load(1, b1);
  load(2, b2);
  if constexpr (XL > 0) window_zero_tail(fw, kv);
  stage_window_store<RT, PITCH, NTH, KWMAX>(sF, fw);
  __syncthreads();  // LDS window staged
  for (int t = 0; t < T; t += 4) {
    load(t + 3, b3);
    compute(t, b0);
    load(t + 4, b0);
    compute(t + 1, b1);
    load(t + 5, b1);
    compute(t + 2, b2);
    load(t + 6, b2);
    compute(t + 3, b3);
  }

  wg_reduce_store<MT, NT, WAVES>(acc, smem, wave, lane, partial + (i64)split * a.I * a.R, a.I, a.R, a.r0, a.rcount,
                                 row_base, CASE_B);
}

// ---------------------------------------------------------------------------
// BF16 MFMA kernel (v_mfma_f32_16x16x32_bf16), case A only (q-contiguous X; the
// last mode runs on the last-to-front copy).  Same work split and the same
// T-then-PRE contraction as the FP32 kernel.  Accumulation is FP32.
//   XIN = uint16_t: X enters as bf16 (one 16-byte load of 8 consecutive j per
//     lane and k-step); the FAST window enters as an exact bf16 hi + lo pair (two
//     MFMAs per step), so the input tensor is the only operand rounded to bf16.
//   XIN = float ("bf16x3"): X enters as fp32 (two 16-byte loads per lane and
//     k-step) and is split in registers into an exact bf16 triple, as is the FAST
//     window; six MFMAs per step form every product term down to 2^-16 of the
//     leading one (x_h f_h; x_h f_l, x_l f_h; x_h f_l2, x_l f_l, x_l2 f_h), so the
//     result has the accuracy of an FP32 MTTKRP of the fp32 tensor.  The BF16 ALS
//     runs its last mode, whose <M, A> gives the fit, this way.
// ---------------------------------------------------------------------------
template <typename XIN, int NT, int WAVES, int MT, int KS, int KWMAX, bool WV, int XL>
__global__ __launch_bounds__(WAVES * 64) void mttkrp_bf16_mfma(const XIN* __restrict__ X, KernArgs a,
                                                               float* __restrict__ partial, AlsJob job, int has_job) {
  constexpr bool X3 = std::is_same<XIN, float>::value;
  static_assert(!X3 || KS == 1, "fp32 input uses one 32-wide k-step per block (register budget)");
  constexpr int NWIN = X3 ? 3 : 2;  // bf16 parts of the FAST window
  constexpr int KB = 32 * KS, PITCH = KWMAX + 8, RT = 16 * NT, NTH = WAVES * 64, ROWS = 16 * MT;
  constexpr int LDS_WIN_U16 = NWIN * RT * PITCH, LDS_RED_F = MT * NT * WAVES * 64 * 4;
  constexpr int LDS_B = (LDS_WIN_U16 * 2 > LDS_RED_F * 4) ? LDS_WIN_U16 * 2 : LDS_RED_F * 4;
  constexpr int LDS_BYTES = LDS_B > JOB_LDS_BYTES ? LDS_B : JOB_LDS_BYTES;
  __shared__ __attribute__((aligned(16))) float smem[LDS_BYTES / 4];
  static_assert(NTH % JOB_THREADS == 0, "the ALS job runs on the first 256 threads");
  if (has_job && blockIdx.x == 0) {
    // the job's workgroup: waves past the job's four end at once (a barrier
    // no longer waits for an ended wave); with one such workgroup per CU the
    // job has its CU to itself
    if (threadIdx.x < JOB_THREADS) als_job_run(job, reinterpret_cast<double*>(smem));
    return;
  }
  if (has_job && job.gate) {
    // start gate: wait (a bounded number of polls) until the job has its Grams
    if (threadIdx.x == 0) {
      for (int it = 0; it < JOB_GATE_POLLS; ++it) {
        if (__hip_atomic_load(job.gate, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == job.gate_stamp) break;
        __builtin_amdgcn_s_sleep(8);
      }
    }
    __syncthreads();
  }
  const int tile = (int)blockIdx.x - (has_job ? 1 : 0);
  const int xt = tile % a.rtiles;
  uint16_t* sHi = reinterpret_cast<uint16_t*>(smem);
  uint16_t* sLo = sHi + RT * PITCH;
  uint16_t* sLo2 = sLo + RT * PITCH;  // X3 only
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar schedule and addresses
  const int g = lane >> 4, c = lane & 15;
  const int split = tile / a.rtiles;
  const int w = split % a.nwin;
  const int ss = split / a.nwin;
  // balanced split of the slow range: the first slow_rem splits take one more
  const i64 s_lo = uniform_i64((i64)ss * a.slow_q + (ss < a.slow_rem ? ss : a.slow_rem));
  const i64 s_hi = s_lo + a.slow_q + (ss < a.slow_rem ? 1 : 0);
  // rows of this window that exist (XL == 1: a ragged fastest mode, whose loads
  // are masked past kv; kv is a multiple of the 8- (bf16) or 4-element (fp32)
  // load vector)
  const int kv = XL == 0 ? a.kw : (int)min((i64)a.kw, a.fast_size - (i64)w * a.kw);
  const int nkb = XL == 0 ? a.kw / KB : (kv + KB - 1) / KB;
  // the workgroup's blocks (slow index, then block within it) in WAVES contiguous,
  // balanced ranges: with 9 slow indices of 4 blocks, 9 blocks per wave rather
  // than 12 for one wave and 8 for the others
  const i64 nblk_wg = (s_hi - s_lo) * nkb;
  const i64 wb0 = uniform_i64(wave * nblk_wg / WAVES), wb1 = (wave + 1) * nblk_wg / WAVES;
  const int T = __builtin_amdgcn_readfirstlane((int)(wb1 - wb0));  // blocks of this wave
  const i64 row_base = (i64)xt * ROWS;

  FastWindow<RT, NTH, KWMAX> fw;  // window loads before the X blocks
  stage_window_load<WV>(fw, a, w, kv);

  i64 rowc[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const i64 rr = row_base + 16 * mt + c;
    rowc[mt] = rr < a.I ? rr : a.I - 1;
  }
  floatx4 acc[MT][NT], Tacc[MT][NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      acc[mt][nt] = floatx4{0.f, 0.f, 0.f, 0.f};
      Tacc[mt][nt] = floatx4{0.f, 0.f, 0.f, 0.f};
    }

  using XV = std::conditional_t<X3, floatx8, bf16x8>;
  struct Buf {
    XV x[MT][KS];
    float pre[NT];
    int kb;
    uint32_t okm;  // XL == 1: bit 2ks + v set where vector v of k-step ks lies below kv
  };

  // per-lane element offsets, fixed for the whole loop; the per-block part of
  // every address is wave-uniform
  i64 loff[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) loff[mt] = rowc[mt] * a.Q + 8 * g;
  // PRE column of each lane: columns past rcount read column r0 (those
  // accumulator columns are never stored)
  const bool has_pre = a.pre != nullptr;
  const float* pre_base = has_pre ? a.pre : kPreOnes;
  const i64 pre_stride = has_pre ? a.R : 0;
  int poff[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) poff[nt] = (has_pre ? a.r0 : 0) + (16 * nt + c < a.rcount ? 16 * nt + c : 0);

  BlockCursor<1, KB, false> cur(a, s_lo + wb0 / nkb, w, nkb, (int)(wb0 % nkb));

  auto load = [&](int t, Buf& b) {
    b.kb = cur.kb;
    const float* prow = pre_base + cur.slow * pre_stride;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) b.pre[nt] = prow[poff[nt]];
    const bool live = t < T;  // past the end: one cache line at X (see the FP32 kernel)
    const XIN* ubase = X + (live ? cur.offset(a) : 0);
    // XL == 1: vectors at k >= kv are loaded from X instead and zeroed in compute();
    // XL == 4 (a zero-padded copy): no masks
    uint32_t okm = 0xffu;
    if constexpr (XL > 0 && XL != 4) {
      okm = 0;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int v = 0; v < 2; ++v)
          okm |= (live && b.kb * KB + 32 * ks + 8 * g + 4 * v < kv) ? 1u << (2 * ks + v) : 0u;
    }
    b.okm = okm;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const XIN* base = ubase + (live ? loff[mt] : 0);
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        if constexpr (X3) {
          const floatx4 v0 = *reinterpret_cast<const floatx4*>(((okm >> (2 * ks)) & 1u) ? base + 32 * ks : X);
          const floatx4 v1 =
              *reinterpret_cast<const floatx4*>(((okm >> (2 * ks + 1)) & 1u) ? base + 32 * ks + 4 : X);
          b.x[mt][ks] = floatx8{v0[0], v0[1], v0[2], v0[3], v1[0], v1[1], v1[2], v1[3]};
        } else {
          b.x[mt][ks] = *reinterpret_cast<const bf16x8*>(((okm >> (2 * ks)) & 1u) ? base + 32 * ks : X);
        }
      }
    }
    if (t + 1 < T) cur.step(a);
  };

  auto compute = [&](int t, const Buf& b) {
    if (t >= T) return;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      if constexpr (X3) {
        bf16x8 xh[MT], xl[MT], xl2[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
          floatx8 xv = b.x[mt][ks];
          if constexpr (XL > 0 && XL != 4) {
#pragma unroll
            for (int e = 0; e < 8; ++e)
              if (!((b.okm >> (2 * ks + e / 4)) & 1u)) xv[e] = 0.f;
          }
          split_bf16x3(xv, xh[mt], xl[mt], xl2[mt]);
        }
        // Two rank tiles at a time, term by term (smallest first): consecutive
        // MFMAs then update 2 * MT different accumulators, so none waits for the
        // previous one's result (one accumulator taking all six terms in a row
        // serialised the matrix pipe on its latency)
        constexpr int NG = NT >= 2 ? 2 : 1;
#pragma unroll
        for (int n0 = 0; n0 < NT; n0 += NG) {
          bf16x8 bh[NG], bl[NG], bl2[NG];
#pragma unroll
          for (int u = 0; u < NG; ++u) {
            const int off = (16 * (n0 + u) + c) * PITCH + b.kb * KB + 32 * ks + 8 * g;
            bh[u] = *reinterpret_cast<const bf16x8*>(&sHi[off]);
            bl[u] = *reinterpret_cast<const bf16x8*>(&sLo[off]);
            bl2[u] = *reinterpret_cast<const bf16x8*>(&sLo2[off]);
          }
#define TENRSO_X3_TERM(XA, XB)                                                                        \
  _Pragma("unroll") for (int u = 0; u < NG; ++u) _Pragma("unroll") for (int mt = 0; mt < MT; ++mt)     \
      Tacc[mt][n0 + u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(XA[mt], XB[u], Tacc[mt][n0 + u], 0, 0, 0);
          TENRSO_X3_TERM(xl2, bh)
          TENRSO_X3_TERM(xl, bl)
          TENRSO_X3_TERM(xh, bl2)
          TENRSO_X3_TERM(xl, bh)
          TENRSO_X3_TERM(xh, bl)
          TENRSO_X3_TERM(xh, bh)
#undef TENRSO_X3_TERM
        }
      } else {
        bf16x8 xv[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
          xv[mt] = b.x[mt][ks];
          if constexpr (XL > 0 && XL != 4) {
            if (!((b.okm >> (2 * ks)) & 1u)) xv[mt] = bf16x8{};
          }
        }
        // two rank tiles at a time, the hi term of all four accumulators before
        // their lo term (no MFMA waits for the one just issued)
        constexpr int NG = NT >= 2 ? 2 : 1;
#pragma unroll
        for (int n0 = 0; n0 < NT; n0 += NG) {
          bf16x8 bh[NG], bl[NG];
#pragma unroll
          for (int u = 0; u < NG; ++u) {
            const int off = (16 * (n0 + u) + c) * PITCH + b.kb * KB + 32 * ks + 8 * g;
            bh[u] = *reinterpret_cast<const bf16x8*>(&sHi[off]);
            bl[u] = *reinterpret_cast<const bf16x8*>(&sLo[off]);
          }
#pragma unroll
          for (int u = 0; u < NG; ++u)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
              Tacc[mt][n0 + u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xv[mt], bh[u], Tacc[mt][n0 + u], 0, 0, 0);
#pragma unroll
          for (int u = 0; u < NG; ++u)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
              Tacc[mt][n0 + u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xv[mt], bl[u], Tacc[mt][n0 + u], 0, 0, 0);
        }
      }
    }
    // a slow index ends, or this wave's range does (it may end inside one)
    if (b.kb == nkb - 1 || t == T - 1) fold_slow<MT, NT>(acc, Tacc, b.pre);
  };

  Buf b0, b1, b2, b3;
  load(0, b0);
  load(1, b1);
  load(2, b2);
  if constexpr (XL > 0) window_zero_tail(fw, kv);
  if constexpr (X3) stage_window_store_bf16x3<RT, PITCH, NTH, KWMAX>(sHi, sLo, sLo2, fw);
  else stage_window_store_bf16x2<RT, PITCH, NTH, KWMAX>(sHi, sLo, fw);
  __syncthreads();  // LDS window staged
  for (int t = 0; t < T; t += 4) {
    load(t + 3, b3);
    compute(t, b0);
    load(t + 4, b0);
    compute(t + 1, b1);
    load(t + 5, b1);
    compute(t + 2, b2);
    load(t + 6, b2);
    compute(t + 3, b3);
  }

  wg_reduce_store<MT, NT, WAVES>(acc, smem, wave, lane, partial + (i64)split * a.I * a.R, a.I, a.R, a.r0, a.rcount,
                                 row_base, false);
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
// A workgroup owns 64 consecutive outputs and splits the split range into
// RED_GROUPS contiguous pieces, one per group of 64 threads (each lane one
// output, so every load of a group is one contiguous 256-byte row segment).
// Each group sums its piece in FP64 with four independent chains; the group
// sums are then added in group order.  The order is fixed, so the result is
// run-to-run identical.  (One thread per output looping over all splits left
// 12 workgroups for a 96^4 rank-32 result, a latency chain of ~170 loads.)
constexpr int RED_GROUPS = 16;
__global__ __launch_bounds__(64 * RED_GROUPS) void reduce_partials_kernel(const float* __restrict__ part, int splits,
                                                                          i64 n, float* __restrict__ o32,
                                                                          double* __restrict__ o64) {
  __shared__ double red[RED_GROUPS][64];
  const int lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
  const i64 idx = (i64)blockIdx.x * 64 + lane;
  double s = 0.0;
  if (idx < n) {
    const int s0 = (int)((i64)grp * splits / RED_GROUPS), s1 = (int)((i64)(grp + 1) * splits / RED_GROUPS);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int sp = s0;
    for (; sp + 4 <= s1; sp += 4) {
      a0 += (double)part[(i64)(sp + 0) * n + idx];
      a1 += (double)part[(i64)(sp + 1) * n + idx];
      a2 += (double)part[(i64)(sp + 2) * n + idx];
      a3 += (double)part[(i64)(sp + 3) * n + idx];
    }
    for (; sp < s1; ++sp) a0 += (double)part[(i64)sp * n + idx];
    s = (a0 + a1) + (a2 + a3);
  }
  red[grp][lane] = s;
  __syncthreads();
  if (grp == 0 && idx < n) {
    double t = red[0][lane];
#pragma unroll
    for (int g = 1; g < RED_GROUPS; ++g) t += red[g][lane];
    if (o32) o32[idx] = (float)t;
    if (o64) o64[idx] = t;
  }
}

// vec: src 16-byte and dst 8-byte aligned (checked on the host), else scalar
__global__ __launch_bounds__(256) void f32_to_bf16_kernel(const float* __restrict__ src, i64 n,
                                                          uint16_t* __restrict__ dst, int vec) {
  const i64 i4 = ((i64)blockIdx.x * 256 + threadIdx.x) * 4;
  if (vec && i4 + 3 < n) {
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

// [rows, K] -> [K, rows] transpose through a 32x32 LDS tile: fp32 -> bf16 (RNE),
// bf16 -> bf16 or fp32 -> fp32.
__device__ __forceinline__ uint16_t to_out(float v, uint16_t*) { return f32_to_bf16_rne(v); }
__device__ __forceinline__ uint16_t to_out(uint16_t v, uint16_t*) { return v; }
__device__ __forceinline__ float to_out(float v, float*) { return v; }

// With B < Bp the output rows are padded: input row r = a*B + b lands at
// a*Bp + b of an output row of extent (rows / B) * Bp (the pad entries are
// zero-filled beforehand); B = Bp leaves the layout plain.
template <typename TIN, typename TOUT>
__global__ __launch_bounds__(256) void transpose_kernel(const TIN* __restrict__ src, i64 rows, i64 K,
                                                        TOUT* __restrict__ dst, i64 B = 1, i64 Bp = 1) {
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
    if (k < K && r < rows) {
      const i64 o = B == Bp ? k * rows + r : k * (rows / B * Bp) + (r / B) * Bp + r % B;
      dst[o] = to_out(tile[tx][yy], dst);
    }
  }
}

// Forward-order Khatri-Rao of n factors: row = (i_0, ..., i_{n-1}) last fastest.
struct KrManyArgs {
  int n;
  i64 R, rows;
  i64 sizes[MAXD];
  const float* F[MAXD];
};
__device__ __forceinline__ void kr_many_body(const KrManyArgs& a, float* __restrict__ dst) {
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
__global__ __launch_bounds__(256) void kr_many_kernel(KrManyArgs a, float* __restrict__ dst) { kr_many_body(a, dst); }

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static inline unsigned cdiv(i64 a, i64 b) { return (unsigned)((a + b - 1) / b); }

int reduce_partials(tenrso_ctx* ctx, const float* partial, int splits, i64 n, float* out32, double* out64) {
  if (n == 0) return TENRSO_OK;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(cdiv(n, 64)), dim3(64 * RED_GROUPS), 0, ctx->stream, partial,
                     splits, n, out32, out64);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

// Several Khatri-Rao products in one launch: product y of grid.y into dst[y].
constexpr int KR_BATCH = MAXD;
struct KrBatchArgs {
  int count;
  KrManyArgs k[KR_BATCH];
  float* dst[KR_BATCH];
};
__global__ __launch_bounds__(256) void kr_batch_kernel(KrBatchArgs b) {
  const int y = blockIdx.y;
  if (y >= b.count) return;
  kr_many_body(b.k[y], b.dst[y]);
}

int mttkrp_pre_batch(tenrso_ctx* ctx, int nd, const i64* shape, const float* const* F, i64 R, int nmodes,
                     const int* modes, bool via_copy_last, const float** pre) {
  KrBatchArgs b{};
  int owner[KR_BATCH];  // the product each requested mode uses, or -1
  i64 offs[KR_BATCH + 1] = {0};
  for (int i = 0; i < nmodes; ++i) {
    pre[i] = nullptr;
    owner[i] = -1;
    const int mode = modes[i];
    // the PRE factors: every other mode but the fastest, in order (the last
    // mode through the copy contracts modes 0 .. nd-3 the same way)
    KrManyArgs k{};
    k.R = R;
    k.rows = 1;
    const bool copy = via_copy_last && mode == nd - 1;
    const int n_oth = nd - 1;
    if (n_oth < 3) continue;
    for (int j = 0; j < nd && k.n < n_oth - 1; ++j) {
      if (!copy && j == mode) continue;
      k.F[k.n] = F[j];
      k.sizes[k.n] = shape[j];
      k.rows *= shape[j];
      ++k.n;
    }
    // shared with an earlier product of the same factors?
    for (int q = 0; q < b.count && owner[i] < 0; ++q) {
      bool same = b.k[q].n == k.n;
      for (int j = 0; same && j < k.n; ++j) same = b.k[q].F[j] == k.F[j] && b.k[q].sizes[j] == k.sizes[j];
      if (same) owner[i] = q;
    }
    if (owner[i] < 0) {
      if (b.count == KR_BATCH) continue;  // more distinct products than one launch takes: the MTTKRP forms its own
      owner[i] = b.count;
      b.k[b.count] = k;
      offs[b.count + 1] = offs[b.count] + k.rows * R;
      ++b.count;
    }
  }
  if (b.count == 0) return TENRSO_OK;
  void* buf = nullptr;
  TENRSO_TRY(ctx_buffer(ctx, SLOT_PRE, (size_t)offs[b.count] * sizeof(float), &buf));
  i64 maxn = 0;
  for (int q = 0; q < b.count; ++q) {
    b.dst[q] = (float*)buf + offs[q];
    maxn = std::max(maxn, b.k[q].rows * R);
  }
  for (int i = 0; i < nmodes; ++i)
    if (owner[i] >= 0) pre[i] = b.dst[owner[i]];
  hipLaunchKernelGGL(kr_batch_kernel, dim3(cdiv(maxn, 256), b.count), dim3(256), 0, ctx->stream, b);
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
    const int vec = (uintptr_t)src % 16 == 0 && (uintptr_t)dst % 8 == 0;
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(cdiv(cdiv(total, 4), 256)), dim3(256), 0, ctx->stream, src, total,
                       dst, vec);
  } else {
    const i64 K = shape[nd - 1], rows = total / K;
    hipLaunchKernelGGL((transpose_kernel<float, uint16_t>), dim3(cdiv(K, 32), cdiv(rows, 32)), dim3(256), 0,
                       ctx->stream, src, rows, K, dst);
  }
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

// [rows, K] fp32 -> [rows, Kp] bf16 (RNE), zero in the pad columns k >= K
// one workgroup per row (no division per element)
__global__ __launch_bounds__(256) void f32_to_bf16_pad_kernel(const float* __restrict__ src, i64 K, i64 Kp,
                                                              uint16_t* __restrict__ dst) {
  const i64 r = blockIdx.x;
  for (i64 k = threadIdx.x; k < Kp; k += 256) dst[r * Kp + k] = k < K ? f32_to_bf16_rne(src[r * K + k]) : (uint16_t)0;
}

int convert_bf16_padded(tenrso_ctx* ctx, const float* src, int nd, const i64* shape, i64 Kp, uint16_t* dst) {
  i64 total = 1;
  for (int k = 0; k < nd; ++k) total *= shape[k];
  const i64 K = shape[nd - 1];
  if (total == 0) return TENRSO_OK;
  const i64 rows = total / K;
  hipLaunchKernelGGL(f32_to_bf16_pad_kernel, dim3((unsigned)rows), dim3(256), 0, ctx->stream, src, K, Kp, dst);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

// one workgroup per row (no division per element)
__global__ __launch_bounds__(256) void f32_pad_kernel(const float* __restrict__ src, i64 K, i64 Kp,
                                                      float* __restrict__ dst) {
  const i64 r = blockIdx.x;
  for (i64 k = threadIdx.x; k < Kp; k += 256) dst[r * Kp + k] = k < K ? src[r * K + k] : 0.0f;
}

// zeros at [B, Bp) of each of n rows of Bp (the pads of a padded copy, nothing else)
__global__ __launch_bounds__(256) void zero_pads_kernel(float* __restrict__ dst, i64 n, i64 B, i64 Bp) {
  const i64 np = Bp - B;
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e >= n * np) return;
  const i64 r = e / np;
  dst[r * Bp + B + (e - r * np)] = 0.0f;
}

int copy_f32_padded(tenrso_ctx* ctx, const float* src, int nd, const i64* shape, i64 Kp, float* dst) {
  i64 total = 1;
  for (int k = 0; k < nd; ++k) total *= shape[k];
  const i64 K = shape[nd - 1];
  if (total == 0) return TENRSO_OK;
  const i64 rows = total / K;
  hipLaunchKernelGGL(f32_pad_kernel, dim3((unsigned)rows), dim3(256), 0, ctx->stream, src, K, Kp, dst);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

int transpose_last_to_front_f32_padded(tenrso_ctx* ctx, const float* src, int nd, const i64* shape, i64 Bp,
                                       float* dst) {
  i64 total = 1;
  for (int k = 0; k < nd; ++k) total *= shape[k];
  if (total == 0) return TENRSO_OK;
  const i64 K = shape[nd - 1], rows = total / K, B = shape[nd - 2];
  const i64 nrow = K * (rows / B);  // rows of Bp in the copy
  if (Bp != B)
    hipLaunchKernelGGL(zero_pads_kernel, dim3(cdiv(nrow * (Bp - B), 256)), dim3(256), 0, ctx->stream, dst, nrow, B,
                       Bp);
  hipLaunchKernelGGL((transpose_kernel<float, float>), dim3(cdiv(K, 32), cdiv(rows, 32)), dim3(256), 0, ctx->stream,
                     src, rows, K, dst, B, Bp);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

int transpose_last_to_front_f32(tenrso_ctx* ctx, const float* src, int nd, const i64* shape, float* dst) {
  i64 total = 1;
  for (int k = 0; k < nd; ++k) total *= shape[k];
  if (total == 0) return TENRSO_OK;
  const i64 K = shape[nd - 1], rows = total / K;
  hipLaunchKernelGGL((transpose_kernel<float, float>), dim3(cdiv(K, 32), cdiv(rows, 32)), dim3(256), 0, ctx->stream,
                     src, rows, K, dst);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

// One pass over an fp32 tensor viewed as [rows, K] (K = last mode) for the BF16
// ALS: the bf16 copy (same layout), the last-to-front fp32 copy ([K, rows]; the
// last mode's MTTKRP contracts the exact tensor, see mttkrp_bf16_mfma) and
// per-workgroup (sum of squares, min) partials, instead of three passes.  64 x 64
// tiles, grid-stride over tiles; each thread loads 4 x 16 bytes (K % 4 == 0, X
// 16-byte aligned), writes its bf16 quads straight out and through an LDS tile
// the transposed fp32 quads.  The bf16 values are f32_to_bf16_rne, as in the
// separate conversion kernel.
__global__ __launch_bounds__(256) void prepare_bf16_kernel(const float* __restrict__ X, i64 rows, i64 K,
                                                           uint16_t* __restrict__ dst, float* __restrict__ dstT,
                                                           double* __restrict__ part) {
  __shared__ float tile[64][64 + 1];
  __shared__ double red[256];
  const int tid = threadIdx.x;
  const i64 tr_n = (rows + 63) / 64, tk_n = (K + 63) / 64;
  double ss = 0.0;
  float mn = FLT_MAX;
  for (i64 t = blockIdx.x; t < tr_n * tk_n; t += gridDim.x) {
    const i64 r0 = (t / tk_n) * 64, k0 = (t % tk_n) * 64;
    // load + direct copy: thread -> (row tid / 16 + 16 j, columns 4 (tid % 16) ..)
    const int c4 = 4 * (tid & 15);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int rr = (tid >> 4) + 16 * j;
      const i64 r = r0 + rr, k = k0 + c4;
      floatx4 v = {0.f, 0.f, 0.f, 0.f};
      if (r < rows && k < K) {
        v = *reinterpret_cast<const floatx4*>(X + r * K + k);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          ss = fma((double)v[q], (double)v[q], ss);
          mn = fminf(mn, v[q]);
        }
        if (dst) {  // null: the transposed copy and the norm only
          ushort4 o;
          o.x = f32_to_bf16_rne(v[0]);
          o.y = f32_to_bf16_rne(v[1]);
          o.z = f32_to_bf16_rne(v[2]);
          o.w = f32_to_bf16_rne(v[3]);
          *reinterpret_cast<ushort4*>(dst + r * K + k) = o;
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) tile[rr][c4 + q] = v[q];
    }
    __syncthreads();
    // transposed quads: thread -> (column kk = tid / 16 + 16 j, rows 4 (tid % 16) ..)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int kk = (tid >> 4) + 16 * j, r4 = 4 * (tid & 15);
      const i64 k = k0 + kk, r = r0 + r4;
      if (k < K && r < rows) {
        const floatx4 o = {tile[r4][kk], tile[r4 + 1][kk], tile[r4 + 2][kk], tile[r4 + 3][kk]};
        float* d = dstT + k * rows + r;
        if (r + 3 < rows && rows % 4 == 0) {
          *reinterpret_cast<floatx4*>(d) = o;
        } else {
          d[0] = o[0];
          if (r + 1 < rows) d[1] = o[1];
          if (r + 2 < rows) d[2] = o[2];
          if (r + 3 < rows) d[3] = o[3];
        }
      }
    }
    __syncthreads();  // tile reused by the next tile of this workgroup
  }
  // (sum, min) partial of the workgroup
  double v = ss;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  const double tot = (red[0] + red[1]) + (red[2] + red[3]);
  __syncthreads();
  red[tid] = (double)mn;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (tid < k) red[tid] = fmin(red[tid], red[tid + k]);
    __syncthreads();
  }
  if (tid == 0) {
    part[blockIdx.x] = tot;
    part[gridDim.x + blockIdx.x] = red[0];
  }
}

int prepare_bf16_als(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, uint16_t* dst, float* dstT,
                     double* tmp, int* nblk_out) {
  i64 total = 1;
  for (int k = 0; k < nd; ++k) total *= shape[k];
  const i64 K = nd >= 1 ? shape[nd - 1] : 1, rows = K ? total / K : 0;
  *nblk_out = 0;
  if (total == 0 || nd < 2 || K % 4 != 0 || (uintptr_t)X % 16 != 0) return TENRSO_ERR_UNSUPPORTED;
  const i64 tiles = cdiv(rows, 64) * cdiv(K, 64);
  const int nblk = (int)std::min<i64>(1024, tiles);
  hipLaunchKernelGGL(prepare_bf16_kernel, dim3(nblk), dim3(256), 0, ctx->stream, X, rows, K, dst, dstT, tmp);
  TENRSO_HIP_TRY(hipGetLastError());
  *nblk_out = nblk;
  return TENRSO_OK;
}

int transpose_last_to_front_u16(tenrso_ctx* ctx, const uint16_t* src, int nd, const i64* shape, uint16_t* dst) {
  i64 total = 1;
  for (int k = 0; k < nd; ++k) total *= shape[k];
  if (total == 0) return TENRSO_OK;
  const i64 K = shape[nd - 1], rows = total / K;
  hipLaunchKernelGGL((transpose_kernel<uint16_t, uint16_t>), dim3(cdiv(K, 32), cdiv(rows, 32)), dim3(256), 0,
                     ctx->stream, src, rows, K, dst);
  TENRSO_HIP_TRY(hipGetLastError());
  return TENRSO_OK;
}

namespace {

template <typename F>
struct KernelArgs;
template <typename... A>
struct KernelArgs<void (*)(A...)> {
  using type = std::tuple<A...>;
};

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

int mttkrp_dense(tenrso_ctx* ctx, int prec, const void* X, const void* X_t, int nd, const i64* shape,
                 const float* const* F, i64 R, int mode, MttkrpOut outk, void* out, MttkrpResult* res,
                 const AlsJob* job, bool xt_f32, const MttkrpScratch* scratch, i64 x_pad, i64 xt_pad,
                 const float* pre_rows) {
  Plan pl = make_plan(nd, shape, mode);
  const bool bf16 = prec == TENRSO_PREC_BF16;
  const float* Fv[MAXD];
  for (int k = 0; k < nd; ++k) Fv[k] = F[k];
  const void* Xuse = X;
  int mode_use = mode;
  i64 shape_use[MAXD];
  for (int k = 0; k < nd; ++k) shape_use[k] = shape[k];

  // BF16 last mode: run mode 0 of the last-to-front copy (same j order); an
  // fp32 copy is contracted exactly (bf16x3 split).  FP32 with a last-to-front
  // fp32 copy given: its last mode also runs as mode 0 of the copy (q-contiguous
  // 16-byte loads instead of the last mode's i-contiguous ones).
  bool x3 = false;
  if ((bf16 || X_t) && mode == nd - 1 && nd >= 2) {
    x3 = bf16 && xt_f32;
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
  size_t scratch_used = 0;  // bytes of the caller's workspace taken by PRE
  // the most split partials that fit in the caller's workspace (or no bound)
  auto max_splits = [&]() -> i64 {
    if (!scratch) return INT64_MAX;
    const size_t per = (size_t)std::max<i64>(1, nout) * sizeof(float);
    return scratch->bytes > scratch_used ? (i64)((scratch->bytes - scratch_used) / per) : 0;
  };
  auto partial_buffer = [&](int nsplits, float** p) -> int {
    if (scratch) {
      *p = reinterpret_cast<float*>(static_cast<char*>(scratch->base) + scratch_used);
      return TENRSO_OK;
    }
    void* pbuf = nullptr;
    TENRSO_TRY(ctx_buffer(ctx, SLOT_PARTIAL, (size_t)nsplits * nout * sizeof(float), &pbuf));
    *p = (float*)pbuf;
    return TENRSO_OK;
  };
  auto too_small = [&](i64 need_splits) {
    set_error("workspace of %zu bytes is too small for this MTTKRP (needs %zu bytes)", scratch->bytes,
              scratch_used + (size_t)need_splits * (size_t)std::max<i64>(1, nout) * sizeof(float));
    return TENRSO_ERR_UNSUPPORTED;
  };
  if (pl.n_oth >= 3 && pre_rows) {
    pre = pre_rows;
  } else if (pl.n_oth >= 3) {
    const float* pf[MAXD];
    i64 ps[MAXD];
    i64 prows = 1;
    for (int m = 0; m < pl.n_oth - 1; ++m) {
      pf[m] = Fv[pl.oth[m]];
      ps[m] = pl.oth_size[m];
      prows *= ps[m];
    }
    void* buf = nullptr;
    const size_t pre_bytes = (size_t)(prows * R) * sizeof(float);
    if (scratch) {
      if (pre_bytes > scratch->bytes) {
        set_error("workspace of %zu bytes cannot hold the %zu bytes of Khatri-Rao rows", scratch->bytes, pre_bytes);
        return TENRSO_ERR_UNSUPPORTED;
      }
      buf = scratch->base;
      scratch_used = (pre_bytes + 255) / 256 * 256;
    } else {
      TENRSO_TRY(ctx_buffer(ctx, SLOT_TMP, pre_bytes, &buf));
    }
    TENRSO_TRY(khatri_rao_many(ctx, pl.n_oth - 1, pf, ps, R, (float*)buf));
    pre = (const float*)buf;
  } else if (pl.n_oth == 2) {
    pre = Fv[pl.oth[0]];
  }
  const float* fastf = pl.n_oth ? Fv[pl.oth[pl.n_oth - 1]] : nullptr;

  const bool caseB = pl.Q == 1;
  // memory extent of the fastest other mode: the BF16 ALS copies of a tensor
  // whose last mode is no multiple of the load vector have that mode padded
  // (zeros), so every row starts on a vector boundary
  const i64 pad = mode == nd - 1 && nd >= 2 ? xt_pad : x_pad;
  const i64 fast_stride = (pad > 0 && !caseB) ? pad : pl.fast_size;
  const i64 Qmem = (pad > 0 && !caseB) ? pl.Q / pl.fast_size * pad : pl.Q;

  KernArgs a{};
  a.I = I;
  a.Q = Qmem;
  a.R = R;
  a.fast_size = pl.fast_size;
  a.fast_stride = fast_stride;
  a.pre = pre;
  a.fast = fastf;

  const int KB = (bf16 && !x3) ? ((pl.fast_size % 64 == 0) ? 64 : 32) : 32;
  constexpr int KWMAX = 128;  // LDS window bound (fp32: 64 x 132 floats)
  // LDS window of the fastest mode: the largest multiple of KB that divides it.
  // A fastest mode no such window divides ("ragged", e.g. 100 or 250) takes
  // windows of a multiple of 32 rows covering it as evenly as possible, the
  // last one partly filled, with the loads past its end masked (XL > 0).
  int kw = 0;
  for (i64 cand = std::min<i64>(KWMAX, pl.fast_size) / KB * KB; cand >= KB; cand -= KB)
    if (pl.fast_size % cand == 0) {
      kw = (int)cand;
      break;
    }
  int XL = 0;
  if (kw == 0 && pl.fast_size > 0) {
    const i64 nw = (pl.fast_size + KWMAX - 1) / KWMAX;
    kw = (int)(((pl.fast_size + nw - 1) / nw + 31) / 32 * 32);
    // masked vectors need the fastest mode (and so every row start) to be a
    // multiple of the load vector: 4 fp32 or 8 bf16 elements; otherwise fp32
    // loads element by element (XL 2) and bf16 takes the generic kernel
    const int vecw = (bf16 && !x3) ? 8 : 4;
    XL = (caseB || fast_stride % vecw == 0) ? 1 : 2;
  }
  // The MFMA kernels read X with 16-byte vectors (case A) or MT-wide vectors
  // along i (case B).  An fp32 view that is only 4-byte aligned (a slice at an
  // odd offset) loads element-wise in case A (XL 2 also takes windows that
  // divide the mode: all its masks are set) and one row per lane in case B; the
  // bf16 kernels need 16-byte alignment, else the generic kernel runs.
  const bool aligned16 = (uintptr_t)Xuse % 16 == 0;
  if (!bf16 && !caseB && !aligned16 && kw > 0) XL = 2;
  // Case A on a zero-padded copy whose rows reach a multiple of 32: every
  // window's last block ends inside the padded row, so X loads need no masks
  // (the pads are zero, and window rows past kv are staged as zeros); only the
  // window load keeps kv (XL 4)
  if (!caseB && pad > 0 && XL == 1 && aligned16 && fast_stride >= (pl.fast_size + 31) / 32 * 32) XL = 4;
  // element-wise fp32 loads become pairs (8 bytes) where every row start is even
  // and X is 8-byte aligned (a fastest mode of 250, say)
  if (XL == 2 && fast_stride % 2 == 0 && Qmem % 2 == 0 && (uintptr_t)Xuse % 8 == 0) XL = 3;
  const bool mfma_ok = pl.n_oth >= 1 && pl.J > 0 && kw > 0 && !(bf16 && caseB) && I > 0 &&
                       (uintptr_t)Xuse % 4 == 0 &&
                       (bf16 ? aligned16 && (XL < 2 || XL == 4) : (caseB || XL >= 2 || aligned16));

  float* partial = nullptr;
  int splits = 1;
  if (mfma_ok) {
    // FP32: 4-wave workgroups, two per CU.  BF16: 8-wave workgroups, one per CU
    // (the VGPR budget admits one), so a fused ALS job gets a CU of its own
    // instead of sharing one with an MTTKRP workgroup (measured: the job at
    // ~22 us beside the MTTKRP against ~14 us alone).
    constexpr int WAVES = 4, WAVES_BF16 = 8;
    const int waves = bf16 ? WAVES_BF16 : WAVES;
    // FP32 at rank <= 32 (NT <= 2), unmasked windows and many rows: 64-row
    // workgroups (MT = 4).  Each B operand (window read and PRE scaling) then
    // feeds four row tiles instead of two, and the last mode loads 4 rows per
    // 16-byte vector instead of 2 per 8-byte one.  The inlined ALS job already
    // sets these instances' register budget, so the larger tile keeps two
    // workgroups per CU.
    const bool mt4 = !bf16 && NT <= 2 && XL == 0 && (I % 64 == 0 || I >= 2048) &&
                     (!caseB || (I % 4 == 0 && (uintptr_t)Xuse % 16 == 0));
    const int MT = mt4 ? 4 : (caseB && (I % 2 != 0 || (uintptr_t)Xuse % 8 != 0)) ? 1 : 2;
    const i64 rows_per_wg = 16 * MT;
    const i64 rtiles = (i64)cdiv(I, rows_per_wg);
    const i64 tiles = rtiles * rchunks;
    a.kw = kw;
    a.nwin = (int)((pl.fast_size + kw - 1) / kw);
    a.n_slow = pl.J / pl.fast_size;
    a.qs = caseB ? 1 : Qmem / fast_stride;
    a.rtiles = (int)rtiles;
    a.win_vec = (R % 4 == 0) && ((uintptr_t)fastf % 16 == 0);
    // the kernel instance for this shape: calls f(kernel) with its function pointer
    auto with_kernel = [&](auto&& f) {
      auto pick = [&](auto wv, auto xl) {
        constexpr bool WV = decltype(wv)::value;
        constexpr int L = decltype(xl)::value;  // XL; 2, 3 and 4 only for fp32 case A
        constexpr int LB = L >= 2 ? 1 : L;      // instances that have no element- or pair-wise form
        constexpr int LX = L == 4 ? 4 : LB;     // bf16 instances: also the unmasked padded form
        if (!bf16) {
          if constexpr (L == 0) {
            if (MT == 4) {
              if (!caseB)
                return NT == 1 ? f(mttkrp_f32_mfma<1, WAVES, 4, false, KWMAX, WV, 0>)
                               : f(mttkrp_f32_mfma<2, WAVES, 4, false, KWMAX, WV, 0>);
              return NT == 1 ? f(mttkrp_f32_mfma<1, WAVES, 4, true, KWMAX, WV, 0>)
                             : f(mttkrp_f32_mfma<2, WAVES, 4, true, KWMAX, WV, 0>);
            }
          }
          if (!caseB) {
            if (NT == 1) return f(mttkrp_f32_mfma<1, WAVES, 2, false, KWMAX, WV, L>);
            if (NT == 2) return f(mttkrp_f32_mfma<2, WAVES, 2, false, KWMAX, WV, L>);
            return f(mttkrp_f32_mfma<4, WAVES, 2, false, KWMAX, WV, L>);
          }
          if (MT == 2) {
            if (NT == 1) return f(mttkrp_f32_mfma<1, WAVES, 2, true, KWMAX, WV, LB>);
            if (NT == 2) return f(mttkrp_f32_mfma<2, WAVES, 2, true, KWMAX, WV, LB>);
            return f(mttkrp_f32_mfma<4, WAVES, 2, true, KWMAX, WV, LB>);
          }
          if (NT == 1) return f(mttkrp_f32_mfma<1, WAVES, 1, true, KWMAX, WV, LB>);
          if (NT == 2) return f(mttkrp_f32_mfma<2, WAVES, 1, true, KWMAX, WV, LB>);
          return f(mttkrp_f32_mfma<4, WAVES, 1, true, KWMAX, WV, LB>);
        }
        if (x3) {
          if (NT == 1) return f(mttkrp_bf16_mfma<float, 1, WAVES_BF16, 2, 1, KWMAX, WV, LX>);
          if (NT == 2) return f(mttkrp_bf16_mfma<float, 2, WAVES_BF16, 2, 1, KWMAX, WV, LX>);
          return f(mttkrp_bf16_mfma<float, 4, WAVES_BF16, 2, 1, KWMAX, WV, LX>);
        }
        if (KB == 64) {  // only a fastest mode divisible by 64 (never ragged)
          if (NT == 1) return f(mttkrp_bf16_mfma<uint16_t, 1, WAVES_BF16, 2, 2, KWMAX, WV, 0>);
          if (NT == 2) return f(mttkrp_bf16_mfma<uint16_t, 2, WAVES_BF16, 2, 2, KWMAX, WV, 0>);
          return f(mttkrp_bf16_mfma<uint16_t, 4, WAVES_BF16, 2, 2, KWMAX, WV, 0>);
        }
        if (NT == 1) return f(mttkrp_bf16_mfma<uint16_t, 1, WAVES_BF16, 2, 1, KWMAX, WV, LX>);
        if (NT == 2) return f(mttkrp_bf16_mfma<uint16_t, 2, WAVES_BF16, 2, 1, KWMAX, WV, LX>);
        return f(mttkrp_bf16_mfma<uint16_t, 4, WAVES_BF16, 2, 1, KWMAX, WV, LX>);
      };
      auto pick_xl = [&](auto wv) {
        if (XL == 0) return pick(wv, std::integral_constant<int, 0>{});
        if (XL == 1) return pick(wv, std::integral_constant<int, 1>{});
        if (XL == 4) return pick(wv, std::integral_constant<int, 4>{});
        if (XL == 3) return pick(wv, std::integral_constant<int, 3>{});
        return pick(wv, std::integral_constant<int, 2>{});
      };
      return a.win_vec ? pick_xl(std::true_type{}) : pick_xl(std::false_type{});
    };
    // workgroups per CU doing MTTKRP work: 2 four-wave FP32 workgroups (the sweet
    // spot swept at 256^3 and 512x512x128), one eight-wave BF16 workgroup, fewer if
    // the instance cannot keep that many resident, or the TENRSO_WG_PER_CU_*
    // override.  When the ALS job shares the launch it takes one of the resident
    // slots, so every workgroup is resident at once.
    int wpc = bf16 ? ctx->wg_per_cu_bf16 : ctx->wg_per_cu_f32;
    int occ = 0;
    {
      const void* kp = nullptr;
      with_kernel([&](auto kern) {
        kp = reinterpret_cast<const void*>(kern);
        return 0;
      });
      for (int i = 0; i < ctx->occ_n; ++i)
        if (ctx->occ_kern[i] == kp) occ = ctx->occ_val[i];
      if (occ <= 0) {
        int nb = 0;
        TENRSO_HIP_TRY(with_kernel([&](auto kern) {
          return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, waves * 64, 0);
        }));
        occ = std::max(1, nb);
        if (ctx->occ_n < 32) {
          ctx->occ_kern[ctx->occ_n] = kp;
          ctx->occ_val[ctx->occ_n++] = occ;
        }
      }
    }
    if (wpc <= 0) wpc = std::min(occ, 2);
    const bool fuse_job = job != nullptr && rchunks == 1;
    const i64 target = (i64)wpc * ctx->num_cus - ((fuse_job && occ <= wpc) ? 1 : 0);
    // each workgroup keeps >= WAVES slow indices
    i64 ssp = std::max<i64>(1, target / std::max<i64>(1, tiles) / a.nwin);
    // Many row tiles but fewer than 3/4 of the resident slots (9216 rows of 32-row
    // tiles: 288 workgroups on 512 slots): about 1.5x the slots in workgroups.
    // Swept on the 9216 x 96 x 96 view of 96^4 at rank 32: 1 split 288
    // workgroups, 3 splits 864 the fastest (more splits cost their reduction).
    if (!fuse_job && tiles * a.nwin * ssp * 4 < target * 3) ssp = cdiv(target * 3, 2 * tiles * a.nwin);
    ssp = std::min<i64>(ssp, a.n_slow >= waves ? a.n_slow / waves : 1);
    ssp = std::min<i64>(ssp, 65535 / a.nwin);
    // a caller's workspace bounds the partials: fewer splits, or refuse
    if (max_splits() < a.nwin) return too_small(a.nwin);
    ssp = std::min<i64>(ssp, max_splits() / a.nwin);
    a.slow_splits = (int)std::max<i64>(1, ssp);
    a.slow_q = a.n_slow / a.slow_splits;
    a.slow_rem = a.n_slow % a.slow_splits;
    splits = a.nwin * a.slow_splits;
    TENRSO_TRY(partial_buffer(splits, &partial));
    if (job && !fuse_job) TENRSO_TRY(als_job_launch(ctx, *job));
    const dim3 block(waves * 64);
    // BF16 with the job fused: the MTTKRP workgroups wait for the job's Gram
    // loads (the job, not the HBM-bound MTTKRP, sets the length of that launch)
    unsigned* gate = nullptr;
    if (bf16 && fuse_job && ctx->job_gate) {
      void* gp = nullptr;
      TENRSO_TRY(ctx_buffer(ctx, SLOT_GATE, sizeof(unsigned), &gp));
      gate = (unsigned*)gp;
    }
    auto launch_main = [&](bool with_job) -> int {
      for (int rc = 0; rc < rchunks; ++rc) {
        const int has_job = (with_job && fuse_job && rc == 0) ? 1 : 0;
        AlsJob jarg{};
        if (has_job) {
          jarg = *job;
          if (gate) {
            jarg.gate = gate;
            jarg.gate_stamp = ++ctx->gate_stamp;
          }
        }
        const dim3 grid((unsigned)(rtiles * splits + has_job), 1, 1);
        a.r0 = rc * RT;
        a.rcount = (int)std::min<i64>(RT, R - a.r0);
        with_kernel([&](auto kern) {
          using XT = std::remove_const_t<std::remove_pointer_t<
              std::tuple_element_t<0, typename KernelArgs<decltype(kern)>::type>>>;
          hipLaunchKernelGGL(kern, grid, block, 0, ctx->stream, (const XT*)Xuse, a, partial, jarg, has_job);
          return 0;
        });
        TENRSO_HIP_TRY(hipGetLastError());
      }
      return TENRSO_OK;
    };
    if (ctx->profile_reps > 0) {
      // profiling: the main kernel alone, `profile_reps` times between two events
      TENRSO_TRY(launch_main(true));
      TENRSO_HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
      for (int rep = 0; rep < ctx->profile_reps; ++rep) TENRSO_TRY(launch_main(false));
      TENRSO_HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
      TENRSO_HIP_TRY(hipEventSynchronize(ctx->ev1));
      float ms = 0.f;
      TENRSO_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
      ctx->last_kernel_ms = (double)ms / ctx->profile_reps;
    } else {
      TENRSO_TRY(launch_main(true));
    }
  } else {
    if (pad > 0 && fast_stride != pl.fast_size) {
      set_error("internal: a padded tensor copy reached the generic MTTKRP");
      return TENRSO_ERR_UNSUPPORTED;
    }
    if (job) TENRSO_TRY(als_job_launch(ctx, *job));
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
    ga.bf16 = (bf16 && !x3) ? 1 : 0;
    const i64 tiles = cdiv(nout, 256);
    splits = (int)std::max<i64>(1, std::min<i64>(pl.J, (2LL * ctx->num_cus + tiles - 1) / std::max<i64>(1, tiles)));
    splits = std::min(splits, 1024);
    if (max_splits() < 1) return too_small(1);
    splits = (int)std::min<i64>(splits, max_splits());
    ga.splits = splits;
    TENRSO_TRY(partial_buffer(splits, &partial));
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
