// mttkrp.hpp -- host-side interface of the dense MTTKRP kernels.
#pragma once
#include "common.hpp"

namespace tenrso {

struct AlsJob;

constexpr int MAXD = 8;

// Output form of a dense MTTKRP launch.
enum class MttkrpOut {
  F32,      // reduce split partials into an fp32 I x R result
  F64,      // reduce split partials into an fp64 I x R result
  PARTIAL   // leave [splits, I, R] fp32 partials (caller reduces, e.g. the ALS solve)
};

struct MttkrpResult {
  const float* partial = nullptr;  // [splits, I, R]
  int splits = 0;
};

// Caller-provided device workspace for the PRE Khatri-Rao rows and the split
// partials (instead of the context's grow-only buffers).  The split count is
// reduced until the partials fit; a shape whose partials cannot fit even
// unsplit is TENRSO_ERR_UNSUPPORTED.  Nothing is allocated.
struct MttkrpScratch {
  void* base = nullptr;  // 256-byte aligned device memory
  size_t bytes = 0;
};

// X: fp32 tensor (prec FP32) or bf16 tensor (prec BF16, uint16 storage).  For BF16
// and mode == ndim-1 the caller passes X_t, the tensor with the last mode moved
// to the front: bf16 (xt_f32 false), or the exact fp32 tensor (xt_f32 true), which
// the MFMA kernel contracts with a three-way bf16 split at FP32 accuracy.
// job (optional): the ALS linear algebra of this mode; it runs as workgroup 0 of the
// MTTKRP launch (or, where it cannot, as its own kernel just before it).
// x_pad / xt_pad (BF16 only, 0 = none): the memory extent of the last mode of X
// and of X_t when those copies have it padded with zeros.
// pre_rows (optional, tensors of order >= 4): the PRE Khatri-Rao rows of this
// mode already formed (mttkrp_pre_batch), instead of a launch that forms them.
int mttkrp_dense(tenrso_ctx* ctx, int prec, const void* X, const void* X_t, int nd, const i64* shape,
                 const float* const* F, i64 R, int mode, MttkrpOut outk, void* out, MttkrpResult* res,
                 const AlsJob* job = nullptr, bool xt_f32 = false, const MttkrpScratch* scratch = nullptr,
                 i64 x_pad = 0, i64 xt_pad = 0, const float* pre_rows = nullptr);

// The PRE Khatri-Rao rows of the MTTKRPs of several modes of one tensor (order
// >= 4) in one launch: pre[i] receives the rows the MFMA MTTKRP of modes[i]
// contracts (via_copy_last: the last mode runs as mode 0 of the last-to-front
// copy), null where the mode needs none; modes with the same rows share them.
// The rows live in the context's SLOT_PRE buffer until the next call.
int mttkrp_pre_batch(tenrso_ctx* ctx, int nd, const i64* shape, const float* const* F, i64 R, int nmodes,
                     const int* modes, bool via_copy_last, const float** pre);

// Size of the file-scope workspace of the no-allocation MTTKRP entry point
// (noalloc.hip): 16 MiB per GPU.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)16 << 20;
// Resolves that workspace's address on `device` (which must be current); called
// once per device by tenrso_ctx_create.
int noalloc_scratch_init(int device);

// Validation shared by the MTTKRP entry points (mode range, sizes, at least 2
// modes), with the reference's messages.
int mttkrp_check_args(int ndim, const int64_t* shape, int64_t rank, int mode);

// The ALS job as its own one-workgroup kernel on the context stream.
int als_job_launch(tenrso_ctx* ctx, const AlsJob& job);

// Sum split partials [splits, n] in FP64 into out (fp32 or fp64).
int reduce_partials(tenrso_ctx* ctx, const float* partial, int splits, i64 n, float* out32, double* out64);

// fp32 -> bf16 (RNE) copy; when `last_to_front` the output is the tensor with its
// last mode moved to the front ([I_{N-1}, I_0, ..., I_{N-2}]).
int convert_bf16(tenrso_ctx* ctx, const float* src, int nd, const i64* shape, bool last_to_front, uint16_t* dst);
// BF16 ALS input in one pass: dst = bf16(X), dstT = the fp32 tensor with its last
// mode moved to the front, and (sum of squares, min) partials in tmp[0, 2 * nblk).
// TENRSO_ERR_UNSUPPORTED (nothing launched) unless the last mode is a multiple of
// 4 and X is 16-byte aligned.  dst may be null: then only dstT and the partials.
int prepare_bf16_als(tenrso_ctx* ctx, const float* X, int nd, const i64* shape, uint16_t* dst, float* dstT,
                     double* tmp, int* nblk_out);

// [.., I_{N-1}] -> [I_{N-1}, ..] (last mode moved to the front), plain copies.
int transpose_last_to_front_u16(tenrso_ctx* ctx, const uint16_t* src, int nd, const i64* shape, uint16_t* dst);
int transpose_last_to_front_f32(tenrso_ctx* ctx, const float* src, int nd, const i64* shape, float* dst);
// Padded layouts for the BF16 ALS of tensors with ragged last modes: the bf16
// copy with its last mode padded to Kp, and the fp32 last-to-front copy with
// its new last mode (shape[nd-2]) padded to Bp; pad entries are zero.
int convert_bf16_padded(tenrso_ctx* ctx, const float* src, int nd, const i64* shape, i64 Kp, uint16_t* dst);
// The fp32 tensor with its last mode padded to Kp (the FP32 ALS of ragged shapes).
int copy_f32_padded(tenrso_ctx* ctx, const float* src, int nd, const i64* shape, i64 Kp, float* dst);
int transpose_last_to_front_f32_padded(tenrso_ctx* ctx, const float* src, int nd, const i64* shape, i64 Bp,
                                       float* dst);

// Forward-order Khatri-Rao of `n` factors (rows = prod sizes) into dst (fp32).
int khatri_rao_many(tenrso_ctx* ctx, int n, const float* const* F, const i64* sizes, i64 R, float* dst);

}  // namespace tenrso
