// sparse.hpp -- device sparse tensor (per-mode sorted COO + CSF export) and sparse MTTKRP.
#pragma once
#include <vector>

#include "common.hpp"

namespace tenrso {

constexpr int MAXD_SP = 8;

struct SpMode {
  i64 rows = 0;                // I_n
  int64_t* rowptr = nullptr;   // I_n + 1 nonzero offsets per output row
  int32_t* row = nullptr;      // nnz: output-mode index of each sorted nonzero
  int32_t* oth = nullptr;      // (nd-1) x nnz: other modes' indices, ascending mode order
  float* vals = nullptr;       // nnz
  // Compressed fibres (the CSF of mode order [n, others ascending] cut above its
  // leaf level): fibre f is the f-th run of sorted nonzeros sharing the prefix
  // (i_n, i_o1, .., i_o(D-1)), D = nd - 1; fpre is D x nfib (SoA: row, then the
  // prefix's other indices).  The leaves are leafe (nnz words: the last other
  // mode's index, bit 31 set on the last nonzero of each fibre) and vals.
  // chunk_fib[c] is the fibre holding nonzero c * SP_CSF_W.
  i64 nfib = 0;
  int32_t* leafe = nullptr;    // nnz
  int32_t* fpre = nullptr;     // D x nfib
  int32_t* chunk_fib = nullptr;
  i64 nchunks = 0;
};

constexpr int SP_CSF_W = 128;  // nonzeros per worker of the fibre kernel

}  // namespace tenrso

struct tenrso_sptensor {
  int nd = 0;
  int64_t nnz = 0;
  int64_t shape[tenrso::MAXD_SP] = {};
  int32_t* idx = nullptr;  // original COO, SoA nd x nnz
  float* vals = nullptr;
  tenrso::SpMode modes[tenrso::MAXD_SP];
};

namespace tenrso {

int sp_create(tenrso_ctx* ctx, int nd, const i64* shape, i64 nnz, const int32_t* idx, const float* vals, bool host,
              tenrso_sptensor** out);
void sp_destroy(tenrso_sptensor* t);
// M = sparse MTTKRP (fp32, shape[mode] x R), deterministic
int sp_mttkrp(tenrso_ctx* ctx, const tenrso_sptensor* t, const float* const* F, i64 R, int mode, float* out);
int sp_export_csf(tenrso_ctx* ctx, const tenrso_sptensor* t, int mode, std::vector<std::vector<int64_t>>& fptr,
                  std::vector<std::vector<int32_t>>& fids, std::vector<float>& vals);
// HiCOO of the COO as uploaded (HiCooTensor::from_coo, hicoo.rs:121-205): block
// coordinate = index / bs, local = index % bs, nonzeros stably sorted by block
// coordinate (lexicographic), duplicates kept.  Outputs are host vectors:
// block_coords (nd x blocks, SoA), block_ptrs (blocks + 1), local (nd x nnz, SoA).
int sp_export_hicoo(tenrso_ctx* ctx, const tenrso_sptensor* t, const i64* bs, std::vector<int64_t>& block_coords,
                    std::vector<int64_t>& block_ptrs, std::vector<int32_t>& local, std::vector<float>& vals);

}  // namespace tenrso
