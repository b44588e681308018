// cp_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference CP decomposition hot path (tenrso, Rust),
// used by tests/, the smoke check and bench.py's cpu_baseline leg as the
// checker.  Nothing in the product (tenrso_amd/) links or calls this file.
//
// Every function cites the reference file:line it restates.  Two instantiations:
//   * T = double : the parity oracle (every correctness bar is checked against it)
//   * T = float  : "reference as written" (sequential f32 sums, no FMA), reported
//                  beside the oracle for information.
// Build: g++ -O2 -ffp-contract=off -fopenmp -shared -fPIC (oracle/build_oracle.py).
// -ffp-contract=off keeps `sum = sum + a*b` as two roundings, like Rust's `*`/`+`.
//
// Parity-unpinned pieces (no reference test fixes them; see the module docs in
// oracle/oracle.py): the per-row `scirs2_linalg::lstsq` solve is restated as a
// Cholesky solve of the SPD normal matrix with the reference's ridge retry, and a
// system that fails both factorisations is a LinalgError, as in the reference;
// `orthonormalize_factor` uses Householder QR (scirs2_linalg::qr's algorithm is
// not in the reference).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using i64 = int64_t;

i64 prod(const i64* s, int n) {
  i64 p = 1;
  for (int k = 0; k < n; ++k) p *= s[k];
  return p;
}

// ---------------------------------------------------------------------------
// unfold_tensor  (tenrso-kernels/src/mttkrp.rs:149-177,
//                 tenrso-core/src/dense/shape_ops.rs:131-157)
// permutation [mode, 0..mode, mode+1..N], C-order copy, reshape (I_mode, prod others):
// column j = sum_m i_{d_m} * prod_{m'>m} I_{d_m'}  (last "other" dim fastest).
// ---------------------------------------------------------------------------
template <class T>
std::vector<T> unfold(const T* X, int nd, const i64* shape, int mode) {
  const i64 total = prod(shape, nd);
  const i64 rows = shape[mode];
  const i64 cols = total / rows;
  std::vector<T> U((size_t)total);
  std::vector<i64> strides(nd);
  strides[nd - 1] = 1;
  for (int k = nd - 2; k >= 0; --k) strides[k] = strides[k + 1] * shape[k + 1];
  std::vector<int> others;
  for (int k = 0; k < nd; ++k)
    if (k != mode) others.push_back(k);
  std::vector<i64> idx(nd, 0);
  // iterate the permuted array in C order: (i_mode, others...)
  for (i64 i = 0; i < rows; ++i) {
    std::fill(idx.begin(), idx.end(), 0);
    idx[mode] = i;
    for (i64 j = 0; j < cols; ++j) {
      i64 off = 0;
      for (int k = 0; k < nd; ++k) off += idx[k] * strides[k];
      U[(size_t)(i * cols + j)] = X[off];
      // increment "others" multi-index, last fastest
      for (int m = (int)others.size() - 1; m >= 0; --m) {
        int d = others[m];
        if (++idx[d] < shape[d]) break;
        idx[d] = 0;
      }
    }
  }
  return U;
}

// khatri_rao_two (mttkrp.rs:209-240) == khatri_rao (khatri_rao.rs:56-89):
// C[ia*J + ib, k] = A[ia,k] * B[ib,k]
template <class T>
std::vector<T> kr_two(const T* A, i64 I, const T* B, i64 J, i64 K) {
  std::vector<T> C((size_t)(I * J * K), T(0));
  for (i64 k = 0; k < K; ++k)
    for (i64 ia = 0; ia < I; ++ia)
      for (i64 ib = 0; ib < J; ++ib) C[(size_t)((ia * J + ib) * K + k)] = A[ia * K + k] * B[ib * K + k];
  return C;
}

// khatri_rao_except_mode (mttkrp.rs:182-206): forward order, left to right.
template <class T>
std::vector<T> kr_except(const T* const* F, int nd, const i64* shape, i64 R, int skip) {
  std::vector<T> res;
  i64 rows = 0;
  bool first = true;
  for (int k = 0; k < nd; ++k) {
    if (k == skip) continue;
    if (first) {
      res.assign(F[k], F[k] + shape[k] * R);
      rows = shape[k];
      first = false;
    } else {
      res = kr_two(res.data(), rows, F[k], shape[k], R);
      rows *= shape[k];
    }
  }
  return res;
}

// mttkrp (mttkrp.rs:74-146): i, r, j triple loop, sequential sum in T.
template <class T>
int mttkrp_faithful(const T* X, int nd, const i64* shape, const T* const* F, i64 R, int mode, T* out) {
  if (mode < 0 || mode >= nd) return 1;
  if (nd < 2) return 2;  // "Need at least 2 factor matrices for MTTKRP" (mttkrp.rs:194-196)
  std::vector<T> U = unfold(X, nd, shape, mode);
  std::vector<T> KR = kr_except(F, nd, shape, R, mode);
  const i64 rows = shape[mode];
  const i64 cols = prod(shape, nd) / rows;
  for (i64 i = 0; i < rows; ++i)
    for (i64 r = 0; r < R; ++r) {
      T sum = T(0);
      for (i64 j = 0; j < cols; ++j) sum = sum + U[(size_t)(i * cols + j)] * KR[(size_t)(j * R + r)];
      out[i * R + r] = sum;
    }
  return 0;
}

// Fast FP64 MTTKRP: same sum as mttkrp.rs:74-146 with a cache-friendly order.
// X viewed as [P, I, Q] (P = prod dims before mode, Q = prod dims after);
// M[i,r] = sum_p L[p,r] * sum_q X[p,i,q] * Rk[q,r], with L / Rk the forward-order
// KR products of the factors before / after the mode.  Row blocks of i run in
// parallel (OpenMP); within a row the p, q order is sequential.
void mttkrp_fast_f64(const double* X, int nd, const i64* shape, const double* const* F, i64 R, int mode,
                     double* out) {
  i64 P = 1, Q = 1;
  for (int k = 0; k < mode; ++k) P *= shape[k];
  for (int k = mode + 1; k < nd; ++k) Q *= shape[k];
  const i64 I = shape[mode];
  // L (P x R) and Rk (Q x R), forward order
  std::vector<double> L((size_t)(P * R), 1.0), Rk((size_t)(Q * R), 1.0);
  {
    std::vector<i64> idx(nd, 0);
    for (i64 p = 0; p < P; ++p) {
      i64 rem = p;
      for (int k = mode - 1; k >= 0; --k) { idx[k] = rem % shape[k]; rem /= shape[k]; }
      for (i64 r = 0; r < R; ++r) {
        double v = 1.0;
        for (int k = 0; k < mode; ++k) v *= F[k][idx[k] * R + r];
        L[(size_t)(p * R + r)] = v;
      }
    }
    for (i64 q = 0; q < Q; ++q) {
      i64 rem = q;
      for (int k = nd - 1; k > mode; --k) { idx[k] = rem % shape[k]; rem /= shape[k]; }
      for (i64 r = 0; r < R; ++r) {
        double v = 1.0;
        for (int k = mode + 1; k < nd; ++k) v *= F[k][idx[k] * R + r];
        Rk[(size_t)(q * R + r)] = v;
      }
    }
  }
#pragma omp parallel
  {
    std::vector<double> t((size_t)R), acc((size_t)R);
#pragma omp for schedule(static)
    for (i64 i = 0; i < I; ++i) {
      std::fill(acc.begin(), acc.end(), 0.0);
      for (i64 p = 0; p < P; ++p) {
        std::fill(t.begin(), t.end(), 0.0);
        const double* xrow = X + (p * I + i) * Q;
        for (i64 q = 0; q < Q; ++q) {
          const double x = xrow[q];
          const double* rk = &Rk[(size_t)(q * R)];
          for (i64 r = 0; r < R; ++r) t[r] += x * rk[r];
        }
        const double* l = &L[(size_t)(p * R)];
        for (i64 r = 0; r < R; ++r) acc[r] += l[r] * t[r];
      }
      for (i64 r = 0; r < R; ++r) out[i * R + r] = acc[r];
    }
  }
}

// mttkrp_blocked_parallel (mttkrp.rs:528-609) restated with OpenMP: the unfolded
// column range is cut into tiles, each tile's partial (I_mode x R) is computed
// independently and the partials are summed in tile order.  Used only as the
// strongest "reference-style" CPU timing; the partial buffers are per tile.
template <class T>
int mttkrp_blocked_parallel(const T* X, int nd, const i64* shape, const T* const* F, i64 R, int mode, i64 tile,
                            T* out) {
  if (tile <= 0) return 3;
  if (mode < 0 || mode >= nd) return 1;
  if (nd < 2) return 2;
  std::vector<T> U = unfold(X, nd, shape, mode);
  std::vector<T> KR = kr_except(F, nd, shape, R, mode);
  const i64 rows = shape[mode];
  const i64 cols = prod(shape, nd) / rows;
  const i64 ntiles = (cols + tile - 1) / tile;
  std::vector<T> parts((size_t)(ntiles * rows * R), T(0));
#pragma omp parallel for schedule(dynamic)
  for (i64 t = 0; t < ntiles; ++t) {
    const i64 c0 = t * tile, c1 = std::min(cols, c0 + tile);
    T* part = &parts[(size_t)(t * rows * R)];
    for (i64 i = 0; i < rows; ++i)
      for (i64 r = 0; r < R; ++r) {
        T sum = T(0);
        for (i64 j = c0; j < c1; ++j) sum = sum + U[(size_t)(i * cols + j)] * KR[(size_t)(j * R + r)];
        part[i * R + r] = sum;
      }
  }
  for (i64 x = 0; x < rows * R; ++x) out[x] = T(0);
  for (i64 t = 0; t < ntiles; ++t)
    for (i64 x = 0; x < rows * R; ++x) out[x] = out[x] + parts[(size_t)(t * rows * R + x)];
  return 0;
}

// compute_gram_matrix (cp/helpers.rs:427-452): G[i,j] = sum_k F[k,i]*F[k,j], sequential.
template <class T>
void gram(const T* Fm, i64 rows, i64 R, T* G) {
  for (i64 i = 0; i < R; ++i)
    for (i64 j = 0; j < R; ++j) {
      T s = T(0);
      for (i64 k = 0; k < rows; ++k) s = s + Fm[k * R + i] * Fm[k * R + j];
      G[i * R + j] = s;
    }
}

// compute_gram_hadamard (cp/helpers.rs:409-424): ones, then *= gram(F_k) for k != skip, ascending.
template <class T>
void gram_hadamard(const T* const* F, int nd, const i64* shape, i64 R, int skip, T* out) {
  for (i64 x = 0; x < R * R; ++x) out[x] = T(1);
  std::vector<T> g((size_t)(R * R));
  for (int k = 0; k < nd; ++k) {
    if (k == skip) continue;
    gram(F[k], shape[k], R, g.data());
    for (i64 x = 0; x < R * R; ++x) out[x] = out[x] * g[(size_t)x];
  }
}

// In-place Cholesky A = L L^T of an SPD n x n matrix (lower triangle used).
template <class T>
bool cholesky(std::vector<T>& A, i64 n) {
  for (i64 j = 0; j < n; ++j) {
    T d = A[(size_t)(j * n + j)];
    for (i64 k = 0; k < j; ++k) d -= A[(size_t)(j * n + k)] * A[(size_t)(j * n + k)];
    if (!(d > T(0)) || !std::isfinite((double)d)) return false;
    d = std::sqrt(d);
    A[(size_t)(j * n + j)] = d;
    for (i64 i = j + 1; i < n; ++i) {
      T s = A[(size_t)(i * n + j)];
      for (i64 k = 0; k < j; ++k) s -= A[(size_t)(i * n + k)] * A[(size_t)(j * n + k)];
      A[(size_t)(i * n + j)] = s / d;
    }
  }
  return true;
}

template <class T>
void chol_solve(const std::vector<T>& Lc, i64 n, T* b) {
  for (i64 i = 0; i < n; ++i) {
    T s = b[i];
    for (i64 k = 0; k < i; ++k) s -= Lc[(size_t)(i * n + k)] * b[k];
    b[i] = s / Lc[(size_t)(i * n + i)];
  }
  for (i64 i = n - 1; i >= 0; --i) {
    T s = b[i];
    for (i64 k = i + 1; k < n; ++k) s -= Lc[(size_t)(k * n + i)] * b[k];
    b[i] = s / Lc[(size_t)(i * n + i)];
  }
}

// solve_least_squares (cp/helpers.rs:455-495): for every row i, x = lstsq(G^T, M[i,:]).
// G^T is the symmetric normal matrix; lstsq's exact solution for SPD G^T is the
// Cholesky solution (restated here).  On failure the reference adds
// T::epsilon()*10R to the diagonal and retries (helpers.rs:477-490) and returns
// CpError::LinalgError if that fails too (helpers.rs:485).  ridge_eps is the
// epsilon of the instantiation being restated (<= 0: the epsilon of T).
// Returns the path taken: 0 = Cholesky, 1 = ridge retry, -1 = LinalgError (out
// untouched).
template <class T>
int solve_ls(const T* M, i64 rows, i64 R, const T* G, T* out, double ridge_eps = 0.0) {
  std::vector<T> Gt((size_t)(R * R));
  for (i64 i = 0; i < R; ++i)
    for (i64 j = 0; j < R; ++j) Gt[(size_t)(i * R + j)] = G[j * R + i];
  std::vector<T> Lc = Gt;
  int path = 0;
  if (!cholesky(Lc, R)) {
    path = 1;
    Lc = Gt;
    const T e = ridge_eps > 0.0 ? (T)ridge_eps : std::numeric_limits<T>::epsilon();
    const T eps = e * (T)(R * 10);
    for (i64 k = 0; k < R; ++k) Lc[(size_t)(k * R + k)] += eps;
    if (!cholesky(Lc, R)) return -1;
  }
  for (i64 i = 0; i < rows; ++i) {
    for (i64 j = 0; j < R; ++j) out[i * R + j] = M[i * R + j];
    chol_solve(Lc, R, out + i * R);
  }
  return path;
}

// compute_norm_squared (cp/helpers.rs:498-510): sequential sum of squares in T.
template <class T>
T norm_sq(const T* X, i64 n) {
  T s = T(0);
  for (i64 i = 0; i < n; ++i) s = s + X[i] * X[i];
  return s;
}

// compute_reconstruction_norm_squared (cp/helpers.rs:533-555)
template <class T>
T recon_norm_sq(const T* const* F, int nd, const i64* shape, i64 R) {
  T norm = T(0);
  for (i64 r = 0; r < R; ++r)
    for (i64 s = 0; s < R; ++s) {
      T cross = T(1);
      for (int k = 0; k < nd; ++k) {
        T ip = T(0);
        for (i64 i = 0; i < shape[k]; ++i) ip = ip + F[k][i * R + r] * F[k][i * R + s];
        cross = cross * ip;
      }
      norm = norm + cross;
    }
  return norm;
}

// compute_inner_product (cp/helpers.rs:558-579): mode-0 MTTKRP, then sum over r, i.
// `fast` selects the cache-friendly FP64 MTTKRP (same value up to rounding).
template <class T>
T inner_product(const T* X, int nd, const i64* shape, const T* const* F, i64 R, bool fast) {
  std::vector<T> M((size_t)(shape[0] * R));
  if constexpr (std::is_same<T, double>::value) {
    if (fast) mttkrp_fast_f64(X, nd, shape, F, R, 0, M.data());
    else mttkrp_faithful(X, nd, shape, F, R, 0, M.data());
  } else {
    (void)fast;
    mttkrp_faithful(X, nd, shape, F, R, 0, M.data());
  }
  T ip = T(0);
  for (i64 r = 0; r < R; ++r)
    for (i64 i = 0; i < shape[0]; ++i) ip = ip + M[(size_t)(i * R + r)] * F[0][i * R + r];
  return ip;
}

// compute_fit (cp/helpers.rs:513-530)
template <class T>
T fit_value(T tensor_norm_sq, T recon, T ip) {
  T err_sq = tensor_norm_sq + recon - T(2) * ip;
  T err = std::sqrt(std::max(err_sq, T(0)));
  T fit = T(1) - err / std::sqrt(tensor_norm_sq);
  return std::min(std::max(fit, T(0)), T(1));
}

// soft_threshold (cp/helpers.rs:610-623)
template <class T>
void soft_threshold(T* Fm, i64 n, T th) {
  for (i64 x = 0; x < n; ++x) {
    T v = Fm[x];
    Fm[x] = v > th ? v - th : (v < -th ? v + th : T(0));
  }
}

// apply_tikhonov_to_gram (cp/helpers.rs:630-702)
template <class T>
void tikhonov(T* G, i64 n, T lam, i64 order) {
  auto at = [&](i64 i, i64 j) -> T& { return G[i * n + j]; };
  if (order == 0) {
    for (i64 i = 0; i < n; ++i) at(i, i) = at(i, i) + lam;
  } else if (order == 1) {
    if (n >= 2)
      for (i64 i = 0; i < n; ++i) {
        T d = (i == 0 || i == n - 1) ? T(1) : T(2);
        at(i, i) = at(i, i) + lam * d;
        if (i + 1 < n) {
          at(i, i + 1) = at(i, i + 1) - lam;
          at(i + 1, i) = at(i + 1, i) - lam;
        }
      }
  } else {
    if (n >= 3) {
      for (i64 i = 0; i < n; ++i) {
        T d = (i == 0 || i == n - 1) ? T(1) : ((i == 1 || i == n - 2) ? T(5) : T(6));
        at(i, i) = at(i, i) + lam * d;
        if (i + 1 < n) {
          T off1 = (i == 0 || i == n - 2) ? T(-2) : T(-4);
          at(i, i + 1) = at(i, i + 1) + lam * off1;
          at(i + 1, i) = at(i + 1, i) + lam * off1;
        }
        if (i + 2 < n) {
          at(i, i + 2) = at(i, i + 2) + lam;
          at(i + 2, i) = at(i + 2, i) + lam;
        }
      }
    } else {
      for (i64 i = 0; i < n; ++i) at(i, i) = at(i, i) + lam;
    }
  }
}

// Householder thin QR, returns Q (rows x R) in place of Fm.  Stands in for
// scirs2_linalg::qr in orthonormalize_factor (cp/helpers.rs:582-607); the sign
// convention of the third-party routine is not visible (parity unpinned).  Q is
// normalised to the QR factorisation with a positive diagonal of R (column k of
// Q negated where the reflection made R[k][k] negative), the unique one for a
// full-rank factor: Q = A chol(A^T A)^-T.
template <class T>
void orthonormalize(T* Fm, i64 m, i64 n) {
  std::vector<double> A((size_t)(m * n));
  for (i64 x = 0; x < m * n; ++x) A[(size_t)x] = Fm[x];
  std::vector<std::vector<double>> vs;
  std::vector<double> rsign((size_t)n, 1.0);
  for (i64 k = 0; k < n && k < m; ++k) {
    std::vector<double> v((size_t)m, 0.0);
    double nrm = 0.0;
    for (i64 i = k; i < m; ++i) nrm += A[(size_t)(i * n + k)] * A[(size_t)(i * n + k)];
    nrm = std::sqrt(nrm);
    double a0 = A[(size_t)(k * n + k)];
    double alpha = a0 > 0 ? -nrm : nrm;
    rsign[(size_t)k] = alpha < 0 ? -1.0 : 1.0;  // R[k][k] = alpha
    for (i64 i = k; i < m; ++i) v[(size_t)i] = A[(size_t)(i * n + k)];
    v[(size_t)k] -= alpha;
    double vn = 0.0;
    for (i64 i = k; i < m; ++i) vn += v[(size_t)i] * v[(size_t)i];
    if (vn > 0) {
      for (i64 j = k; j < n; ++j) {
        double d = 0.0;
        for (i64 i = k; i < m; ++i) d += v[(size_t)i] * A[(size_t)(i * n + j)];
        d = 2.0 * d / vn;
        for (i64 i = k; i < m; ++i) A[(size_t)(i * n + j)] -= d * v[(size_t)i];
      }
    }
    vs.push_back(v);
  }
  // Q = H_0 H_1 ... H_{n-1} applied to the first n columns of I
  std::vector<double> Q((size_t)(m * n), 0.0);
  for (i64 i = 0; i < std::min(m, n); ++i) Q[(size_t)(i * n + i)] = 1.0;
  for (i64 k = (i64)vs.size() - 1; k >= 0; --k) {
    const auto& v = vs[(size_t)k];
    double vn = 0.0;
    for (i64 i = k; i < m; ++i) vn += v[(size_t)i] * v[(size_t)i];
    if (vn <= 0) continue;
    for (i64 j = 0; j < n; ++j) {
      double d = 0.0;
      for (i64 i = k; i < m; ++i) d += v[(size_t)i] * Q[(size_t)(i * n + j)];
      d = 2.0 * d / vn;
      for (i64 i = k; i < m; ++i) Q[(size_t)(i * n + j)] -= d * v[(size_t)i];
    }
  }
  for (i64 i = 0; i < m; ++i)
    for (i64 j = 0; j < n; ++j) Fm[i * n + j] = (T)(Q[(size_t)(i * n + j)] * rsign[(size_t)j]);
}

}  // namespace

// ===========================================================================
// C ABI for the Python loader (oracle/oracle.py).  All arrays are host,
// C-contiguous, row-major.
// ===========================================================================
extern "C" {

struct orc_constraints {
  int nonnegative;      // CpConstraints::nonnegative (cp/types.rs:198-225)
  int orthogonal;       // CpConstraints::orthogonal
  double l2;            // effective_l2()  (cp/types.rs:288-293)
  double l1;            // effective_l1()  (cp/types.rs:296-298)
  int tikhonov_order;   // >0 selects apply_tikhonov_to_gram (cp/core.rs:330-335)
  double tikhonov_lambda;
};

struct orc_result {
  int iters;
  int reason;           // 0 FitTolerance, 1 MaxIterations, 2 Oscillation, 3 TimeLimit (cp/types.rs:312-322)
  int oscillation_count;
  double fit;
  double final_fit_change;
  int solve_path_max;   // worst solve_ls path taken (0 Cholesky, 1 ridge retry)
  double seconds;       // wall time of the ALS loop
};

int orc_mttkrp_f64(const double* X, int nd, const int64_t* shape, const double* const* F, int64_t R, int mode,
                   double* out) {
  return mttkrp_faithful(X, nd, shape, F, R, mode, out);
}
int orc_mttkrp_f32(const float* X, int nd, const int64_t* shape, const float* const* F, int64_t R, int mode,
                   float* out) {
  return mttkrp_faithful(X, nd, shape, F, R, mode, out);
}
int orc_mttkrp_fast_f64(const double* X, int nd, const int64_t* shape, const double* const* F, int64_t R, int mode,
                        double* out) {
  if (mode < 0 || mode >= nd || nd < 2) return 1;
  mttkrp_fast_f64(X, nd, shape, F, R, mode, out);
  return 0;
}
int orc_mttkrp_blocked_parallel_f32(const float* X, int nd, const int64_t* shape, const float* const* F, int64_t R,
                                    int mode, int64_t tile, float* out) {
  return mttkrp_blocked_parallel(X, nd, shape, F, R, mode, tile, out);
}
int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_khatri_rao_f64(const double* A, int64_t I, const double* B, int64_t J, int64_t K, double* C) {
  auto v = kr_two(A, I, B, J, K);
  std::memcpy(C, v.data(), v.size() * sizeof(double));
}
void orc_khatri_rao_f32(const float* A, int64_t I, const float* B, int64_t J, int64_t K, float* C) {
  auto v = kr_two(A, I, B, J, K);
  std::memcpy(C, v.data(), v.size() * sizeof(float));
}

void orc_gram_f64(const double* Fm, int64_t rows, int64_t R, double* G) { gram(Fm, rows, R, G); }
void orc_gram_hadamard_f64(const double* const* F, int nd, const int64_t* shape, int64_t R, int skip, double* out) {
  gram_hadamard(F, nd, shape, R, skip, out);
}
int orc_solve_ls_f64(const double* M, int64_t rows, int64_t R, const double* G, double* out, double ridge_eps) {
  return solve_ls(M, rows, R, G, out, ridge_eps);
}
void orc_tikhonov_f64(double* G, int64_t n, double lam, int64_t order) { tikhonov(G, n, lam, order); }
void orc_soft_threshold_f64(double* Fm, int64_t n, double th) { soft_threshold(Fm, n, th); }
void orc_orthonormalize_f64(double* Fm, int64_t m, int64_t n) { orthonormalize(Fm, m, n); }
double orc_norm_sq_f64(const double* X, int64_t n) { return norm_sq(X, n); }
float orc_norm_sq_f32(const float* X, int64_t n) { return norm_sq(X, n); }
double orc_recon_norm_sq_f64(const double* const* F, int nd, const int64_t* shape, int64_t R) {
  return recon_norm_sq(F, nd, shape, R);
}
double orc_fit_f64(const double* X, int nd, const int64_t* shape, const double* const* F, int64_t R,
                   double tensor_norm_sq, int fast) {
  return fit_value(tensor_norm_sq, recon_norm_sq(F, nd, shape, R), inner_product(X, nd, shape, F, R, fast != 0));
}

// cp_als (cp/core.rs:59-182) and cp_als_constrained (cp/core.rs:225-408) with the
// initial factors supplied by the caller (the reference draws them from an
// unseeded RNG, cp/helpers.rs:212-231).  `fast` selects the cache-friendly FP64
// MTTKRP (only for T = double).  time_limit < 0 means None.  ridge_eps: see
// solve_ls.  Xlead (may be null): the tensor the MTTKRPs of modes 0..N-2
// contract instead of X (the model of an ALS whose leading modes read a rounded
// copy of X while the last mode, the norm and the fit use X itself).
// Returns 13 (LinalgError) when a solve fails after its ridge retry.
}  // extern "C"
template <class T>
int cp_als_impl(const T* X, int nd, const int64_t* shape, int64_t R, T** F, int max_iters, double tol,
                double time_limit, const orc_constraints* c, int fast, double* fit_history, orc_result* res,
                double ridge_eps = 0.0, const T* Xlead = nullptr) {
  if (R == 0) return 10;
  for (int k = 0; k < nd; ++k)
    if (R > shape[k]) return 10;                       // InvalidRank (core.rs:82-91)
  if (!(tol >= 0.0 && tol < 1.0)) return 11;           // InvalidTolerance (core.rs:93-95)
  const i64 total = prod(shape, nd);
  orc_constraints cc{};
  if (c) cc = *c;
  if (cc.nonnegative)
    for (i64 x = 0; x < total; ++x)
      if (X[x] < T(0)) return 12;                      // NonnegativeViolation (core.rs:273-275)
  if (cc.nonnegative)
    for (int k = 0; k < nd; ++k)
      for (i64 x = 0; x < shape[k] * R; ++x) F[k][x] = std::max(F[k][x], T(0));   // core.rs:281-285
  const T tnorm = norm_sq(X, total);
  T prev_fit = T(0), fit = T(0);
  int iters = 0, osc = 0, reason = 1, worst = 0;
  T final_change = T(0);
  std::vector<T> M, G((size_t)(R * R));
  auto t0 = std::chrono::steady_clock::now();
  for (int iter = 0; iter < max_iters; ++iter) {
    if (time_limit >= 0.0) {
      double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (el > time_limit) { reason = 3; break; }
    }
    iters = iter + 1;
    for (int mode = 0; mode < nd; ++mode) {
      M.assign((size_t)(shape[mode] * R), T(0));
      const T* Xm = (Xlead && mode < nd - 1) ? Xlead : X;
      if constexpr (std::is_same<T, double>::value) {
        if (fast) mttkrp_fast_f64(Xm, nd, shape, F, R, mode, M.data());
        else mttkrp_faithful(Xm, nd, shape, (const T* const*)F, R, mode, M.data());
      } else {
        mttkrp_faithful(Xm, nd, shape, (const T* const*)F, R, mode, M.data());
      }
      gram_hadamard((const T* const*)F, nd, shape, R, mode, G.data());
      if (cc.tikhonov_order > 0) tikhonov(G.data(), R, (T)cc.tikhonov_lambda, cc.tikhonov_order);
      else if (cc.l2 > 0.0)
        for (i64 i = 0; i < R; ++i) G[(size_t)(i * R + i)] += (T)cc.l2;
      const int path = solve_ls(M.data(), shape[mode], R, G.data(), F[mode], ridge_eps);
      if (path < 0) return 13;  // CpError::LinalgError (helpers.rs:485)
      worst = std::max(worst, path);
      if (cc.l1 > 0.0) soft_threshold(F[mode], shape[mode] * R, (T)cc.l1);
      if (cc.nonnegative)
        for (i64 x = 0; x < shape[mode] * R; ++x) F[mode][x] = std::max(F[mode][x], T(0));
      if (cc.orthogonal) orthonormalize(F[mode], shape[mode], R);
    }
    fit = fit_value(tnorm, recon_norm_sq((const T* const*)F, nd, shape, R),
                    inner_product(X, nd, shape, (const T* const*)F, R, fast != 0));
    if (fit_history) fit_history[iter] = (double)fit;
    if (iter > 0 && fit < prev_fit) osc += 1;
    T change = std::fabs(fit - prev_fit);
    final_change = change;
    if (iter > 0 && change < (T)tol) { reason = 0; break; }
    if (osc > 5 && iter > 10) { reason = 2; break; }
    prev_fit = fit;
  }
  if (res) {
    res->iters = iters;
    res->reason = reason;
    res->oscillation_count = osc;
    res->fit = (double)fit;
    res->final_fit_change = (double)final_change;
    res->solve_path_max = worst;
    res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return 0;
}

extern "C" {
int orc_cp_als_f64(const double* X, int nd, const int64_t* shape, int64_t R, double** F, int max_iters, double tol,
                   double time_limit, const orc_constraints* c, int fast, double* hist, orc_result* res,
                   double ridge_eps, const double* Xlead) {
  return cp_als_impl(X, nd, shape, R, F, max_iters, tol, time_limit, c, fast, hist, res, ridge_eps, Xlead);
}
int orc_cp_als_f32(const float* X, int nd, const int64_t* shape, int64_t R, float** F, int max_iters, double tol,
                   double time_limit, const orc_constraints* c, double* hist, orc_result* res, double ridge_eps) {
  return cp_als_impl(X, nd, shape, R, F, max_iters, tol, time_limit, c, 0, hist, res, ridge_eps);
}

// cp_reconstruct (tenrso-kernels/src/outer.rs:223-274 with outer_product 109-136,
// outer_product_weighted 159-170): result += w_r * prod_k v_k[i_k], r ascending.
void orc_cp_reconstruct_f64(const double* const* F, int nd, const int64_t* shape, int64_t R, const double* w,
                            double* out) {
  const i64 total = prod(shape, nd);
  for (i64 x = 0; x < total; ++x) out[x] = 0.0;
  std::vector<i64> idx(nd);
  for (i64 r = 0; r < R; ++r) {
    std::fill(idx.begin(), idx.end(), 0);
    for (i64 x = 0; x < total; ++x) {
      double p = 1.0;
      for (int k = 0; k < nd; ++k) p = p * F[k][idx[k] * R + r];
      if (w) p = p * w[r];
      out[x] = out[x] + p;
      for (int k = nd - 1; k >= 0; --k) {
        if (++idx[k] < shape[k]) break;
        idx[k] = 0;
      }
    }
  }
}

// extract_weights (cp/types.rs:421-446): normalise columns, accumulate norms.
// eps is T::epsilon() of the instantiation being modelled (:433).
void orc_extract_weights_f64(double** F, int nd, const int64_t* shape, int64_t R, double* w, double eps) {
  for (i64 r = 0; r < R; ++r) w[r] = 1.0;
  for (int k = 0; k < nd; ++k)
    for (i64 r = 0; r < R; ++r) {
      double s = 0.0;
      for (i64 i = 0; i < shape[k]; ++i) s = s + F[k][i * R + r] * F[k][i * R + r];
      double n = std::sqrt(s);
      if (n > eps) {
        w[r] = w[r] * n;
        for (i64 i = 0; i < shape[k]; ++i) F[k][i * R + r] = F[k][i * R + r] / n;
      }
    }
}

// CpReconstructionGrad::compute_factor_gradients (tenrso-ad/src/grad.rs:81-120):
// dA_n = unfold(G, n) . KR_except(n) (forward order, grad.rs:126-149), times weights.
int orc_cp_factor_grads_f64(const double* Gt, int nd, const int64_t* shape, const double* const* F, int64_t R,
                            const double* w, double** out) {
  for (int mode = 0; mode < nd; ++mode) {
    int rc = mttkrp_faithful(Gt, nd, shape, F, R, mode, out[mode]);
    if (rc) return rc;
    if (w)
      for (i64 i = 0; i < shape[mode]; ++i)
        for (i64 r = 0; r < R; ++r) out[mode][i * R + r] = out[mode][i * R + r] * w[r];
  }
  return 0;
}

// Sparse COO MTTKRP (the reference has no sparse MTTKRP; semantics follow
// CooTensor::to_dense (tenrso-sparse/src/coo.rs:267-286) composed with mttkrp):
// M[i_n, r] += x * prod_{k != n} A_k[i_k, r], FP64 accumulation.  Indices are
// SoA int32 (idx[k*nnz + e]).  Duplicates, if any, are summed (the dense form of
// a deduplicated tensor, coo.rs:320-348).
void orc_coo_mttkrp_f64(const int32_t* idx, const float* vals, int64_t nnz, int nd, const int64_t* shape,
                        const double* const* F, int64_t R, int mode, double* out) {
  for (i64 x = 0; x < shape[mode] * R; ++x) out[x] = 0.0;
#pragma omp parallel
  {
    std::vector<double> loc((size_t)(shape[mode] * R), 0.0);
#pragma omp for schedule(static)
    for (i64 e = 0; e < nnz; ++e) {
      const i64 row = idx[(i64)mode * nnz + e];
      const double v = vals[e];
      for (i64 r = 0; r < R; ++r) {
        double p = v;
        for (int k = 0; k < nd; ++k)
          if (k != mode) p *= F[k][(i64)idx[(i64)k * nnz + e] * R + r];
        loc[(size_t)(row * R + r)] += p;
      }
    }
#pragma omp critical
    for (i64 x = 0; x < shape[mode] * R; ++x) out[x] += loc[(size_t)x];
  }
}

}  // extern "C"
