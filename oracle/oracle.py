"""ctypes loader + numpy wrappers for liboracle.so (TEST INFRASTRUCTURE ONLY)."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

__all__ = [
    "lib", "mttkrp", "mttkrp_fast", "mttkrp_blocked_parallel", "khatri_rao", "gram", "gram_hadamard",
    "solve_ls", "tikhonov", "soft_threshold", "orthonormalize", "norm_sq", "recon_norm_sq", "fit",
    "cp_als", "cp_reconstruct", "extract_weights", "nmode_product", "nmode_products_seq", "cp_factor_grads", "coo_mttkrp", "num_threads", "svd_init",
    "hicoo_from_coo",
    "OrcConstraints", "REASONS", "LinalgError", "F32_EPS",
]

REASONS = ["FitTolerance", "MaxIterations", "Oscillation", "TimeLimit"]

_lib = None


class OrcConstraints(C.Structure):
    _fields_ = [("nonnegative", C.c_int), ("orthogonal", C.c_int), ("l2", C.c_double), ("l1", C.c_double),
                ("tikhonov_order", C.c_int), ("tikhonov_lambda", C.c_double)]


class OrcResult(C.Structure):
    _fields_ = [("iters", C.c_int), ("reason", C.c_int), ("oscillation_count", C.c_int), ("fit", C.c_double),
                ("final_fit_change", C.c_double), ("solve_path_max", C.c_int), ("seconds", C.c_double)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build_oracle
            build_oracle.build()
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _shape(shape):
    return (C.c_int64 * len(shape))(*shape)


def _ptrs(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


def _prep(X, factors, dtype):
    X = np.ascontiguousarray(X, dtype=dtype)
    F = [np.ascontiguousarray(f, dtype=dtype) for f in factors]
    return X, F


def num_threads():
    return lib().orc_num_threads()


def mttkrp(X, factors, mode, dtype=np.float64):
    """Faithful restatement of tenrso_kernels::mttkrp (mttkrp.rs:74-146)."""
    X, F = _prep(X, factors, dtype)
    R = F[0].shape[1]
    out = np.zeros((X.shape[mode], R), dtype=dtype)
    fn = lib().orc_mttkrp_f64 if dtype == np.float64 else lib().orc_mttkrp_f32
    rc = fn(_p(X), X.ndim, _shape(X.shape), _ptrs(F), C.c_int64(R), mode, _p(out))
    if rc:
        raise ValueError(f"oracle mttkrp failed rc={rc}")
    return out


def mttkrp_fast(X, factors, mode):
    """Cache-friendly FP64 MTTKRP (OpenMP); same sum as mttkrp.rs:74-146."""
    X, F = _prep(X, factors, np.float64)
    R = F[0].shape[1]
    out = np.zeros((X.shape[mode], R), dtype=np.float64)
    rc = lib().orc_mttkrp_fast_f64(_p(X), X.ndim, _shape(X.shape), _ptrs(F), C.c_int64(R), mode, _p(out))
    if rc:
        raise ValueError(f"oracle mttkrp_fast failed rc={rc}")
    return out


def mttkrp_blocked_parallel(X, factors, mode, tile=64):
    """OpenMP restatement of mttkrp_blocked_parallel (mttkrp.rs:528-609), f32."""
    X, F = _prep(X, factors, np.float32)
    R = F[0].shape[1]
    out = np.zeros((X.shape[mode], R), dtype=np.float32)
    rc = lib().orc_mttkrp_blocked_parallel_f32(_p(X), X.ndim, _shape(X.shape), _ptrs(F), C.c_int64(R), mode,
                                               C.c_int64(tile), _p(out))
    if rc:
        raise ValueError(f"oracle mttkrp_blocked_parallel failed rc={rc}")
    return out


def khatri_rao(A, B, dtype=np.float64):
    """khatri_rao (khatri_rao.rs:56-89)."""
    A = np.ascontiguousarray(A, dtype=dtype)
    B = np.ascontiguousarray(B, dtype=dtype)
    assert A.shape[1] == B.shape[1]
    out = np.zeros((A.shape[0] * B.shape[0], A.shape[1]), dtype=dtype)
    fn = lib().orc_khatri_rao_f64 if dtype == np.float64 else lib().orc_khatri_rao_f32
    fn(_p(A), C.c_int64(A.shape[0]), _p(B), C.c_int64(B.shape[0]), C.c_int64(A.shape[1]), _p(out))
    return out


def gram(Fm):
    """compute_gram_matrix (cp/helpers.rs:427-452)."""
    Fm = np.ascontiguousarray(Fm, dtype=np.float64)
    G = np.zeros((Fm.shape[1], Fm.shape[1]))
    lib().orc_gram_f64(_p(Fm), C.c_int64(Fm.shape[0]), C.c_int64(Fm.shape[1]), _p(G))
    return G


def gram_hadamard(factors, skip):
    """compute_gram_hadamard (cp/helpers.rs:409-424)."""
    F = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
    R = F[0].shape[1]
    out = np.zeros((R, R))
    shape = [f.shape[0] for f in F]
    lib().orc_gram_hadamard_f64(_ptrs(F), len(F), _shape(shape), C.c_int64(R), skip, _p(out))
    return out


class LinalgError(ValueError):
    """CpError::LinalgError (cp/types.rs:13-35): a solve failed after the ridge retry."""


def solve_ls(M, G, ridge_eps=None):
    """solve_least_squares (cp/helpers.rs:455-495); returns (X, path), path 0 for the
    plain solve and 1 for the ridge retry (Gamma + eps 10R I, eps = ridge_eps or the
    FP64 epsilon).  Raises LinalgError when the retry fails too (helpers.rs:485)."""
    M = np.ascontiguousarray(M, dtype=np.float64)
    G = np.ascontiguousarray(G, dtype=np.float64)
    out = np.zeros_like(M)
    path = lib().orc_solve_ls_f64(_p(M), C.c_int64(M.shape[0]), C.c_int64(M.shape[1]), _p(G), _p(out),
                                  C.c_double(0.0 if ridge_eps is None else float(ridge_eps)))
    if path < 0:
        raise LinalgError("Linear algebra error: lstsq failed after the ridge retry")
    return out, path


def tikhonov(G, lam, order):
    """apply_tikhonov_to_gram (cp/helpers.rs:630-702)."""
    G = np.array(G, dtype=np.float64, order="C")
    lib().orc_tikhonov_f64(_p(G), C.c_int64(G.shape[0]), C.c_double(lam), C.c_int64(order))
    return G


def soft_threshold(Fm, th):
    """soft_threshold (cp/helpers.rs:610-623)."""
    Fm = np.array(Fm, dtype=np.float64, order="C")
    lib().orc_soft_threshold_f64(_p(Fm), C.c_int64(Fm.size), C.c_double(th))
    return Fm


def orthonormalize(Fm):
    Fm = np.array(Fm, dtype=np.float64, order="C")
    lib().orc_orthonormalize_f64(_p(Fm), C.c_int64(Fm.shape[0]), C.c_int64(Fm.shape[1]))
    return Fm


lib_restype_done = False


def _restypes():
    global lib_restype_done
    if not lib_restype_done:
        L = lib()
        L.orc_norm_sq_f64.restype = C.c_double
        L.orc_norm_sq_f32.restype = C.c_float
        L.orc_recon_norm_sq_f64.restype = C.c_double
        L.orc_fit_f64.restype = C.c_double
        lib_restype_done = True


def norm_sq(X, dtype=np.float64):
    """compute_norm_squared (cp/helpers.rs:498-510), sequential in dtype."""
    _restypes()
    X = np.ascontiguousarray(X, dtype=dtype)
    fn = lib().orc_norm_sq_f64 if dtype == np.float64 else lib().orc_norm_sq_f32
    return float(fn(_p(X), C.c_int64(X.size)))


def recon_norm_sq(factors):
    """compute_reconstruction_norm_squared (cp/helpers.rs:533-555)."""
    _restypes()
    F = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
    shape = [f.shape[0] for f in F]
    return float(lib().orc_recon_norm_sq_f64(_ptrs(F), len(F), _shape(shape), C.c_int64(F[0].shape[1])))


def fit(X, factors, tensor_norm_sq=None, fast=True):
    """compute_fit (cp/helpers.rs:513-530) in FP64."""
    _restypes()
    X, F = _prep(X, factors, np.float64)
    if tensor_norm_sq is None:
        tensor_norm_sq = norm_sq(X)
    return float(lib().orc_fit_f64(_p(X), X.ndim, _shape(X.shape), _ptrs(F), C.c_int64(F[0].shape[1]),
                                   C.c_double(tensor_norm_sq), int(fast)))


F32_EPS = float(np.finfo(np.float32).eps)


def cp_als(X, init_factors, max_iters, tol, time_limit=None, nonnegative=False, orthogonal=False, l2=0.0,
           l1=0.0, tikhonov_order=0, tikhonov_lambda=0.0, dtype=np.float64, fast=True, ridge_eps=F32_EPS,
           x_lead=None):
    """cp_als / cp_als_constrained (cp/core.rs:59-182, 225-408) from explicit initial factors.

    ridge_eps: the T::epsilon() of the ridge retry (helpers.rs:477-490).  It
    defaults to f32's, the epsilon of the T = f32 instantiation that the device
    library replaces; None takes the epsilon of `dtype`.
    x_lead: optional tensor that the MTTKRPs of modes 0..N-2 contract instead of
    X (FP64 model of the BF16 ALS: leading modes on bf16(X), the last mode, the
    norm and the fit on X itself).
    Returns (factors, info) where info mirrors CpDecomp + ConvergenceInfo.
    Raises LinalgError when a solve fails after its ridge retry.
    """
    X, F = _prep(X, init_factors, dtype)
    F = [f.copy() for f in F]
    R = F[0].shape[1]
    cons = OrcConstraints(int(nonnegative), int(orthogonal), float(l2), float(l1), int(tikhonov_order),
                          float(tikhonov_lambda))
    res = OrcResult()
    hist = np.zeros(max(max_iters, 1), dtype=np.float64)
    tl = -1.0 if time_limit is None else float(time_limit)
    eps = C.c_double(0.0 if ridge_eps is None else float(ridge_eps))
    if dtype == np.float64:
        lead = None
        if x_lead is not None:
            lead = np.ascontiguousarray(x_lead, dtype=np.float64)
            if lead.shape != X.shape:
                raise ValueError("x_lead must have the shape of X")
        rc = lib().orc_cp_als_f64(_p(X), X.ndim, _shape(X.shape), C.c_int64(R), _ptrs(F), max_iters,
                                  C.c_double(tol), C.c_double(tl), C.byref(cons), int(fast), _p(hist),
                                  C.byref(res), eps, None if lead is None else _p(lead))
    else:
        if x_lead is not None:
            raise ValueError("x_lead is supported by the FP64 oracle only")
        rc = lib().orc_cp_als_f32(_p(X), X.ndim, _shape(X.shape), C.c_int64(R), _ptrs(F), max_iters,
                                  C.c_double(tol), C.c_double(tl), C.byref(cons), _p(hist), C.byref(res), eps)
    if rc == 13:
        raise LinalgError("Linear algebra error: lstsq failed after the ridge retry")
    if rc:
        raise ValueError({10: "InvalidRank", 11: "InvalidTolerance", 12: "NonnegativeViolation"}.get(rc, rc))
    info = dict(iters=res.iters, reason=REASONS[res.reason], oscillation_count=res.oscillation_count,
                oscillated=res.oscillation_count > 0, fit=res.fit, final_fit_change=res.final_fit_change,
                fit_history=hist[:res.iters].copy(), solve_path_max=res.solve_path_max, seconds=res.seconds)
    return F, info


def cp_reconstruct(factors, weights=None):
    """cp_reconstruct (tenrso-kernels/src/outer.rs:223-274)."""
    F = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
    shape = [f.shape[0] for f in F]
    out = np.zeros(shape, dtype=np.float64)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    lib().orc_cp_reconstruct_f64(_ptrs(F), len(F), _shape(shape), C.c_int64(F[0].shape[1]),
                                 None if w is None else _p(w), _p(out))
    return out


def extract_weights(factors, eps=float(np.finfo(np.float32).eps)):
    """CpDecomp::extract_weights (cp/types.rs:421-446); returns (factors, weights).

    FP64 arithmetic; eps is the T::epsilon() threshold of the instantiation being
    modelled (:433): f32's by default (the instantiation the device replaces)."""
    F = [np.array(f, dtype=np.float64, order="C") for f in factors]
    shape = [f.shape[0] for f in F]
    R = F[0].shape[1]
    w = np.zeros(R)
    lib().orc_extract_weights_f64(_ptrs(F), len(F), _shape(shape), C.c_int64(R), _p(w), C.c_double(eps))
    return F, w


def nmode_product(X, M, mode):
    """nmode_product (tenrso-kernels/src/nmode.rs:72-130) in FP64: Y = X x_mode M,
    Y[.., j, ..] = sum_i M[j, i] X[.., i, ..], with the reference's checks and
    messages (:82-98).  numpy tensordot over the mode, moved back into place."""
    X = np.asarray(X, dtype=np.float64)
    M = np.asarray(M, dtype=np.float64)
    if mode >= X.ndim:
        raise ValueError(f"Mode {mode} out of bounds for tensor with rank {X.ndim}")
    if M.shape[1] != X.shape[mode]:
        raise ValueError(f"Matrix columns ({M.shape[1]}) must match tensor mode-{mode} size ({X.shape[mode]})")
    return np.moveaxis(np.tensordot(M, X, axes=(1, mode)), 0, mode)


def nmode_products_seq(X, pairs):
    """nmode_products_seq (nmode.rs:241-255): the products (matrix, mode) in order."""
    Y = np.asarray(X, dtype=np.float64)
    for M, mode in pairs:
        Y = nmode_product(Y, M, mode)
    return Y


def cp_factor_grads(grad_output, factors, weights=None):
    """CpReconstructionGrad::compute_factor_gradients (tenrso-ad/src/grad.rs:81-120)."""
    G, F = _prep(grad_output, factors, np.float64)
    R = F[0].shape[1]
    outs = [np.zeros((G.shape[k], R)) for k in range(G.ndim)]
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    rc = lib().orc_cp_factor_grads_f64(_p(G), G.ndim, _shape(G.shape), _ptrs(F), C.c_int64(R),
                                       None if w is None else _p(w), _ptrs(outs))
    if rc:
        raise ValueError(f"oracle cp_factor_grads failed rc={rc}")
    return outs


def coo_mttkrp(indices, values, shape, factors, mode):
    """FP64 COO MTTKRP: dense semantics of CooTensor::to_dense (coo.rs:267-286) + mttkrp.

    indices: (ndim, nnz) int32 SoA; values: (nnz,) float32.
    """
    idx = np.ascontiguousarray(indices, dtype=np.int32)
    vals = np.ascontiguousarray(values, dtype=np.float32)
    F = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
    R = F[0].shape[1]
    out = np.zeros((shape[mode], R))
    lib().orc_coo_mttkrp_f64(_p(idx), _p(vals), C.c_int64(vals.size), len(shape), _shape(shape), _ptrs(F),
                             C.c_int64(R), mode, _p(out))
    return out


def svd_init(X, rank, kind):
    """initialize_factors InitStrategy::Svd / Nnsvd / LeverageScore
    (tenrso-decomp cp/helpers.rs:232-373) in FP64, with numpy's economy SVD of each
    mode-k unfolding in place of the third-party one (`svd(&unfolded, false, None)`).

    The SVD's signs are arbitrary; here, as on the device, each u_r is turned so
    that its first entry of largest magnitude is positive (v_r with it).  Nnsvd
    does not depend on that choice.  Columns past min(rank, I_k, J_k) are NaN: the
    reference fills them from an unseeded thread_rng.
    kind: "svd", "nnsvd" or "leverage_score".
    """
    X = np.asarray(X, dtype=np.float64)
    out = []
    for k, n in enumerate(X.shape):
        Xk = np.moveaxis(X, k, 0).reshape(n, -1)
        U, s, Vt = np.linalg.svd(Xk, full_matrices=False)
        ar = min(rank, U.shape[1], Vt.shape[0])
        for r in range(ar):
            i = int(np.argmax(np.abs(U[:, r])))
            if U[i, r] < 0:
                U[:, r] = -U[:, r]
                Vt[r] = -Vt[r]
        F = np.full((n, rank), np.nan)
        if kind == "svd":
            F[:, :ar] = U[:, :ar]
        elif kind == "nnsvd":
            for r in range(ar):
                u, v = U[:, r], Vt[r]
                up, un = np.where(u > 0, u, 0.0), np.where(u > 0, 0.0, -u)   # split_sign (helpers.rs:380-398)
                vp, vn = np.where(v > 0, v, 0.0), np.where(v > 0, 0.0, -v)
                pp = np.linalg.norm(up) * np.linalg.norm(vp)
                npr = np.linalg.norm(un) * np.linalg.norm(vn)
                F[:, r] = up * np.sqrt(s[r] * pp) if pp >= npr else un * np.sqrt(s[r] * npr)
        elif kind == "leverage_score":
            lev = (U[:, :ar] ** 2).sum(axis=1) / ar
            tot = lev.sum()
            if tot > np.finfo(np.float32).eps:
                lev = lev / tot
            for r in range(ar):
                F[:, r] = U[:, r] * np.sqrt(s[r]) * np.sqrt(lev * n)
        else:
            raise ValueError(kind)
        out.append(F)
    return out


def hicoo_from_coo(indices, values, shape, block_shape):
    """HiCooTensor::from_coo (tenrso-sparse/src/hicoo.rs:121-205) restated with numpy:
    block = index // bs, local = index % bs, a stable sort of the nonzeros by block
    coordinate (lexicographic; sort_by is stable, so a block keeps COO order),
    block_ptrs at each change of block plus the final nnz.  Returns
    (block_coords (ndim, B), block_ptrs (B + 1), local (ndim, nnz), values)."""
    idx = np.asarray(indices, np.int64).reshape(len(shape), -1)
    vals = np.asarray(values)
    bs = np.asarray(block_shape, np.int64)
    if bs.size != len(shape) or (bs <= 0).any():
        raise ValueError("invalid block shape")
    nnz = vals.size
    if nnz == 0:
        return np.zeros((len(shape), 0), np.int64), np.zeros(1, np.int64), np.zeros((len(shape), 0), np.int64), vals
    blk = idx // bs[:, None]
    loc = idx % bs[:, None]
    order = np.lexsort(blk[::-1])  # lexsort is stable
    sb, sl, sv = blk[:, order], loc[:, order], vals[order]
    change = np.ones(nnz, bool)
    change[1:] = (sb[:, 1:] != sb[:, :-1]).any(axis=0)
    starts = np.nonzero(change)[0]
    return sb[:, starts], np.append(starts, nnz).astype(np.int64), sl, sv
