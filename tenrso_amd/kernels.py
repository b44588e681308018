"""tenrso-kernels mirror: MTTKRP, Khatri-Rao and CP reconstruction on gfx950.

Same names, argument meaning and error behaviour as the reference functions
(tenrso-kernels/src/mttkrp.rs, khatri_rao.rs, outer.rs, utils.rs), for T = f32:

* numpy inputs take the host path (``*_host`` C entry points: copy in, compute on
  the GPU, copy out) and return numpy arrays;
* torch CUDA tensors take the device path on torch's current stream and return
  torch tensors (no host round trip).

Validation failures raise ``ValueError`` carrying the reference's message;
device failures raise ``TenrsoError``.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib

__all__ = [
    "mttkrp", "mttkrp_blocked", "mttkrp_blocked_parallel", "mttkrp_fused", "mttkrp_fused_parallel",
    "mttkrp_all_modes", "mttkrp_bf16", "to_bf16", "khatri_rao", "khatri_rao_parallel", "khatri_rao_blocked",
    "khatri_rao_blocked_parallel", "cp_reconstruct", "extract_weights", "mttkrp_streaming",
    "nmode_product", "nmode_products_seq", "tucker_operator", "tucker_operator_ordered", "tucker_reconstruct",
]


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _is_cuda(x) -> bool:
    return _is_torch(x) and bool(x.is_cuda)


def _to_numpy(x):
    """A torch tensor (any device) as a host numpy array (other inputs unchanged)."""
    return x.detach().cpu().numpy() if _is_torch(x) else x


def _dev_f32(t, device):
    """`t` (numpy array or torch tensor, any float dtype, any device) as a
    contiguous float32 CUDA tensor on `device`: what the device entry points read
    through raw pointers.  Factors of another dtype or on another device (the
    reference's usual T = f64, CPU tensors) are converted, never reinterpreted."""
    import torch
    if _is_torch(t):
        if t.is_complex():
            raise TypeError("complex factors are not supported")
        return t.detach().to(device=device, dtype=torch.float32).contiguous()
    return torch.as_tensor(np.ascontiguousarray(t, dtype=np.float32), device=device)


def _i64s(vals):
    return (C.c_int64 * len(vals))(*[int(v) for v in vals])


def _check_mttkrp(shape, factors, mode):
    """mttkrp.rs:85-122 validation (same order, same messages) via tenrso_mttkrp_check."""
    nf = len(factors)
    if nf == 0:
        raise ValueError(f"Number of factor matrices (0) must match tensor rank ({len(shape)})")
    rows = [int(f.shape[0]) for f in factors]
    cols = [int(f.shape[1]) if len(f.shape) == 2 else -1 for f in factors]
    rc = lib().tenrso_mttkrp_check(len(shape), _i64s(shape), nf, _i64s(rows), _i64s(cols), int(mode))
    if rc != _lib.OK:
        raise ValueError(_lib.last_error())
    return cols[0]


def _np32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def mttkrp(tensor, factors, mode: int):
    """tenrso_kernels::mttkrp (mttkrp.rs:74-146): V = X_(mode) . KR(factors except mode).

    Returns an (I_mode, R) array/tensor.
    """
    shape = tuple(int(s) for s in tensor.shape)
    R = _check_mttkrp(shape, factors, mode)
    if _is_torch(tensor) and not tensor.is_cuda:
        # a torch CPU tensor takes the host path and gets a torch CPU result
        import torch
        return torch.from_numpy(mttkrp(_to_numpy(tensor), [_to_numpy(f) for f in factors], mode))
    if _is_torch(tensor):
        import torch
        X = tensor.contiguous()
        if X.dtype != torch.float32:
            raise TypeError("device mttkrp expects a float32 CUDA tensor (the library computes in f32)")
        F = [_dev_f32(f, X.device) for f in factors]
        out = torch.empty((shape[mode], R), dtype=torch.float32, device=X.device)
        ctx = _lib.torch_context(X)
        ptrs = (C.c_void_p * len(F))(*[f.data_ptr() for f in F])
        check(lib().tenrso_mttkrp_f32(ctx.ptr, X.data_ptr(), len(shape), _i64s(shape), ptrs, R, mode,
                                      out.data_ptr()))
        return out
    X = _np32(tensor)
    F = [_np32(_to_numpy(f)) for f in factors]
    out = np.empty((shape[mode], R), dtype=np.float32)
    ctx = _lib.context_for(0)
    ptrs = (C.c_void_p * len(F))(*[f.ctypes.data for f in F])
    check(lib().tenrso_mttkrp_f32_host(ctx.ptr, X.ctypes.data, len(shape), _i64s(shape), ptrs, R, mode,
                                       out.ctypes.data))
    return out


def _tile_check(tile_size):
    if int(tile_size) == 0:
        raise ValueError("Tile size must be greater than 0")


def mttkrp_blocked(tensor, factors, mode: int, tile_size: int):
    """mttkrp_blocked (mttkrp.rs:395-498): same contract as mttkrp; tiling is internal."""
    _tile_check(tile_size)
    return mttkrp(tensor, factors, mode)


def mttkrp_blocked_parallel(tensor, factors, mode: int, tile_size: int):
    """mttkrp_blocked_parallel (mttkrp.rs:528-609)."""
    _tile_check(tile_size)
    return mttkrp(tensor, factors, mode)


def mttkrp_fused(tensor, factors, mode: int):
    """mttkrp_fused (mttkrp.rs:783-878): KR formed on the fly (as every path here)."""
    return mttkrp(tensor, factors, mode)


def mttkrp_fused_parallel(tensor, factors, mode: int):
    """mttkrp_fused_parallel (mttkrp.rs:903-992)."""
    return mttkrp(tensor, factors, mode)


def mttkrp_all_modes(tensor, factors):
    """mttkrp_all_modes (utils.rs:282-311)."""
    return [mttkrp(tensor, factors, m) for m in range(len(tensor.shape))]


def to_bf16(tensor):
    """Round a float32 CUDA tensor to bfloat16 (RNE) with the library's conversion kernel."""
    import torch
    X = tensor.contiguous()
    if not _is_cuda(X) or X.dtype != torch.float32:
        raise TypeError("to_bf16 expects a float32 CUDA tensor")
    out = torch.empty(X.shape, dtype=torch.bfloat16, device=X.device)
    ctx = _lib.torch_context(X)
    check(lib().tenrso_convert_f32_to_bf16(ctx.ptr, X.data_ptr(), X.numel(), out.data_ptr()))
    return out


def mttkrp_bf16(tensor_bf16, factors, mode: int):
    """BF16-input / FP32-accumulate MFMA MTTKRP (no reference counterpart).

    tensor_bf16: torch.bfloat16 CUDA tensor; factors: float32 CUDA tensors.
    """
    import torch
    shape = tuple(int(s) for s in tensor_bf16.shape)
    R = _check_mttkrp(shape, factors, mode)
    X = tensor_bf16.contiguous()
    if not _is_cuda(X) or X.dtype != torch.bfloat16:
        raise TypeError("mttkrp_bf16 expects a bfloat16 CUDA tensor (see to_bf16)")
    F = [_dev_f32(f, X.device) for f in factors]
    out = torch.empty((shape[mode], R), dtype=torch.float32, device=X.device)
    ctx = _lib.torch_context(X)
    ptrs = (C.c_void_p * len(F))(*[f.data_ptr() for f in F])
    check(lib().tenrso_mttkrp_bf16(ctx.ptr, X.data_ptr(), len(shape), _i64s(shape), ptrs, R, mode,
                                   out.data_ptr()))
    return out


def khatri_rao(a, b):
    """khatri_rao (khatri_rao.rs:56-89): C[ia*J + ib, k] = A[ia,k] * B[ib,k].

    The reference panics on a column mismatch; here that raises ValueError with
    the same message.
    """
    if a.shape[1] != b.shape[1]:
        raise ValueError(f"Number of columns must match: A has {a.shape[1]} columns, B has {b.shape[1]} columns")
    I, K = int(a.shape[0]), int(a.shape[1])
    J = int(b.shape[0])
    if _is_torch(a) and not a.is_cuda:
        import torch
        return torch.from_numpy(khatri_rao(_to_numpy(a), _to_numpy(b)))
    if _is_torch(a):
        import torch
        A = _dev_f32(a, a.device)
        B = _dev_f32(b, a.device)
        out = torch.empty((I * J, K), dtype=torch.float32, device=A.device)
        ctx = _lib.torch_context(A)
        check(lib().tenrso_khatri_rao_f32(ctx.ptr, A.data_ptr(), I, B.data_ptr(), J, K, out.data_ptr()))
        return out
    A, B = _np32(a), _np32(_to_numpy(b))
    out = np.empty((I * J, K), dtype=np.float32)
    if out.size:
        ctx = _lib.context_for(0)
        check(lib().tenrso_khatri_rao_f32_host(ctx.ptr, A.ctypes.data, I, B.ctypes.data, J, K, out.ctypes.data))
    return out


def khatri_rao_parallel(a, b):
    """khatri_rao_parallel (khatri_rao.rs:117-157): bit-identical to khatri_rao."""
    return khatri_rao(a, b)


def khatri_rao_blocked(a, b, block_size: int):
    """khatri_rao_blocked (khatri_rao.rs:193-255): panics on block_size 0."""
    if a.shape[1] != b.shape[1]:
        raise ValueError(f"Number of columns must match: A has {a.shape[1]} columns, B has {b.shape[1]} columns")
    if int(block_size) == 0:
        raise ValueError(f"Block size must be greater than 0, got {block_size}")
    return khatri_rao(a, b)


def khatri_rao_blocked_parallel(a, b, block_size: int):
    """khatri_rao_blocked_parallel (khatri_rao.rs:279-357)."""
    return khatri_rao_blocked(a, b, block_size)


def cp_reconstruct(factors, weights=None):
    """cp_reconstruct (outer.rs:223-274): sum_r w_r a_r^(0) o ... o a_r^(N-1)."""
    if len(factors) == 0:
        raise ValueError("Need at least one factor matrix for CP reconstruction")
    R = int(factors[0].shape[1])
    for i, f in enumerate(factors):
        if int(f.shape[1]) != R:
            raise ValueError(f"Factor {i} has {f.shape[1]} columns, expected {R}")
    if weights is not None and len(weights) != R:
        raise ValueError(f"Weights length {len(weights)} must match rank {R}")
    shape = [int(f.shape[0]) for f in factors]
    if _is_cuda(factors[0]):
        import torch
        device = factors[0].device
        F = [_dev_f32(f, device) for f in factors]
        w = None if weights is None else _dev_f32(weights, device)
        out = torch.empty(shape, dtype=torch.float32, device=device)
        ctx = _lib.torch_context(F[0])
        ptrs = (C.c_void_p * len(F))(*[f.data_ptr() for f in F])
        check(lib().tenrso_cp_reconstruct_f32(ctx.ptr, len(F), _i64s(shape), ptrs, R,
                                              None if w is None else w.data_ptr(), out.data_ptr()))
        return out
    import torch
    F = [_dev_f32(f, "cuda") for f in factors]
    w = None if weights is None else _dev_f32(weights, "cuda")
    out = cp_reconstruct(F, w).cpu()
    return out if _is_torch(factors[0]) else out.numpy()


def extract_weights(factors):
    """CpDecomp::extract_weights (cp/types.rs:421-446) on the device.

    Returns (normalised factors, weights): every column with norm > f32 epsilon
    divided by its norm, weights = product of the column norms over the factors.
    CUDA float32 contiguous factors are normalised in place; other inputs are
    copied to the device and the results come back in the input's kind (numpy or
    torch CPU).
    """
    if len(factors) == 0:
        raise ValueError("Need at least one factor matrix to extract weights")
    R = int(factors[0].shape[1])
    for i, f in enumerate(factors):
        if len(f.shape) != 2 or int(f.shape[1]) != R:
            raise ValueError(f"Factor {i} has {f.shape[1] if len(f.shape) == 2 else '?'} columns, expected {R}")
    import torch
    on_dev = _is_cuda(factors[0])
    device = factors[0].device if on_dev else torch.device("cuda")
    F = []
    for f in factors:
        if _is_cuda(f) and f.device == device and f.dtype == torch.float32 and f.is_contiguous():
            F.append(f)
        else:
            F.append(_dev_f32(f, device))
    w = torch.empty(R, dtype=torch.float32, device=device)
    ctx = _lib.torch_context(F[0])
    ptrs = (C.c_void_p * len(F))(*[f.data_ptr() for f in F])
    check(lib().tenrso_cp_extract_weights_f32(ctx.ptr, len(F), _i64s([f.shape[0] for f in F]), ptrs, R,
                                              w.data_ptr()))
    if on_dev:
        return F, w
    if _is_torch(factors[0]):
        return [f.cpu() for f in F], w.cpu()
    return [f.cpu().numpy() for f in F], w.cpu().numpy()


def mttkrp_streaming(tensor, factors, mode: int, device_budget_bytes: int, return_stats: bool = False):
    """Out-of-core MTTKRP (the streaming MTTKRP that tenrso-ooc's example
    mttkrp_streaming.rs:66-110 describes and leaves as a call of the in-memory
    mttkrp): the tensor stays in HOST memory (a numpy array or a torch CPU tensor,
    pinned or pageable) and streams through at most ``device_budget_bytes`` of
    device memory in mode-0 slabs, double-buffered against the MFMA MTTKRP of the
    previous slab.  Returns the (I_mode, R) result on the host (numpy, or torch
    CPU for a torch input) and, with ``return_stats``, a dict: seconds, h2d_GBs
    (tensor bytes moved per second), slab_rows, slabs, pinned (the input was
    pinned), host_copy_s (time spent staging a pageable input)."""
    shape = tuple(int(s) for s in tensor.shape)
    R = _check_mttkrp(shape, factors, mode)
    is_torch = _is_torch(tensor)
    if is_torch and tensor.is_cuda:
        raise TypeError("mttkrp_streaming takes a host tensor (use mttkrp for device tensors)")
    if is_torch:
        import torch
        X = tensor.detach()
        if X.dtype != torch.float32 or not X.is_contiguous():
            X = X.to(torch.float32).contiguous()
        xptr = X.data_ptr()
    else:
        X = _np32(tensor)
        xptr = X.ctypes.data
    F = [_np32(_to_numpy(f)) for f in factors]
    out = np.empty((shape[mode], R), dtype=np.float32)
    stats = np.zeros(8, dtype=np.float64)
    ctx = _lib.context_for(0)
    ptrs = (C.c_void_p * len(F))(*[f.ctypes.data for f in F])
    check(lib().tenrso_mttkrp_f32_streaming(ctx.ptr, xptr, len(shape), _i64s(shape), ptrs, R, int(mode),
                                            out.ctypes.data, int(device_budget_bytes), stats.ctypes.data))
    if is_torch:
        import torch
        out = torch.from_numpy(out)
    if not return_stats:
        return out
    return out, {"seconds": float(stats[0]), "h2d_GBs": float(stats[1]), "slab_rows": int(stats[2]),
                 "slabs": int(stats[3]), "pinned": bool(stats[4]), "host_copy_s": float(stats[5])}


# ---------------------------------------------------------------- n-mode products
def _ttm_chain(tensor, mats, modes):
    """Y = tensor x_{modes[0]} mats[0] x_{modes[1]} mats[1] ... in one C call
    (tenrso_nmode_products_seq_f32: batched FP32 GEMMs on the device, the
    intermediates in the context's buffers).  Checks every product against the
    shape the previous ones leave, with the messages of nmode.rs:82-98.  A CUDA
    tensor stays on its device (torch's current stream); a host tensor is copied
    to the current device and the result back."""
    import torch
    shape = [int(s) for s in tensor.shape]
    nd = len(shape)
    cur = list(shape)
    for M, m in zip(mats, modes):
        if m < 0 or m >= nd:
            raise ValueError(f"Mode {m} out of bounds for tensor with rank {nd}")
        if int(M.shape[1]) != cur[m]:
            raise ValueError(f"Matrix columns ({int(M.shape[1])}) must match tensor mode-{m} size ({cur[m]})")
        cur[m] = int(M.shape[0])
    on_dev = _is_cuda(tensor)
    dev = tensor.device if on_dev else torch.device("cuda", torch.cuda.current_device())
    X = _dev_f32(tensor, dev)
    Ms = [_dev_f32(M, dev) for M in mats]
    Y = torch.empty(cur, dtype=torch.float32, device=dev)
    ctx = _lib.torch_context(X)
    n = len(Ms)
    ptrs = (C.c_void_p * max(n, 1))(*[M.data_ptr() for M in Ms])
    out_shape = (C.c_int64 * max(nd, 1))()
    check(lib().tenrso_nmode_products_seq_f32(
        ctx.ptr, X.data_ptr(), nd, _i64s(shape), n, ptrs, _i64s([M.shape[0] for M in Ms] or [0]),
        _i64s([M.shape[1] for M in Ms] or [0]), (C.c_int * max(n, 1))(*[int(m) for m in modes]),
        Y.data_ptr(), out_shape))
    if list(out_shape)[:nd] != cur:
        raise RuntimeError(f"n-mode chain shape {list(out_shape)[:nd]} != {cur}")
    if on_dev:
        return Y
    out = Y.cpu()
    return out if _is_torch(tensor) else out.numpy()


def nmode_product(tensor, matrix, mode: int):
    """nmode_product (nmode.rs:72-130): Y = X x_mode M, M of shape (J, I_mode),
    Y[.., j, ..] = sum_i M[j, i] X[.., i, ..] (FP32 GEMM on the device)."""
    return _ttm_chain(tensor, [matrix], [int(mode)])


def nmode_products_seq(tensor, matrices):
    """nmode_products_seq (nmode.rs:241-255): ``matrices`` is a sequence of
    (matrix, mode) pairs applied in order."""
    return _ttm_chain(tensor, [m for m, _ in matrices], [int(k) for _, k in matrices])


def tucker_operator_ordered(tensor, mode_matrix_pairs):
    """tucker_operator_ordered (nmode.rs:517-530): (mode, matrix) pairs in order."""
    return _ttm_chain(tensor, [m for _, m in mode_matrix_pairs], [int(k) for k, _ in mode_matrix_pairs])


def tucker_operator(tensor, factor_matrices: dict):
    """tucker_operator (nmode.rs:442-485): the products of a {mode: matrix} map,
    modes checked first, then applied in ascending order of the ratio
    rows / I_mode (the most shrinking first).  The reference's ties follow its
    hash map's order; here ties go by ascending mode."""
    nd = len(tensor.shape)
    for m in factor_matrices:
        if m < 0 or m >= nd:
            raise ValueError(f"Mode {m} out of bounds for tensor with rank {nd}")
    order = sorted(factor_matrices, key=lambda m: (int(factor_matrices[m].shape[0]) / int(tensor.shape[m]), m))
    return _ttm_chain(tensor, [factor_matrices[m] for m in order], order)


def tucker_reconstruct(core, factors):
    """tucker_reconstruct (nmode.rs:576-612): X = G x_0 U_0 x_1 U_1 ... through
    tucker_operator, with the reference's checks and messages."""
    nd = len(core.shape)
    if len(factors) != nd:
        raise ValueError(f"Number of factor matrices ({len(factors)}) must match core rank ({nd})")
    for i, f in enumerate(factors):
        if int(f.shape[1]) != int(core.shape[i]):
            raise ValueError(f"Factor matrix {i} has {int(f.shape[1])} columns, expected {int(core.shape[i])} "
                             f"(core mode-{i} size)")
    return tucker_operator(core, dict(enumerate(factors)))
