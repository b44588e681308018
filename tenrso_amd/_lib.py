"""ctypes binding of libtenrso_hip.so (the C ABI of include/tenrso_hip.h).

The library is loaded from this package directory.  There is no fallback: if the
library is missing, or no gfx950 GPU is present, every compute call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtenrso_hip.so")

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_INVALID_RANK = 2
ERR_INVALID_TOLERANCE = 3
ERR_SHAPE_MISMATCH = 4
ERR_LINALG = 5
ERR_NONNEGATIVE = 6
ERR_INVALID_REGULARIZATION = 7
ERR_DEVICE = 8
ERR_UNSUPPORTED = 9

PREC_FP32 = 0
PREC_BF16 = 1
INIT_SVD, INIT_NNSVD, INIT_LEVERAGE = 2, 3, 4  # enum tenrso_init_kind
AD_MTTKRP, AD_CP_RECONSTRUCT, AD_CP_LOSS = 0, 1, 2  # enum tenrso_ad_kind

REASONS = ["FitTolerance", "MaxIterations", "Oscillation", "TimeLimit"]

# exported symbols, with ctypes signatures (kept in sync with include/tenrso_hip.h)
_vp, _i, _i64, _d = C.c_void_p, C.c_int, C.c_int64, C.c_double


class CpOptions(C.Structure):
    _fields_ = [("rank", C.c_int64), ("max_iters", C.c_int32), ("precision", C.c_int32), ("tol", C.c_double),
                ("time_limit_s", C.c_double), ("nonnegative", C.c_int32), ("validate_nonneg_input", C.c_int32),
                ("orthogonal", C.c_int32), ("tikhonov_order", C.c_int32), ("l2", C.c_double), ("l1", C.c_double),
                ("tikhonov_lambda", C.c_double), ("fit_from_mode0", C.c_int32), ("reserved", C.c_int32)]


class CpResult(C.Structure):
    _fields_ = [("iters", C.c_int32), ("reason", C.c_int32), ("oscillation_count", C.c_int32),
                ("oscillated", C.c_int32), ("fit", C.c_double), ("final_fit_change", C.c_double),
                ("solve_path_max", C.c_int32), ("reserved", C.c_int32), ("seconds", C.c_double)]


SIGNATURES = {
    "tenrso_last_error": ([], C.c_char_p),
    "tenrso_version": ([], C.c_char_p),
    "tenrso_ctx_create": ([_i, _i, _vp, C.POINTER(_vp)], _i),
    "tenrso_ctx_destroy": ([_vp], _i),
    "tenrso_ctx_synchronize": ([_vp], _i),
    "tenrso_ctx_set_timing": ([_vp, _i], _i),
    "tenrso_ctx_last_kernel_ms": ([_vp, C.POINTER(_d)], _i),
    "tenrso_mttkrp_check": ([_i, _vp, _i, _vp, _vp, _i], _i),
    "tenrso_mttkrp_f32": ([_vp, _vp, _i, _vp, _vp, _i64, _i, _vp], _i),
    "tenrso_mttkrp_f32_host": ([_vp, _vp, _i, _vp, _vp, _i64, _i, _vp], _i),
    "tenrso_mttkrp_f32_f64out": ([_vp, _vp, _i, _vp, _vp, _i64, _i, _vp], _i),
    "tenrso_mttkrp_bf16": ([_vp, _vp, _i, _vp, _vp, _i64, _i, _vp], _i),
    "tenrso_convert_f32_to_bf16": ([_vp, _vp, _i64, _vp], _i),
    "tenrso_mttkrp_profile": ([_vp, _i, _vp, _i, _vp, _vp, _i64, _i, _i, C.POINTER(_d)], _i),
    "tenrso_mttkrp_f32_noalloc": ([_vp, _vp, _i, _vp, _vp, _i64, _i, _vp], _i),
    "tenrso_mttkrp_all_modes_f32": ([_vp, _vp, _i, _vp, _vp, _i64, _vp], _i),
    "tenrso_mttkrp_f32_streaming": ([_vp, _vp, _i, _vp, _vp, _i64, _i, _vp, _i64, _vp], _i),
    "tenrso_nmode_product_f32": ([_vp, _vp, _i, _vp, _vp, _i64, _i64, _i, _vp], _i),
    "tenrso_nmode_products_seq_f32": ([_vp, _vp, _i, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "tenrso_khatri_rao_f32": ([_vp, _vp, _i64, _vp, _i64, _i64, _vp], _i),
    "tenrso_khatri_rao_f32_host": ([_vp, _vp, _i64, _vp, _i64, _i64, _vp], _i),
    "tenrso_cp_reconstruct_f32": ([_vp, _i, _vp, _vp, _i64, _vp, _vp], _i),
    "tenrso_cp_extract_weights_f32": ([_vp, _i, _vp, _vp, _i64, _vp], _i),
    "tenrso_cp_als_f32": ([_vp, _vp, _i, _vp, _vp, C.POINTER(CpOptions), C.POINTER(CpResult), _vp], _i),
    "tenrso_cp_als_f32_host": ([_vp, _vp, _i, _vp, _vp, C.POINTER(CpOptions), C.POINTER(CpResult), _vp], _i),
    "tenrso_cp_init_svd": ([_vp, _vp, _i, _vp, _i64, _i, C.c_uint64, _vp], _i),
    "tenrso_mttkrp_vjp_f32": ([_vp, _vp, _i, _vp, _vp, _i64, _i, _vp, _vp, _vp], _i),
    "tenrso_cp_factor_grads_f32": ([_vp, _vp, _i, _vp, _vp, _i64, _vp, _vp], _i),
    "tenrso_cp_loss_grad_f32": ([_vp, _vp, _i, _vp, _vp, _i64, _vp, C.POINTER(_d)], _i),
    "tenrso_cp_gd_f32": ([_vp, _vp, _i, _vp, _vp, _i64, _i, _d, _vp], _i),
    "tenrso_ad_op_create": ([_vp, _i, _vp, _i, _vp, _vp, _i64, _i, _vp, C.POINTER(_vp)], _i),
    "tenrso_ad_op_destroy": ([_vp], _i),
    "tenrso_ad_op_name": ([_vp], C.c_char_p),
    "tenrso_ad_op_num_inputs": ([_vp], _i),
    "tenrso_ad_op_num_outputs": ([_vp], _i),
    "tenrso_ad_op_input_numel": ([_vp, _i], _i64),
    "tenrso_ad_op_output_numel": ([_vp, _i], _i64),
    "tenrso_ad_op_forward": ([_vp, _vp], _i),
    "tenrso_ad_op_backward": ([_vp, _vp, _vp], _i),
    "tenrso_sp_create": ([_vp, _i, _vp, _i64, _vp, _vp, _i, C.POINTER(_vp)], _i),
    "tenrso_sp_destroy": ([_vp], _i),
    "tenrso_sp_mode_arrays": ([_vp, _vp, _i, _vp, _vp, _vp, _vp], _i),
    "tenrso_sp_csf": ([_vp, _vp, _i, _vp, _vp, _vp, _vp], _i),
    "tenrso_sp_hicoo": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "tenrso_sp_fibres": ([_vp, _i, _vp], _i),
    "tenrso_sp_mttkrp_f32": ([_vp, _vp, _vp, _i64, _i, _vp], _i),
    "tenrso_sp_cp_als_f32": ([_vp, _vp, _vp, C.POINTER(CpOptions), C.POINTER(CpResult), _vp], _i),
}

_lib = None
_lock = threading.Lock()


class TenrsoError(RuntimeError):
    """A failing C-ABI call; `code` is the tenrso_status value."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def lib():
    """Load libtenrso_hip.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise ImportError(f"{LIB_PATH} is missing: build it with `python -m tenrso_amd.build_lib` "
                                      "(or __graft_entry__.build())")
                L = C.CDLL(LIB_PATH)
                for name, (args, res) in SIGNATURES.items():
                    fn = getattr(L, name)
                    fn.argtypes = args
                    fn.restype = res
                _lib = L
    return _lib


def last_error() -> str:
    return lib().tenrso_last_error().decode()


def check(rc: int):
    if rc != OK:
        raise TenrsoError(rc, last_error())


class Context:
    """A tenrso_ctx: one HIP stream (owned, or borrowed from torch) + device scratch."""

    def __init__(self, device: int = 0, stream: int | None = None):
        self.ptr = C.c_void_p()
        borrow = 0 if stream is None else 1
        check(lib().tenrso_ctx_create(device, borrow, C.c_void_p(stream or 0), C.byref(self.ptr)))
        self.device = device

    def synchronize(self):
        check(lib().tenrso_ctx_synchronize(self.ptr))

    def set_timing(self, on: bool):
        check(lib().tenrso_ctx_set_timing(self.ptr, int(on)))

    def last_kernel_ms(self) -> float:
        v = C.c_double()
        check(lib().tenrso_ctx_last_kernel_ms(self.ptr, C.byref(v)))
        return v.value

    def close(self):
        if self.ptr:
            lib().tenrso_ctx_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_contexts: dict = {}


def context_for(device: int = 0, stream: int | None = None) -> Context:
    """Cached context per (device, stream).  stream=None: a private stream (host APIs)."""
    key = (device, stream)
    ctx = _contexts.get(key)
    if ctx is None:
        ctx = Context(device, stream)
        _contexts[key] = ctx
    return ctx


def torch_context(tensor) -> Context:
    """Context that runs on torch's current stream of the tensor's device."""
    import torch
    dev = tensor.device.index if tensor.device.index is not None else torch.cuda.current_device()
    return context_for(dev, torch.cuda.current_stream(dev).cuda_stream)
