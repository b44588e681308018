"""tenrso-decomp CP mirror: cp_als / cp_als_constrained on gfx950.

Types and rules follow tenrso-decomp/src/cp (types.rs, core.rs, helpers.rs):
``CpError`` variants and messages, ``InitStrategy``, ``RegularizationType``,
``CpConstraints`` (effective_l2 / effective_l1), ``ConvergenceReason``,
``ConvergenceInfo`` and ``CpDecomp`` (reconstruct, extract_weights).

The ALS iteration itself runs in the HIP library (``tenrso_cp_als_f32``).  The
reference draws initial factors from an unseeded ``thread_rng``
(cp/helpers.rs:212-231); here they are drawn from numpy's PCG64 with an optional
``seed`` (None = fresh entropy), and ``cp_als_with_init`` takes explicit factors.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib
from ._lib import CpOptions, CpResult, lib

__all__ = [
    "CpError", "InvalidRank", "InvalidTolerance", "LinalgError", "ShapeMismatch", "ConvergenceFailed",
    "NonnegativeViolation", "InvalidRegularization", "InitStrategy", "RegularizationType", "CpConstraints",
    "ConvergenceReason", "ConvergenceInfo", "CpDecomp", "cp_als", "cp_als_constrained", "cp_als_with_init",
    "initialize_factors",
]


# ---------------------------------------------------------------- errors (cp/types.rs:13-35)
class CpError(Exception):
    pass


class InvalidRank(CpError):
    def __init__(self, rank):
        super().__init__(f"Invalid rank: {rank}")
        self.rank = rank


class InvalidTolerance(CpError):
    def __init__(self, tol):
        super().__init__(f"Invalid tolerance: {tol}")
        self.tol = tol


class LinalgError(CpError):
    def __init__(self, msg):
        super().__init__(f"Linear algebra error: {msg}")


class ShapeMismatch(CpError):
    def __init__(self, msg):
        super().__init__(f"Shape mismatch: {msg}")


class ConvergenceFailed(CpError):
    def __init__(self, iters):
        super().__init__(f"Convergence failed after {iters} iterations")


class NonnegativeViolation(CpError):
    def __init__(self):
        super().__init__("Non-negative constraint violated: input tensor contains negative values")


class InvalidRegularization(CpError):
    def __init__(self, msg):
        super().__init__(f"Invalid regularization parameter: {msg}")


# ---------------------------------------------------------------- config types
class InitStrategy(enum.Enum):
    """cp/types.rs:38-59."""
    Random = "random"              # U[0, 1)
    RandomNormal = "random_normal"  # N(0, 1)
    Svd = "svd"
    Nnsvd = "nnsvd"
    LeverageScore = "leverage_score"


@dataclass(frozen=True)
class RegularizationType:
    """cp/types.rs:64-193.  kind in {"None", "L2", "L1", "ElasticNet", "Tikhonov"}."""
    kind: str = "None"
    lam: float = 0.0
    alpha: float = 0.0
    order: int = 0

    @staticmethod
    def none():
        return RegularizationType("None")

    @staticmethod
    def l2(lam):
        return RegularizationType("L2", lam)

    @staticmethod
    def l1(lam):
        return RegularizationType("L1", lam)

    @staticmethod
    def elastic_net(lam, alpha):
        return RegularizationType("ElasticNet", lam, alpha)

    @staticmethod
    def tikhonov(lam, order):
        return RegularizationType("Tikhonov", lam, 0.0, order)

    def validate(self):
        """RegularizationType::validate (types.rs:117-168)."""
        if self.kind == "None":
            return
        name = {"L2": "L2", "L1": "L1", "ElasticNet": "ElasticNet", "Tikhonov": "Tikhonov"}[self.kind]
        if self.lam < 0.0:
            raise InvalidRegularization(f"{name} lambda must be >= 0, got {self.lam}")
        if self.kind == "ElasticNet" and not (0.0 <= self.alpha <= 1.0):
            raise InvalidRegularization(f"ElasticNet alpha must be in [0, 1], got {self.alpha}")

    def l2_component(self):
        """types.rs:171-182."""
        return {"None": 0.0, "L2": self.lam, "L1": 0.0, "ElasticNet": (1.0 - self.alpha) * self.lam,
                "Tikhonov": self.lam if self.order == 0 else 0.0}[self.kind]

    def l1_component(self):
        """types.rs:185-193."""
        return {"None": 0.0, "L2": 0.0, "L1": self.lam, "ElasticNet": self.alpha * self.lam,
                "Tikhonov": 0.0}[self.kind]


@dataclass
class CpConstraints:
    """cp/types.rs:198-309."""
    nonnegative: bool = False
    l2_reg: float = 0.0
    orthogonal: bool = False
    regularization: RegularizationType = field(default_factory=RegularizationType.none)
    validate_nonneg_input: bool = True

    @staticmethod
    def nonnegative_():
        return CpConstraints(nonnegative=True)

    @staticmethod
    def l2_regularized(lam):
        return CpConstraints(l2_reg=lam, regularization=RegularizationType.l2(lam))

    @staticmethod
    def l1_regularized(lam):
        return CpConstraints(regularization=RegularizationType.l1(lam))

    @staticmethod
    def elastic_net(lam, alpha):
        return CpConstraints(regularization=RegularizationType.elastic_net(lam, alpha))

    @staticmethod
    def tikhonov(lam, order):
        return CpConstraints(regularization=RegularizationType.tikhonov(lam, order))

    @staticmethod
    def orthogonal_():
        return CpConstraints(orthogonal=True)

    def effective_l2(self):
        """types.rs:288-293."""
        return self.l2_reg if self.regularization.kind == "None" else self.regularization.l2_component()

    def effective_l1(self):
        return self.regularization.l1_component()


class ConvergenceReason(enum.Enum):
    FitTolerance = 0
    MaxIterations = 1
    Oscillation = 2
    TimeLimit = 3


@dataclass
class ConvergenceInfo:
    """cp/types.rs:328-344."""
    fit_history: List[float]
    reason: ConvergenceReason
    oscillated: bool
    oscillation_count: int
    final_fit_change: float


@dataclass
class CpDecomp:
    """cp/types.rs:349-446."""
    factors: list
    weights: Optional[object]
    fit: float
    iters: int
    convergence: Optional[ConvergenceInfo]
    solve_path_max: int = 0
    seconds: float = 0.0

    def reconstruct(self, shape):
        """CpDecomp::reconstruct (types.rs:384-415) via the device cp_reconstruct."""
        from .kernels import cp_reconstruct
        if len(self.factors) != len(shape):
            raise ValueError(f"Shape rank mismatch: expected {len(self.factors)} modes, got {len(shape)}")
        for i, f in enumerate(self.factors):
            if int(f.shape[0]) != int(shape[i]):
                raise ValueError(f"Mode-{i} size mismatch: expected {shape[i]}, got {f.shape[0]}")
        return cp_reconstruct(self.factors, self.weights)

    def extract_weights(self):
        """CpDecomp::extract_weights (types.rs:421-446): unit-norm columns, weights =
        product of the column norms, computed by the device kernel
        (tenrso_cp_extract_weights_f32)."""
        from .kernels import extract_weights
        self.factors, self.weights = extract_weights(self.factors)


# ---------------------------------------------------------------- entry points
_SVD_KINDS = {"svd": _lib.INIT_SVD, "nnsvd": _lib.INIT_NNSVD, "leverage_score": _lib.INIT_LEVERAGE}


def initialize_factors(shape, rank, init=InitStrategy.Random, seed=None, tensor=None):
    """initialize_factors (cp/helpers.rs:192-377).

    Random / RandomNormal are drawn here (seeded; the reference uses an unseeded
    thread_rng).  Svd / Nnsvd / LeverageScore need the tensor: the SVD of every
    unfolding runs on the device (tenrso_cp_init_svd; mode sizes <= 1024, rank
    <= 64).  Returns float32 numpy arrays.
    """
    rng = np.random.default_rng(seed)
    if init == InitStrategy.Random:
        return [rng.random((n, rank)).astype(np.float32) for n in shape]
    if init == InitStrategy.RandomNormal:
        return [rng.standard_normal((n, rank)).astype(np.float32) for n in shape]
    if init.value in _SVD_KINDS:
        if tensor is None:
            raise ValueError(f"{init} initialisation needs the tensor")
        import torch
        X = tensor if type(tensor).__module__.startswith("torch") else torch.from_numpy(
            np.ascontiguousarray(tensor, dtype=np.float32))
        if not X.is_cuda:
            X = X.to("cuda")
        X = X.float().contiguous()
        if tuple(X.shape) != tuple(shape):
            raise ShapeMismatch(f"tensor shape {tuple(X.shape)} does not match {tuple(shape)}")
        F = [torch.empty((int(n), int(rank)), dtype=torch.float32, device=X.device) for n in shape]
        ctx = _lib.torch_context(X)
        shp = (C.c_int64 * len(shape))(*[int(n) for n in shape])
        ptrs = (C.c_void_p * len(F))(*[f.data_ptr() for f in F])
        seed_u = int(rng.integers(0, 2**63)) if seed is None else int(seed)
        rc = lib().tenrso_cp_init_svd(ctx.ptr, X.data_ptr(), len(shape), shp, int(rank), _SVD_KINDS[init.value],
                                      seed_u, ptrs)
        if rc != _lib.OK:
            msg = _lib.last_error()
            mk = _ERRORS.get(rc)
            if mk:
                raise mk(msg)
            raise _lib.TenrsoError(rc, msg)
        return [f.cpu().numpy() for f in F]
    raise ValueError(f"unknown initialisation {init}")


def _validate(shape, rank, tol):
    """cp/core.rs:82-95 (same order)."""
    if rank == 0:
        raise InvalidRank(rank)
    for s in shape:
        if rank > s:
            raise InvalidRank(rank)
    if not (0.0 <= tol < 1.0):
        raise InvalidTolerance(tol)


_ERRORS = {
    _lib.ERR_INVALID_RANK: lambda m: InvalidRank(m.split(": ", 1)[-1]),
    _lib.ERR_INVALID_TOLERANCE: lambda m: InvalidTolerance(m.split(": ", 1)[-1]),
    _lib.ERR_SHAPE_MISMATCH: lambda m: ShapeMismatch(m.split(": ", 1)[-1]),
    _lib.ERR_LINALG: lambda m: LinalgError(m.split(": ", 1)[-1]),
    _lib.ERR_NONNEGATIVE: lambda m: NonnegativeViolation(),
    _lib.ERR_INVALID_REGULARIZATION: lambda m: InvalidRegularization(m.split(": ", 1)[-1]),
}


def cp_als_with_init(tensor, init_factors, max_iters, tol, time_limit=None, constraints=None, precision="fp32",
                     fit_from_mode0=False):
    """CP-ALS from explicit initial factors (cp/core.rs:59-182 / 225-408 semantics).

    tensor: numpy float32 array (host path) or float32 CUDA tensor (device path).
    precision: "fp32" (drop-in) or "bf16" (BF16-input / FP32-accumulate MTTKRP).
    time_limit: seconds or datetime.timedelta or None.
    """
    shape = tuple(int(s) for s in tensor.shape)
    if len(init_factors) != len(shape):
        raise ShapeMismatch(f"Number of factor matrices ({len(init_factors)}) must match tensor rank ({len(shape)})")
    rank = int(init_factors[0].shape[1])
    _validate(shape, rank, tol)
    for k, f in enumerate(init_factors):
        if tuple(int(d) for d in f.shape) != (shape[k], rank):
            raise ShapeMismatch(f"Initial factor {k} has shape {tuple(f.shape)}, expected ({shape[k]}, {rank})")
    cons = constraints or CpConstraints()
    cons.regularization.validate()
    if cons.regularization.kind == "None" and cons.l2_reg < 0.0:
        raise InvalidTolerance(cons.l2_reg)  # cp/core.rs:255-258
    if hasattr(time_limit, "total_seconds"):
        time_limit = time_limit.total_seconds()
    reg = cons.regularization
    opts = CpOptions()
    opts.rank = rank
    opts.max_iters = int(max_iters)
    opts.precision = {"fp32": _lib.PREC_FP32, "bf16": _lib.PREC_BF16}[precision]
    opts.tol = float(tol)
    opts.time_limit_s = -1.0 if time_limit is None else float(time_limit)
    opts.nonnegative = int(cons.nonnegative)
    opts.validate_nonneg_input = int(cons.validate_nonneg_input)
    opts.orthogonal = int(cons.orthogonal)
    opts.tikhonov_order = int(reg.order) if (reg.kind == "Tikhonov" and reg.order > 0) else 0
    opts.tikhonov_lambda = float(reg.lam) if opts.tikhonov_order else 0.0
    opts.l2 = float(cons.effective_l2()) if not opts.tikhonov_order else 0.0
    opts.l1 = float(cons.effective_l1())
    opts.fit_from_mode0 = int(bool(fit_from_mode0))
    res = CpResult()
    hist = np.zeros(max(int(max_iters), 1), dtype=np.float64)
    shp = (C.c_int64 * len(shape))(*shape)
    if type(tensor).__module__.startswith("torch"):
        import torch
        X = tensor.contiguous()
        if X.dtype != torch.float32 or not X.is_cuda:
            raise TypeError("device cp_als expects a float32 CUDA tensor")
        # the result factors: one copy of all initial factors (one launch), each
        # factor a contiguous view of it
        src = [torch.as_tensor(f, device=X.device).float().reshape(-1) for f in init_factors]
        flat = torch.cat(src)
        F, off = [], 0
        for f in init_factors:
            n = int(f.shape[0]) * rank
            F.append(flat[off:off + n].view(int(f.shape[0]), rank))
            off += n
        ctx = _lib.torch_context(X)
        ptrs = (C.c_void_p * len(F))(*[f.data_ptr() for f in F])
        rc = lib().tenrso_cp_als_f32(ctx.ptr, X.data_ptr(), len(shape), shp, ptrs, C.byref(opts), C.byref(res),
                                     hist.ctypes.data)
    else:
        X = np.ascontiguousarray(tensor, dtype=np.float32)
        F = [np.array(f, dtype=np.float32, order="C", copy=True) for f in init_factors]
        ctx = _lib.context_for(0)
        ptrs = (C.c_void_p * len(F))(*[f.ctypes.data for f in F])
        rc = lib().tenrso_cp_als_f32_host(ctx.ptr, X.ctypes.data, len(shape), shp, ptrs, C.byref(opts),
                                          C.byref(res), hist.ctypes.data)
    if rc != _lib.OK:
        msg = _lib.last_error()
        mk = _ERRORS.get(rc)
        if mk:
            raise mk(msg)
        raise _lib.TenrsoError(rc, msg)
    info = ConvergenceInfo(fit_history=hist[:res.iters].tolist(), reason=ConvergenceReason(res.reason),
                           oscillated=bool(res.oscillated), oscillation_count=int(res.oscillation_count),
                           final_fit_change=float(res.final_fit_change))
    return CpDecomp(factors=F, weights=None, fit=float(res.fit), iters=int(res.iters), convergence=info,
                    solve_path_max=int(res.solve_path_max), seconds=float(res.seconds))


def cp_als(tensor, rank, max_iters, tol, init=InitStrategy.Random, time_limit=None, seed=None, precision="fp32"):
    """tenrso_decomp::cp_als (cp/core.rs:59-182)."""
    shape = tuple(int(s) for s in tensor.shape)
    _validate(shape, int(rank), tol)
    F = initialize_factors(shape, int(rank), init, seed, tensor=tensor)
    return cp_als_with_init(tensor, F, max_iters, tol, time_limit=time_limit, precision=precision)


def cp_als_constrained(tensor, rank, max_iters, tol, init=InitStrategy.Random, constraints=None, time_limit=None,
                       seed=None, precision="fp32"):
    """tenrso_decomp::cp_als_constrained (cp/core.rs:225-408)."""
    shape = tuple(int(s) for s in tensor.shape)
    _validate(shape, int(rank), tol)
    F = initialize_factors(shape, int(rank), init, seed, tensor=tensor)
    return cp_als_with_init(tensor, F, max_iters, tol, time_limit=time_limit, constraints=constraints,
                            precision=precision)
