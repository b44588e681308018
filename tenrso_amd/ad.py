"""tenrso-ad mirror for the CP path: VJP rules on gfx950.

* ``CpReconstructionGrad`` (tenrso-ad/src/grad.rs:41-120): factor gradients of
  X^ = sum_r w_r a_1^r o ... o a_N^r, i.e. dA_n = mttkrp(dL/dX^, A, n) diag(w).
* ``MttkrpVjp``: VjpOp-shaped rule (tenrso-ad/src/vjp.rs:30-44) for
  M = mttkrp(X, A, mode); ``vjp(Mbar)`` returns [dX] + [dA_m for m != mode].
* ``DeviceAdOperation``: AdOperation-shaped rules (tenrso-ad/src/hooks.rs:44-74)
  held by the library (``tenrso_ad_op_*``): mttkrp, cp_reconstruct, cp_loss.
* ``CpLoss``: the cp_loss rule, L = 0.5 ||X - [[A]]||^2, gradients
  X - Xhat and A_n Gamma_n - mttkrp(X, A, n).
* ``cp_gd``: gradient-descent CP fit on the device.
* ``check_gradient`` / ``GradCheckConfig`` (tenrso-ad/src/gradcheck.rs:43-268):
  the reference's element-wise central finite-difference checker (host logic,
  callables of numpy arrays).

Every compute call takes float32 CUDA tensors and runs on torch's current
stream through the HIP library; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, lib
from .kernels import _check_mttkrp, _i64s, mttkrp

__all__ = ["CpReconstructionGrad", "MttkrpVjp", "CpLoss", "DeviceAdOperation", "cp_loss_and_grad", "cp_gd",
           "GradCheckConfig", "GradCheckResult", "check_gradient", "AdContext"]


def _dev(t, what):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32):
        raise TypeError(f"{what} must be a float32 CUDA tensor")
    return t.contiguous()


def _ptrs(ts):
    return (C.c_void_p * len(ts))(*[0 if t is None else t.data_ptr() for t in ts])


class CpReconstructionGrad:
    """grad.rs:41-120.  ``factors`` are (I_n, R) CUDA tensors, ``weights`` (R,) or None."""

    def __init__(self, factors, weights=None):
        self.factors = [_dev(f, "factor") for f in factors]
        self.weights = None if weights is None else _dev(weights, "weights")

    def compute_factor_gradients(self, grad_output):
        """grad.rs:81-120: [mttkrp(grad_output, factors, n) * weights for each mode n]."""
        import torch
        n_modes = len(self.factors)
        if grad_output.dim() != n_modes:
            raise ValueError(f"Gradient rank {grad_output.dim()} doesn't match number of factors {n_modes}")
        G = _dev(grad_output, "grad_output")
        shape = tuple(int(s) for s in G.shape)
        R = _check_mttkrp(shape, self.factors, 0)
        grads = [torch.empty((shape[n], R), dtype=torch.float32, device=G.device) for n in range(n_modes)]
        if R == 0:
            return grads
        ctx = _lib.torch_context(G)
        check(lib().tenrso_cp_factor_grads_f32(ctx.ptr, G.data_ptr(), n_modes, _i64s(shape), _ptrs(self.factors), R,
                                               None if self.weights is None else self.weights.data_ptr(),
                                               _ptrs(grads)))
        return grads


class MttkrpVjp:
    """VJP of M = mttkrp(X, factors, mode) (vjp.rs:30-44 shape: ``vjp(output_grad) -> [grads]``).

    Inputs, in order: X, then factors[m] for m != mode (factors[mode] does not
    enter the MTTKRP).  ``vjp`` returns the gradients in that order.
    """

    def __init__(self, tensor, factors, mode: int):
        self.X = _dev(tensor, "tensor")
        self.factors = [_dev(f, "factor") for f in factors]
        self.mode = int(mode)
        self.shape = tuple(int(s) for s in self.X.shape)
        self.R = _check_mttkrp(self.shape, self.factors, self.mode)

    def forward(self):
        return mttkrp(self.X, self.factors, self.mode)

    def vjp(self, output_grad):
        import torch
        Mbar = _dev(output_grad, "output_grad")
        nd, mode, R = len(self.shape), self.mode, self.R
        if tuple(Mbar.shape) != (self.shape[mode], R):
            raise ValueError(f"output_grad shape {tuple(Mbar.shape)} doesn't match {(self.shape[mode], R)}")
        dX = torch.empty(self.shape, dtype=torch.float32, device=self.X.device)
        dF = [None if m == mode else torch.empty((self.shape[m], R), dtype=torch.float32, device=self.X.device)
              for m in range(nd)]
        if R == 0:
            dX.zero_()
            return [dX] + [d.zero_() for d in dF if d is not None]
        ctx = _lib.torch_context(self.X)
        check(lib().tenrso_mttkrp_vjp_f32(ctx.ptr, self.X.data_ptr(), nd, _i64s(self.shape), _ptrs(self.factors), R,
                                          mode, Mbar.data_ptr(), _ptrs(dF), dX.data_ptr()))
        return [dX] + [d for d in dF if d is not None]

    # AdOperation shape (hooks.rs:44-74)
    def backward(self, output_grads):
        return self.vjp(output_grads[0])

    def name(self):
        return "mttkrp"

    def num_inputs(self):
        return len(self.shape)

    def num_outputs(self):
        return 1


def cp_loss_and_grad(tensor, factors, with_grad: bool = True):
    """L = 0.5 ||X - [[A_1..A_N]]||^2 and dL/dA_n = A_n Gamma_n - mttkrp(X, A, n).

    Gram matrices, Gamma_n, the A_n Gamma_n product and the loss are FP64.
    Returns (loss: float, grads: list of (I_n, R) tensors or None).
    """
    import torch
    X = _dev(tensor, "tensor")
    F = [_dev(f, "factor") for f in factors]
    shape = tuple(int(s) for s in X.shape)
    R = _check_mttkrp(shape, F, 0)
    grads = [torch.empty((shape[n], R), dtype=torch.float32, device=X.device) for n in range(len(shape))] \
        if with_grad else None
    loss = C.c_double(0.0)
    ctx = _lib.torch_context(X)
    check(lib().tenrso_cp_loss_grad_f32(ctx.ptr, X.data_ptr(), len(shape), _i64s(shape), _ptrs(F), R,
                                        _ptrs(grads) if grads is not None else None, C.byref(loss)))
    return loss.value, grads


class DeviceAdOperation:
    """AdOperation-shaped rule (hooks.rs:44-74) held by the HIP library
    (``tenrso_ad_op_*``): ``forward()``, ``backward(output_grads, needs=None)``,
    ``vjp(output_grad)`` (the VjpOp shape, vjp.rs:30-44), ``name()``,
    ``num_inputs()``, ``num_outputs()``.  The inputs are borrowed by the library,
    so this object keeps them alive.  ``needs`` (one bool per input) limits which
    gradients are computed; the others come back as None."""

    def __init__(self, kind, tensor, factors, mode=0, weights=None):
        import torch
        self.factors = [_dev(f, "factor") for f in factors]
        self.X = None if tensor is None else _dev(tensor, "tensor")
        self.weights = None if weights is None else _dev(weights, "weights")
        if self.X is not None:
            self.shape = tuple(int(s) for s in self.X.shape)
            self.device = self.X.device
        else:
            self.shape = tuple(int(f.shape[0]) for f in self.factors)
            self.device = self.factors[0].device
        self.R = _check_mttkrp(self.shape, self.factors, int(mode) if kind == _lib.AD_MTTKRP else 0)
        self._ctx = _lib.torch_context(self.factors[0])
        self._ptr = C.c_void_p()
        check(lib().tenrso_ad_op_create(self._ctx.ptr, int(kind), None if self.X is None else self.X.data_ptr(),
                                        len(self.shape), _i64s(self.shape), _ptrs(self.factors), self.R, int(mode),
                                        None if self.weights is None else self.weights.data_ptr(),
                                        C.byref(self._ptr)))
        self.kind = int(kind)
        self.mode = int(mode)
        self._torch = torch

    def _input_shape(self, k):
        has_x = self.kind != _lib.AD_CP_RECONSTRUCT
        if has_x and k == 0:
            return self.shape
        m = k - 1 if has_x else k
        return (self.shape[m], self.R)

    def _output_shape(self):
        if self.kind == _lib.AD_MTTKRP:
            return (self.shape[self.mode], self.R)
        if self.kind == _lib.AD_CP_RECONSTRUCT:
            return self.shape
        return (1,)

    def name(self):
        return lib().tenrso_ad_op_name(self._ptr).decode()

    def num_inputs(self):
        return int(lib().tenrso_ad_op_num_inputs(self._ptr))

    def num_outputs(self):
        return int(lib().tenrso_ad_op_num_outputs(self._ptr))

    def forward(self):
        out = self._torch.empty(self._output_shape(), dtype=self._torch.float32, device=self.device)
        check(lib().tenrso_ad_op_forward(self._ptr, _ptrs([out])))
        return [out]

    def backward(self, output_grads, needs=None):
        g = output_grads[0]
        if not (isinstance(g, self._torch.Tensor) and g.is_cuda):
            g = self._torch.tensor(g, dtype=self._torch.float32, device=self.device)
        g = _dev(g.reshape(self._output_shape()).to(self._torch.float32), "output gradient")
        n_in = self.num_inputs()
        needs = [True] * n_in if needs is None else list(needs)
        grads = [self._torch.empty(self._input_shape(k), dtype=self._torch.float32, device=self.device)
                 if needs[k] else None for k in range(n_in)]
        check(lib().tenrso_ad_op_backward(self._ptr, _ptrs([g]), _ptrs(grads)))
        return grads

    def vjp(self, output_grad):
        return self.backward([output_grad])

    def __del__(self):
        try:
            if self._ptr:
                lib().tenrso_ad_op_destroy(self._ptr)
                self._ptr = C.c_void_p()
        except Exception:
            pass


class CpLoss(DeviceAdOperation):
    """AdOperation-shaped (hooks.rs:44-74) rule for L = 0.5 ||X - [[A_1..A_N]]||^2:
    inputs [X, A_1..A_N], one output (the loss); backward(g) returns
    [g (X - Xhat), g (A_n Gamma_n - mttkrp(X, A, n)) for each n], on the device."""

    def __init__(self, tensor, factors):
        super().__init__(_lib.AD_CP_LOSS, tensor, factors)


class AdContext:
    """Minimal tape (hooks.rs:80-224 GenericAdAdapter): register AdOperation-shaped
    rules, run backward from one of them, read the input gradients back."""

    def __init__(self):
        self._ops = []
        self._grads = {}

    def register_operation(self, op) -> int:
        self._ops.append(op)
        return len(self._ops) - 1

    def backward(self, op_id: int, grad):
        if not 0 <= op_id < len(self._ops):
            raise KeyError(f"Operation {op_id} not found")
        for k, g in enumerate(self._ops[op_id].backward([grad])):
            self._grads[(op_id, k)] = g

    def get_gradient(self, op_id: int, input_index: int):
        return self._grads.get((op_id, input_index))

    def clear(self):
        self._ops.clear()
        self._grads.clear()


def cp_gd(tensor, factors, steps: int, lr: float = 0.0):
    """``steps`` simultaneous gradient steps A_n -= lr dL/dA_n on the device, in place
    on ``factors``.  lr <= 0 picks 1 / max_n ||Gamma_n||_F at every step.  Returns
    the loss history (loss at the point each step starts from)."""
    X = _dev(tensor, "tensor")
    for f in factors:
        _dev(f, "factor")
        if not f.is_contiguous():
            raise ValueError("cp_gd updates factors in place; they must be contiguous")
    shape = tuple(int(s) for s in X.shape)
    R = _check_mttkrp(shape, factors, 0)
    hist = np.zeros(max(int(steps), 0), dtype=np.float64)
    ctx = _lib.torch_context(X)
    check(lib().tenrso_cp_gd_f32(ctx.ptr, X.data_ptr(), len(shape), _i64s(shape), _ptrs(factors), R, int(steps),
                                 float(lr), hist.ctypes.data if hist.size else None))
    return hist


# ---------------------------------------------------------------- gradcheck mirror
@dataclass
class GradCheckConfig:
    """gradcheck.rs:43-70 (same defaults)."""
    epsilon: float = 1e-5
    rtol: float = 1e-3
    atol: float = 1e-5
    use_central_diff: bool = True
    verbose: bool = False


@dataclass
class GradCheckResult:
    passed: bool
    max_abs_diff: float
    max_rel_diff: float
    num_failures: int
    num_elements: int


def check_gradient(f, df, x, grad_y, config: GradCheckConfig | None = None) -> GradCheckResult:
    """gradcheck.rs:107-268: compare df(x, grad_y) with the element-wise finite
    difference <grad_y, (f(x+h e_i) - f(x-h e_i)) / 2h>.  An element fails when
    abs_diff > atol and rel_diff > rtol, rel_diff = abs_diff / |numerical| when
    |numerical| > machine epsilon of x's dtype, else abs_diff."""
    config = config or GradCheckConfig()
    x = np.asarray(x)
    grad_y = np.asarray(grad_y)
    analytical = np.asarray(df(x, grad_y))
    if analytical.shape != x.shape:
        raise ValueError(f"Gradient shape {list(analytical.shape)} doesn't match input shape {list(x.shape)}")
    eps = x.dtype.type(config.epsilon)
    numerical = np.zeros_like(x)
    flat = x.reshape(-1)
    for i in range(flat.size):
        xp = flat.copy()
        xp[i] = xp[i] + eps
        yp = np.asarray(f(xp.reshape(x.shape)))
        if config.use_central_diff:
            xm = flat.copy()
            xm[i] = xm[i] - eps
            ym = np.asarray(f(xm.reshape(x.shape)))
            numerical.reshape(-1)[i] = np.sum(grad_y * ((yp - ym) * (x.dtype.type(1) / (eps + eps))))
        else:
            y0 = np.asarray(f(x))
            numerical.reshape(-1)[i] = np.sum(grad_y * ((yp - y0) * (x.dtype.type(1) / eps)))
    mach = np.finfo(x.dtype).eps
    abs_diff = np.abs(analytical - numerical)
    rel_diff = np.where(np.abs(numerical) > mach, abs_diff / np.where(np.abs(numerical) > mach,
                                                                       np.abs(numerical), 1), abs_diff)
    fails = (abs_diff > config.atol) & (rel_diff > config.rtol)
    if config.verbose:
        for idx in zip(*np.nonzero(fails)):
            print(f"Gradient mismatch at {list(idx)}: analytical={analytical[idx]}, numerical={numerical[idx]}, "
                  f"abs_diff={abs_diff[idx]}, rel_diff={rel_diff[idx]}")
    return GradCheckResult(passed=not fails.any(), max_abs_diff=float(abs_diff.max(initial=0.0)),
                           max_rel_diff=float(rel_diff.max(initial=0.0)), num_failures=int(fails.sum()),
                           num_elements=int(x.size))
