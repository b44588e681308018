"""GPU test of the no-allocation MTTKRP entry point (tenrso_mttkrp_f32_noalloc).

Called through the C ABI with ctypes on a context created beforehand, it must
agree with tenrso_mttkrp_f32 (the entry point it mirrors) and with the FP64
oracle to the FP32 bar (relative Frobenius <= 1e-5) on two calls in a row, the
second with a smaller input (and, there, a 4-way tensor whose Khatri-Rao rows
also live in the workspace); the device's free memory must not change across
the calls; a shape whose workspace exceeds 16 MiB is refused before any launch.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import oracle  # noqa: E402
from tenrso_amd import _lib  # noqa: E402

SCRATCH_BYTES = 16 << 20


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _hip():
    L = C.CDLL("libamdhip64.so")
    L.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    L.hipMemGetInfo.restype = C.c_int
    return L


def _free_bytes(hip):
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _call(entry, ctx, Xd, shape, Fd, rank, mode, out):
    fptrs = (C.c_void_p * len(Fd))(*[f.data_ptr() for f in Fd])
    shp = (C.c_int64 * len(shape))(*shape)
    return entry(ctx.ptr, Xd.data_ptr(), len(shape), shp, fptrs, rank, mode, out.data_ptr())


def test_noalloc_mttkrp_matches_mttkrp_f32():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L = _lib.lib()
    hip = _hip()
    ctx_ref = _lib.Context(0)   # its own buffers for tenrso_mttkrp_f32
    ctx_na = _lib.Context(0)    # set up, never given a buffer
    cases = [((256, 256, 256), 64), ((24, 32, 16, 32), 6)]  # second call: smaller input
    rng = np.random.default_rng(2024)
    inputs = []
    for shape, rank in cases:
        X = rng.standard_normal(shape).astype(np.float32)
        F = [rng.standard_normal((n, rank)).astype(np.float32) for n in shape]
        Xd = torch.from_numpy(X).cuda()
        Fd = [torch.from_numpy(f).cuda() for f in F]
        outs = [torch.full((n, rank), np.nan, dtype=torch.float32, device="cuda") for n in shape]
        refs = [torch.empty((n, rank), dtype=torch.float32, device="cuda") for n in shape]
        inputs.append((shape, rank, X, F, Xd, Fd, outs, refs))
    # the entry point it mirrors first: the HIP runtime's own first-launch set-up
    # then happens before the free-memory reading
    for shape, rank, X, F, Xd, Fd, outs, refs in inputs:
        for mode in range(len(shape)):
            _lib.check(_call(L.tenrso_mttkrp_f32, ctx_ref, Xd, shape, Fd, rank, mode, refs[mode]))
    _lib.check(L.tenrso_ctx_synchronize(ctx_ref.ptr))
    torch.cuda.synchronize()
    free0 = _free_bytes(hip)
    for shape, rank, X, F, Xd, Fd, outs, refs in inputs:
        print(f"[case] noalloc mttkrp {shape} rank {rank}", flush=True)
        for mode in range(len(shape)):
            rc = _call(L.tenrso_mttkrp_f32_noalloc, ctx_na, Xd, shape, Fd, rank, mode, outs[mode])
            assert rc == _lib.OK, _lib.last_error()
        _lib.check(L.tenrso_ctx_synchronize(ctx_na.ptr))
    free1 = _free_bytes(hip)
    assert free1 == free0, (free0, free1)
    for shape, rank, X, F, Xd, Fd, outs, refs in inputs:
        X64, F64 = X.astype(np.float64), [f.astype(np.float64) for f in F]
        for mode in range(len(shape)):
            got, ref = outs[mode].cpu().numpy(), refs[mode].cpu().numpy()
            exact = oracle.mttkrp_fast(X64, F64, mode)
            print(f"  mode {mode}: vs mttkrp_f32 {rel(got, ref):.2e}, vs fp64 {rel(got, exact):.2e}", flush=True)
            assert np.isfinite(got).all()
            assert rel(got, exact) <= 1e-5
            assert rel(got, ref) <= 1e-5
    # a workspace need beyond 16 MiB (70000 x 64 fp32 outputs = 17.9 MB even unsplit)
    shape, rank = (70000, 32, 2), 64
    assert shape[0] * rank * 4 > SCRATCH_BYTES
    Xd = torch.zeros(shape, dtype=torch.float32, device="cuda")
    Fd = [torch.zeros((n, rank), dtype=torch.float32, device="cuda") for n in shape]
    out = torch.zeros((shape[0], rank), dtype=torch.float32, device="cuda")
    rc = _call(L.tenrso_mttkrp_f32_noalloc, ctx_na, Xd, shape, Fd, rank, 0, out)
    assert rc == _lib.ERR_UNSUPPORTED
    assert "too small" in _lib.last_error()
    ctx_ref.close()
    ctx_na.close()
