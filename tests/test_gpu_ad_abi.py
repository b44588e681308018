"""GPU tests of the reverse-mode C entry points with NULL and partial outputs, and
of the AdOperation-shaped rules (tenrso_ad_op_*; tenrso-ad/src/hooks.rs:44-74,
vjp.rs:30-44), all called through the C ABI, against the FP64 oracle:

- mttkrp VJP: dX = cp_reconstruct(A with A_mode := Mbar), dA_m =
  mttkrp(X, A with A_mode := Mbar, m); pinned by FP64 finite differences in
  tests/test_gpu_ad.py and here by the oracle's cp_reconstruct and mttkrp;
- cp_reconstruct: dA_n = oracle.cp_factor_grads(G, A, w) (grad.rs:81-120);
- cp_loss: L = 0.5 ||X - Xhat||^2 from the oracle's reconstruction,
  dX = g (X - Xhat), dA_n = g oracle.cp_factor_grads(Xhat - X, A).
Bars: FP32 results within 1e-5 relative (Frobenius) of FP64; buffers whose
gradient is not requested are left untouched (NaN sentinels stay NaN).
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import oracle  # noqa: E402
from tenrso_amd import _lib  # noqa: E402
from tenrso_amd.ad import CpLoss, DeviceAdOperation  # noqa: E402


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def nan_like(shape):
    return torch.full(shape, float("nan"), dtype=torch.float32, device="cuda")


def ptrs(ts):
    return (C.c_void_p * len(ts))(*[0 if t is None else t.data_ptr() for t in ts])


def i64s(v):
    return (C.c_int64 * len(v))(*v)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _problem(shape, R, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal(shape).astype(np.float32)
    F = [rng.standard_normal((n, R)).astype(np.float32) for n in shape]
    return rng, X, F


def test_mttkrp_vjp_entry_null_and_partial_outputs():
    print("[case] tenrso_mttkrp_vjp_f32 NULL / partial outputs", flush=True)
    shape, R, mode = (20, 24, 32), 5, 1
    rng, X, F = _problem(shape, R, 1)
    Mbar = rng.standard_normal((shape[mode], R)).astype(np.float32)
    Xd, Fd, Md = dev(X), [dev(f) for f in F], dev(Mbar)
    F64 = [f.astype(np.float64) for f in F]
    Fm = list(F64)
    Fm[mode] = Mbar.astype(np.float64)
    ref_dX = oracle.cp_reconstruct(Fm)
    ref_dA = {m: oracle.mttkrp_fast(X.astype(np.float64), Fm, m) for m in range(3) if m != mode}
    L = _lib.lib()
    ctx = _lib.torch_context(Xd)
    # dX only (dfactors NULL)
    dX = nan_like(shape)
    _lib.check(L.tenrso_mttkrp_vjp_f32(ctx.ptr, Xd.data_ptr(), 3, i64s(shape), ptrs(Fd), R, mode, Md.data_ptr(),
                                       None, dX.data_ptr()))
    assert rel(dX.cpu().numpy(), ref_dX) <= 1e-5
    # dA_2 only (dX NULL, dA_0 NULL); the dA_mode slot is never written
    d0, d1, d2 = nan_like((shape[0], R)), nan_like((shape[1], R)), nan_like((shape[2], R))
    _lib.check(L.tenrso_mttkrp_vjp_f32(ctx.ptr, Xd.data_ptr(), 3, i64s(shape), ptrs(Fd), R, mode, Md.data_ptr(),
                                       ptrs([None, d1, d2]), None))
    torch.cuda.synchronize()
    assert torch.isnan(d1).all() and torch.isnan(d0).all()
    assert rel(d2.cpu().numpy(), ref_dA[2]) <= 1e-5
    # everything
    _lib.check(L.tenrso_mttkrp_vjp_f32(ctx.ptr, Xd.data_ptr(), 3, i64s(shape), ptrs(Fd), R, mode, Md.data_ptr(),
                                       ptrs([d0, None, d2]), dX.data_ptr()))
    assert rel(d0.cpu().numpy(), ref_dA[0]) <= 1e-5 and rel(dX.cpu().numpy(), ref_dX) <= 1e-5


def test_factor_grads_loss_grad_gd_entries_null_and_partial():
    print("[case] factor grads / loss grad / gd NULL and partial outputs", flush=True)
    shape, R = (16, 20, 24), 4
    rng, X, F = _problem(shape, R, 2)
    w = rng.uniform(0.5, 2.0, R).astype(np.float32)
    G = rng.standard_normal(shape).astype(np.float32)
    Xd, Fd, Gd, wd = dev(X), [dev(f) for f in F], dev(G), dev(w)
    F64 = [f.astype(np.float64) for f in F]
    L = _lib.lib()
    ctx = _lib.torch_context(Xd)
    # cp_factor_grads: only mode 1 requested
    ref = oracle.cp_factor_grads(G.astype(np.float64), F64, w.astype(np.float64))
    g = [nan_like((n, R)) for n in shape]
    _lib.check(L.tenrso_cp_factor_grads_f32(ctx.ptr, Gd.data_ptr(), 3, i64s(shape), ptrs(Fd), R, wd.data_ptr(),
                                            ptrs([None, g[1], None])))
    torch.cuda.synchronize()
    assert torch.isnan(g[0]).all() and torch.isnan(g[2]).all()
    assert rel(g[1].cpu().numpy(), ref[1]) <= 1e-5
    # cp_loss_grad: loss only (grads NULL), then one gradient
    Xh = oracle.cp_reconstruct(F64)
    loss_ref = 0.5 * float(np.sum((X.astype(np.float64) - Xh) ** 2))
    gref = oracle.cp_factor_grads(Xh - X.astype(np.float64), F64)
    loss = C.c_double()
    _lib.check(L.tenrso_cp_loss_grad_f32(ctx.ptr, Xd.data_ptr(), 3, i64s(shape), ptrs(Fd), R, None, C.byref(loss)))
    assert abs(loss.value - loss_ref) <= 1e-6 * loss_ref
    _lib.check(L.tenrso_cp_loss_grad_f32(ctx.ptr, Xd.data_ptr(), 3, i64s(shape), ptrs(Fd), R,
                                         ptrs([g[0], None, None]), None))
    torch.cuda.synchronize()
    assert rel(g[0].cpu().numpy(), gref[0]) <= 1e-5 and torch.isnan(g[2]).all()
    # cp_gd: the loss history is optional; the factors take the same steps either way
    Fa = [f.clone() for f in Fd]
    Fb = [f.clone() for f in Fd]
    hist = np.zeros(5)
    _lib.check(L.tenrso_cp_gd_f32(ctx.ptr, Xd.data_ptr(), 3, i64s(shape), ptrs(Fa), R, 5, 0.0,
                                  hist.ctypes.data))
    _lib.check(L.tenrso_cp_gd_f32(ctx.ptr, Xd.data_ptr(), 3, i64s(shape), ptrs(Fb), R, 5, 0.0, None))
    torch.cuda.synchronize()
    assert all(torch.equal(a, b) for a, b in zip(Fa, Fb))
    assert abs(hist[0] - loss_ref) <= 1e-6 * loss_ref and hist[-1] < hist[0]


def test_ad_op_mttkrp_rule():
    print("[case] tenrso_ad_op mttkrp", flush=True)
    shape, R, mode = (12, 16, 10, 32), 6, 2
    rng, X, F = _problem(shape, R, 3)
    Mbar = rng.standard_normal((shape[mode], R)).astype(np.float32)
    op = DeviceAdOperation(_lib.AD_MTTKRP, dev(X), [dev(f) for f in F], mode=mode)
    assert op.name() == "mttkrp" and op.num_inputs() == 5 and op.num_outputs() == 1
    assert [_lib.lib().tenrso_ad_op_input_numel(op._ptr, k) for k in range(5)] == \
        [int(np.prod(shape))] + [n * R for n in shape]
    assert _lib.lib().tenrso_ad_op_output_numel(op._ptr, 0) == shape[mode] * R
    assert _lib.lib().tenrso_ad_op_input_numel(op._ptr, 5) == -1
    X64, F64 = X.astype(np.float64), [f.astype(np.float64) for f in F]
    (M,) = op.forward()
    assert rel(M.cpu().numpy(), oracle.mttkrp_fast(X64, F64, mode)) <= 1e-5
    Fm = list(F64)
    Fm[mode] = Mbar.astype(np.float64)
    grads = op.vjp(dev(Mbar))
    assert rel(grads[0].cpu().numpy(), oracle.cp_reconstruct(Fm)) <= 1e-5
    for m in range(4):
        g = grads[1 + m].cpu().numpy()
        if m == mode:
            assert np.array_equal(g, np.zeros_like(g))
        else:
            assert rel(g, oracle.mttkrp_fast(X64, Fm, m)) <= 1e-5
    # partial: only dA_0 and the (zero) dA_mode
    part = op.backward([dev(Mbar)], needs=[False, True, False, True, False])
    assert part[0] is None and part[2] is None and part[4] is None
    assert torch.equal(part[1], grads[1]) and torch.equal(part[3], grads[3])
    # a NULL gradient array computes nothing and succeeds
    _lib.check(_lib.lib().tenrso_ad_op_backward(op._ptr, ptrs([dev(Mbar)]), None))


def test_ad_op_cp_reconstruct_rule():
    print("[case] tenrso_ad_op cp_reconstruct", flush=True)
    shape, R = (14, 18, 22), 7
    rng, _, F = _problem(shape, R, 4)
    w = rng.uniform(0.5, 2.0, R).astype(np.float32)
    G = rng.standard_normal(shape).astype(np.float32)
    op = DeviceAdOperation(_lib.AD_CP_RECONSTRUCT, None, [dev(f) for f in F], weights=dev(w))
    assert op.name() == "cp_reconstruct" and op.num_inputs() == 3 and op.num_outputs() == 1
    F64, w64 = [f.astype(np.float64) for f in F], w.astype(np.float64)
    (Xh,) = op.forward()
    assert rel(Xh.cpu().numpy(), oracle.cp_reconstruct(F64, w64)) <= 1e-5
    ref = oracle.cp_factor_grads(G.astype(np.float64), F64, w64)
    grads = op.vjp(dev(G))
    for n in range(3):
        assert rel(grads[n].cpu().numpy(), ref[n]) <= 1e-5
    part = op.backward([dev(G)], needs=[False, False, True])
    assert part[0] is None and part[1] is None and torch.equal(part[2], grads[2])


def test_ad_op_cp_loss_rule():
    print("[case] tenrso_ad_op cp_loss", flush=True)
    shape, R = (20, 16, 24), 5
    rng, X, F = _problem(shape, R, 5)
    F = [0.5 * f for f in F]
    op = CpLoss(dev(X), [dev(f) for f in F])
    assert op.name() == "cp_loss" and op.num_inputs() == 4 and op.num_outputs() == 1
    X64, F64 = X.astype(np.float64), [f.astype(np.float64) for f in F]
    Xh = oracle.cp_reconstruct(F64)
    loss_ref = 0.5 * float(np.sum((X64 - Xh) ** 2))
    (Ld,) = op.forward()
    assert Ld.shape == (1,) and abs(float(Ld.item()) - loss_ref) <= 1e-6 * loss_ref
    g = 2.5
    grads = op.backward([torch.tensor([g], device="cuda")])
    assert rel(grads[0].cpu().numpy(), g * (X64 - Xh)) <= 1e-5
    fref = oracle.cp_factor_grads(Xh - X64, F64)
    for n in range(3):
        assert rel(grads[1 + n].cpu().numpy(), g * fref[n]) <= 1e-5
    part = op.backward([g], needs=[True, False, True, False])
    assert part[1] is None and part[3] is None
    assert torch.equal(part[0], grads[0]) and torch.equal(part[2], grads[2])


def test_ad_op_create_errors():
    L = _lib.lib()
    Xd = torch.zeros((4, 5, 6), dtype=torch.float32, device="cuda")
    Fd = [torch.zeros((n, 3), dtype=torch.float32, device="cuda") for n in (4, 5, 6)]
    ctx = _lib.torch_context(Xd)
    out = C.c_void_p()
    rc = L.tenrso_ad_op_create(ctx.ptr, 7, Xd.data_ptr(), 3, i64s((4, 5, 6)), ptrs(Fd), 3, 0, None, C.byref(out))
    assert rc == _lib.ERR_INVALID_ARGUMENT and "unknown AD operation kind" in _lib.last_error()
    rc = L.tenrso_ad_op_create(ctx.ptr, _lib.AD_CP_LOSS, None, 3, i64s((4, 5, 6)), ptrs(Fd), 3, 0, None,
                               C.byref(out))
    assert rc == _lib.ERR_INVALID_ARGUMENT and "needs the tensor X" in _lib.last_error()
    rc = L.tenrso_ad_op_create(ctx.ptr, _lib.AD_MTTKRP, Xd.data_ptr(), 3, i64s((4, 5, 6)), ptrs(Fd), 3, 3, None,
                               C.byref(out))
    assert rc == _lib.ERR_INVALID_ARGUMENT and "Mode 3 out of bounds for tensor with rank 3" in _lib.last_error()
    assert not out.value
