"""GPU tests of the n-mode products (TTM) and Tucker's TTM chain
(tenrso_nmode_product_f32 / tenrso_nmode_products_seq_f32; reference
tenrso-kernels/src/nmode.rs:72-130, 241-255, 442-612 and the HOOI chain of
tenrso-decomp/src/tucker.rs:1122-1140).

The reference's own cases (nmode.rs tests: the mode-0 values, identity, the
shape cases, the error messages) run through the device path exactly; random
tensors of 2 to 5 modes against the FP64 oracle at every mode (relative
Frobenius <= 1e-5: FP32 GEMM with contractions of at most a few hundred
terms); the Tucker core chain G = X x_0 U_0^T x_1 U_1^T x_2 U_2^T and the
reconstruction X' = G x_0 U_0 x_1 U_1 x_2 U_2 against the oracle, plus the
projection identity: for X = G0 x U (orthonormal U) the chain returns G0."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import oracle  # noqa: E402
from tenrso_amd import _lib, kernels  # noqa: E402


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def test_reference_cases():
    print("[case] nmode reference cases", flush=True)
    T = np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]], dtype=np.float32)
    M = np.array([[1.0, 0.0], [0.0, 1.0], [1.0, 1.0]], dtype=np.float32)
    assert np.array_equal(kernels.nmode_product(T, M, 0), [[1, 2, 3], [4, 5, 6], [5, 7, 9]])
    X = np.arange(24, dtype=np.float32).reshape(2, 3, 4)
    assert np.array_equal(kernels.nmode_product(X, np.eye(3, dtype=np.float32), 1), X)
    M5 = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [0, 1, 1]], np.float32)
    Y = kernels.nmode_product(X, M5, 1)
    assert Y.shape == (2, 5, 4)
    assert np.array_equal(Y, oracle.nmode_product(X, M5, 1))
    m1, m2 = np.eye(2, dtype=np.float32), np.array([[1.0, 0, 0], [0, 1.0, 0]], np.float32)
    Y = kernels.nmode_products_seq(X, [(m1, 0), (m2, 1)])
    assert Y.shape == (2, 2, 4)
    assert np.array_equal(Y, X[:, :2, :])
    # tucker_operator: single mode keeps the shape; several modes in ratio order
    assert kernels.tucker_operator(T, {0: np.eye(2, dtype=np.float32)}).shape == (2, 3)
    Y = kernels.tucker_operator(X, {0: m1, 1: m2})
    assert Y.shape == (2, 2, 4) and np.array_equal(Y, X[:, :2, :])
    with pytest.raises(ValueError, match="Mode 5 out of bounds for tensor with rank 2"):
        kernels.nmode_product(T, np.ones((1, 2), np.float32), 5)
    with pytest.raises(ValueError, match=r"Matrix columns \(2\) must match tensor mode-1 size \(3\)"):
        kernels.nmode_product(T, np.ones((1, 2), np.float32), 1)
    with pytest.raises(ValueError, match=r"Number of factor matrices \(1\) must match core rank \(2\)"):
        kernels.tucker_reconstruct(T, [np.eye(2, dtype=np.float32)])


def test_c_abi_checks_and_messages():
    """The C entry points refuse what nmode.rs:82-98 refuses, with its messages,
    before any device work (also for a product later in a chain)."""
    print("[case] nmode C-ABI checks", flush=True)
    ctx = _lib.context_for(0)
    X = torch.zeros((2, 3), device="cuda")
    M = torch.zeros((4, 2), device="cuda")
    Y = torch.zeros(16, device="cuda")
    shape = (C.c_int64 * 2)(2, 3)
    rc = _lib.lib().tenrso_nmode_product_f32(ctx.ptr, X.data_ptr(), 2, shape, M.data_ptr(), 4, 2, 2, Y.data_ptr())
    assert rc == _lib.ERR_INVALID_ARGUMENT and "Mode 2 out of bounds for tensor with rank 2" in _lib.last_error()
    rc = _lib.lib().tenrso_nmode_product_f32(ctx.ptr, X.data_ptr(), 2, shape, M.data_ptr(), 4, 2, 1, Y.data_ptr())
    assert rc == _lib.ERR_INVALID_ARGUMENT
    assert "Matrix columns (2) must match tensor mode-1 size (3)" in _lib.last_error()
    # chain: the second product meets mode 0 resized to 4 by the first
    ptrs = (C.c_void_p * 2)(M.data_ptr(), M.data_ptr())
    out_shape = (C.c_int64 * 2)()
    rc = _lib.lib().tenrso_nmode_products_seq_f32(ctx.ptr, X.data_ptr(), 2, shape, 2, ptrs, (C.c_int64 * 2)(4, 4),
                                                  (C.c_int64 * 2)(2, 2), (C.c_int * 2)(0, 0), Y.data_ptr(),
                                                  out_shape)
    assert rc == _lib.ERR_INVALID_ARGUMENT
    assert "Matrix columns (2) must match tensor mode-0 size (4)" in _lib.last_error()


@pytest.mark.parametrize("shape,rows", [((30, 40, 50), 17), ((7, 1, 9, 5), 3), ((64, 48), 96), ((3, 4, 5, 6, 7), 2),
                                        ((128, 96, 80), 64), ((1, 300), 5)])
def test_every_mode_vs_oracle(shape, rows):
    print(f"[case] nmode {shape} -> {rows} rows", flush=True)
    rng = np.random.default_rng(sum(shape) + rows)
    X = rng.standard_normal(shape).astype(np.float32)
    Xd = torch.from_numpy(X).cuda()
    for mode in range(len(shape)):
        M = rng.standard_normal((rows, shape[mode])).astype(np.float32)
        ref = oracle.nmode_product(X, M, mode)
        got = kernels.nmode_product(Xd, torch.from_numpy(M).cuda(), mode)
        assert isinstance(got, torch.Tensor) and got.is_cuda and tuple(got.shape) == ref.shape
        e = rel(got.cpu().numpy(), ref)
        print(f"  mode {mode}: rel {e:.2e}", flush=True)
        assert e <= 1e-5
        assert rel(kernels.nmode_product(X, M, mode), ref) <= 1e-5  # host path


def test_empty_products():
    """Zero rows: an empty result; an empty contracted mode: a zero result."""
    print("[case] nmode empty", flush=True)
    X = np.ones((3, 4, 5), np.float32)
    assert kernels.nmode_product(X, np.ones((0, 4), np.float32), 1).shape == (3, 0, 5)
    Z = np.ones((3, 0, 5), np.float32)
    Y = kernels.nmode_product(Z, np.ones((6, 0), np.float32), 1)
    assert Y.shape == (3, 6, 5) and not Y.any()
    assert np.array_equal(kernels.nmode_products_seq(X, []), X)


def test_tucker_core_chain_and_reconstruct():
    """Tucker's TTM chain: G = X x_0 U0^T x_1 U1^T x_2 U2^T, then X' = G x U."""
    print("[case] tucker chain 96x80x64 -> 24x20x16", flush=True)
    rng = np.random.default_rng(5)
    dims, ranks = (96, 80, 64), (24, 20, 16)
    U = [np.linalg.qr(rng.standard_normal((n, r)))[0].astype(np.float32) for n, r in zip(dims, ranks)]
    G0 = rng.standard_normal(ranks).astype(np.float32)
    X = oracle.nmode_products_seq(G0, [(U[k], k) for k in range(3)])
    X = (X + 1e-3 * rng.standard_normal(dims)).astype(np.float32)
    Xd = torch.from_numpy(X).cuda()
    Ut = [torch.from_numpy(np.ascontiguousarray(u.T)).cuda() for u in U]
    G = kernels.tucker_operator_ordered(Xd, [(k, Ut[k]) for k in range(3)])
    Gref = oracle.nmode_products_seq(X, [(U[k].T, k) for k in range(3)])
    assert rel(G.cpu().numpy(), Gref) <= 1e-5
    assert rel(G.cpu().numpy(), G0) <= 1e-2  # projection identity up to the noise
    Xr = kernels.tucker_reconstruct(G, [torch.from_numpy(u).cuda() for u in U])
    Xref = oracle.nmode_products_seq(Gref, [(U[k], k) for k in range(3)])
    assert rel(Xr.cpu().numpy(), Xref) <= 1e-5
    # HOOI's chain for mode 1: every other mode, descending (tucker.rs:1122-1140)
    Y = kernels.nmode_products_seq(Xd, [(Ut[2], 2), (Ut[0], 0)])
    assert tuple(Y.shape) == (24, 80, 16)
    assert rel(Y.cpu().numpy(), oracle.nmode_products_seq(X, [(U[2].T, 2), (U[0].T, 0)])) <= 1e-5
