"""GPU parity of the MTTKRP variant entry points and of CpDecomp.extract_weights.

- mttkrp_blocked / _blocked_parallel / _fused / _fused_parallel
  (tenrso-kernels/src/mttkrp.rs:395-498, 528-609, 783-878, 903-992) share the
  mttkrp contract; the reference's own tests hold them equal to the standard
  mttkrp (mttkrp.rs:617-720 for tiles {1, 2, 5, 10, 20}, 1046-1145 for the
  fused forms, tests/integration_tests.rs:539-575 at 80^3).  Here each is held
  to the FP64 oracle (relative Frobenius <= 1e-5, the FP32 bar) and to the
  device mttkrp bit for bit.
- mttkrp_all_modes (utils.rs:282-311), through the Python mirror and through
  the C entry tenrso_mttkrp_all_modes_f32 called with ctypes.
- extract_weights (cp/types.rs:421-446) against the FP64 oracle restatement.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import oracle  # noqa: E402
from tenrso_amd import _lib, decomp, kernels  # noqa: E402


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


VARIANT_SHAPES = [((5, 6, 7), 3), ((20, 32, 24), 8), ((80, 80, 80), 16), ((16, 12, 32, 8), 5)]


@pytest.mark.parametrize("shape,rank", VARIANT_SHAPES)
def test_mttkrp_variants_match_standard_and_oracle(shape, rank):
    print(f"[case] mttkrp variants {shape} rank {rank}", flush=True)
    rng = np.random.default_rng(sum(shape) + rank)
    X = rng.standard_normal(shape).astype(np.float32)
    F = [rng.standard_normal((n, rank)).astype(np.float32) for n in shape]
    Xd, Fd = dev(X), [dev(f) for f in F]
    X64, F64 = X.astype(np.float64), [f.astype(np.float64) for f in F]
    for mode in range(len(shape)):
        ref = oracle.mttkrp_fast(X64, F64, mode)
        std = kernels.mttkrp(Xd, Fd, mode)
        assert rel(std.cpu().numpy(), ref) <= 1e-5
        outs = {"fused": kernels.mttkrp_fused(Xd, Fd, mode),
                "fused_parallel": kernels.mttkrp_fused_parallel(Xd, Fd, mode)}
        for tile in (1, 2, 5, 10, 20):
            outs[f"blocked{tile}"] = kernels.mttkrp_blocked(Xd, Fd, mode, tile)
            outs[f"blocked_parallel{tile}"] = kernels.mttkrp_blocked_parallel(Xd, Fd, mode, tile)
        for name, got in outs.items():
            assert got.shape == (shape[mode], rank), name
            assert torch.equal(got, std), name
            assert rel(got.cpu().numpy(), ref) <= 1e-5, name
        # the oracle's f32 restatement of the blocked-parallel summation order
        # agrees to f32 rounding
        assert rel(oracle.mttkrp_blocked_parallel(X, F, mode, tile=5), ref) <= 1e-6
    for fn in (kernels.mttkrp_blocked, kernels.mttkrp_blocked_parallel):
        with pytest.raises(ValueError, match="Tile size must be greater than 0"):
            fn(Xd, Fd, 0, 0)


@pytest.mark.parametrize("shape,rank", [((24, 32, 40), 6), ((64, 64, 64), 32), ((9, 10, 11, 32), 4)])
def test_mttkrp_all_modes_python_and_c_entry(shape, rank):
    print(f"[case] mttkrp_all_modes {shape} rank {rank}", flush=True)
    rng = np.random.default_rng(7 * rank)
    X = rng.standard_normal(shape).astype(np.float32)
    F = [rng.standard_normal((n, rank)).astype(np.float32) for n in shape]
    Xd, Fd = dev(X), [dev(f) for f in F]
    refs = [oracle.mttkrp_fast(X.astype(np.float64), [f.astype(np.float64) for f in F], m)
            for m in range(len(shape))]
    py = kernels.mttkrp_all_modes(Xd, Fd)
    host = kernels.mttkrp_all_modes(X, F)
    outs = [torch.empty((n, rank), dtype=torch.float32, device="cuda") for n in shape]
    ctx = _lib.torch_context(Xd)
    fptrs = (C.c_void_p * len(Fd))(*[f.data_ptr() for f in Fd])
    optrs = (C.c_void_p * len(outs))(*[o.data_ptr() for o in outs])
    shp = (C.c_int64 * len(shape))(*shape)
    _lib.check(_lib.lib().tenrso_mttkrp_all_modes_f32(ctx.ptr, Xd.data_ptr(), len(shape), shp, fptrs, rank, optrs))
    torch.cuda.synchronize()
    assert len(py) == len(host) == len(shape)
    for m in range(len(shape)):
        assert rel(py[m].cpu().numpy(), refs[m]) <= 1e-5
        assert rel(host[m], refs[m]) <= 1e-5
        assert rel(outs[m].cpu().numpy(), refs[m]) <= 1e-5
        assert torch.equal(outs[m], py[m])
    # the C entry validates like mttkrp: a mode count below 2 is rejected
    rc = _lib.lib().tenrso_mttkrp_all_modes_f32(ctx.ptr, Xd.data_ptr(), 1, shp, fptrs, rank, optrs)
    assert rc == _lib.ERR_INVALID_ARGUMENT
    assert "at least 2 factor matrices" in _lib.last_error()


def _factors_with_small_columns(shape, R, rng):
    F = [rng.standard_normal((n, R)).astype(np.float32) * rng.uniform(0.1, 10.0, R).astype(np.float32)
         for n in shape]
    F[0][:, 1] = 0.0                 # zero column: norm 0 <= eps, left as is, weight unchanged
    F[1][:, 2] = 1e-9                # norm below f32 epsilon, left as is
    return F


@pytest.mark.parametrize("shape,R", [((30, 20, 25), 5), ((256, 256, 256), 64), ((7, 9, 5, 6), 70),
                                     ((10000, 8000, 6000), 32)])
def test_extract_weights_vs_oracle(shape, R):
    """Device extract_weights against the FP64 restatement of cp/types.rs:421-446:
    normalised factors and weights within 1e-6 relative (fp32 outputs of FP64
    norms), the columns at or below f32 epsilon left untouched, and the weighted
    reconstruction unchanged (small shapes: the reconstruction is the invariant
    extract_weights preserves)."""
    print(f"[case] extract_weights {shape} rank {R}", flush=True)
    rng = np.random.default_rng(R)
    F = _factors_with_small_columns(shape, R, rng)
    ref_F, ref_w = oracle.extract_weights(F)
    Fd = [dev(f) for f in F]
    got_F, got_w = kernels.extract_weights(Fd)
    assert all(g is d for g, d in zip(got_F, Fd))  # in place on contiguous fp32 CUDA factors
    assert rel(got_w.cpu().numpy(), ref_w) <= 1e-6
    for g, r, f in zip(got_F, ref_F, F):
        assert rel(g.cpu().numpy(), r) <= 1e-6
    assert np.array_equal(got_F[0][:, 1].cpu().numpy(), F[0][:, 1])
    assert np.array_equal(got_F[1][:, 2].cpu().numpy(), F[1][:, 2])
    if np.prod(shape) <= 1_000_000:
        before = oracle.cp_reconstruct([f.astype(np.float64) for f in F])
        after = kernels.cp_reconstruct(got_F, got_w).cpu().numpy()
        assert rel(after, before) <= 1e-5
    # numpy and torch CPU factors come back as numpy / torch CPU with the same values
    hF, hw = kernels.extract_weights([f.copy() for f in F])
    assert isinstance(hw, np.ndarray) and np.array_equal(hw, got_w.cpu().numpy())
    tF, tw = kernels.extract_weights([torch.from_numpy(f.copy()) for f in F])
    assert not tw.is_cuda and torch.equal(tw, got_w.cpu())
    assert all(np.array_equal(h, g.cpu().numpy()) for h, g in zip(hF, got_F))


def test_cpdecomp_extract_weights_after_cp_als():
    """CpDecomp.extract_weights on a cp_als result (the reference's usage): unit
    columns, weights = product of norms, reconstruction unchanged."""
    print("[case] CpDecomp.extract_weights after cp_als", flush=True)
    rng = np.random.default_rng(3)
    X = rng.standard_normal((20, 24, 28)).astype(np.float32)
    init = [rng.standard_normal((n, 4)).astype(np.float32) for n in X.shape]
    cp = decomp.cp_als_with_init(dev(X), [dev(f) for f in init], 5, 0.0)
    raw = [f.clone() for f in cp.factors]
    rec0 = cp.reconstruct(X.shape).cpu().numpy()
    cp.extract_weights()
    ref_F, ref_w = oracle.extract_weights([f.cpu().numpy() for f in raw])
    assert rel(cp.weights.cpu().numpy(), ref_w) <= 1e-6
    for f in cp.factors:
        assert np.allclose(torch.linalg.vector_norm(f, dim=0).cpu().numpy(), 1.0, atol=1e-6)
    assert rel(cp.reconstruct(X.shape).cpu().numpy(), rec0) <= 1e-5


def test_extract_weights_rejects_bad_input():
    with pytest.raises(ValueError, match="Need at least one factor matrix"):
        kernels.extract_weights([])
    with pytest.raises(ValueError, match="columns, expected 3"):
        kernels.extract_weights([np.ones((4, 3), np.float32), np.ones((5, 2), np.float32)])
    ctx = _lib.context_for(0)
    shp = (C.c_int64 * 1)(4)
    rc = _lib.lib().tenrso_cp_extract_weights_f32(ctx.ptr, 0, shp, None, 3, None)
    assert rc == _lib.ERR_INVALID_ARGUMENT
