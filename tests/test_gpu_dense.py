"""GPU parity of the dense path: HIP library (through the C ABI) vs the FP64 oracle.

Bars (written in each test): FP32 MTTKRP relative Frobenius error <= 1e-5 vs FP64;
Khatri-Rao bit-exact vs the f32 restatement; CP-ALS fit within 1e-4 of the FP64
oracle run from the same initial factors.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import oracle  # noqa: E402
from tenrso_amd import decomp, kernels, synth  # noqa: E402


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def f64(fs):
    return [np.asarray(f, dtype=np.float64) for f in fs]


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    print(f"[gpu] {torch.cuda.get_device_name(0)}", flush=True)


# shapes: fast MFMA path (fastest other mode % 32 == 0), ragged rows, fallback path,
# 2-way, 4-way and 5-way tensors; 64-row tiles at rank <= 32 (rows a multiple of 64,
# or >= 2048: mode 0 of (2050, 8, 32), the last mode of (8, 16, 2052) with a partial
# last tile, both modes 0 and 2 of (128, 96, 64))
SHAPES = [
    ((2, 3, 4), 2), ((4, 5, 6), 3), ((40, 64, 32), 8), ((17, 32, 64), 5), ((64, 96, 32), 16), ((33, 50, 70), 7),
    ((128, 64), 20), ((24, 32, 16, 32), 6), ((6, 5, 4, 3, 32), 3), ((100, 64, 64), 64), ((20, 32, 32), 70),
    ((128, 96, 64), 32), ((2050, 8, 32), 16), ((8, 16, 2052), 32),
]


@pytest.mark.parametrize("shape,rank", SHAPES)
def test_mttkrp_f32_all_modes_vs_fp64(shape, rank):
    print(f"[case] mttkrp f32 shape={shape} rank={rank}", flush=True)
    rng = np.random.default_rng(hash(shape) % 2**32)
    X = rng.standard_normal(shape).astype(np.float32)
    F = [rng.standard_normal((n, rank)).astype(np.float32) for n in shape]
    Xd, Fd = dev(X), [dev(f) for f in F]
    for mode in range(len(shape)):
        got = kernels.mttkrp(Xd, Fd, mode).cpu().numpy()
        ref = oracle.mttkrp_fast(X.astype(np.float64), f64(F), mode)
        assert got.shape == ref.shape
        assert rel(got, ref) <= 1e-5, (mode, rel(got, ref))
        # host path (numpy in, numpy out) gives the same values
        if np.prod(shape) <= 100_000:
            host = kernels.mttkrp(X, F, mode)
            assert rel(host, ref) <= 1e-5


def test_inputs_of_other_dtype_or_device_are_converted():
    """Factors that are float64 CUDA tensors (the reference's usual T = f64),
    torch CPU tensors or numpy arrays are converted to float32 on X's device
    before any launch, never passed through as raw pointers; a torch CPU tensor
    takes the host path.  Every result matches the FP64 oracle."""
    print("[case] dtype / device conversion of factors", flush=True)
    rng = np.random.default_rng(41)
    X = rng.standard_normal((24, 32, 40)).astype(np.float32)
    F = [rng.standard_normal((n, 6)) for n in X.shape]  # float64
    Xd = dev(X)
    for mode in range(3):
        ref = oracle.mttkrp_fast(X.astype(np.float64), F, mode)
        variants = {
            "f64 cuda": [torch.from_numpy(f).cuda() for f in F],
            "f32 cpu": [torch.from_numpy(f.astype(np.float32)) for f in F],
            "f64 numpy": F,
            "f16 cuda": [torch.from_numpy(f).cuda().half() for f in F],
        }
        for name, fs in variants.items():
            got = kernels.mttkrp(Xd, fs, mode)
            assert got.is_cuda and got.dtype == torch.float32
            tol = 1e-5 if name != "f16 cuda" else 5e-3
            assert rel(got.cpu().numpy(), ref) <= tol, (name, mode)
        host = kernels.mttkrp(torch.from_numpy(X), [torch.from_numpy(f) for f in F], mode)
        assert not host.is_cuda and rel(host.numpy(), ref) <= 1e-5
        b16 = kernels.mttkrp_bf16(kernels.to_bf16(Xd), [torch.from_numpy(f) for f in F], mode)
        assert rel(b16.cpu().numpy(), ref) <= 5e-3
    A = rng.standard_normal((7, 3))
    B = rng.standard_normal((5, 3))
    kr_ref = oracle.khatri_rao(A.astype(np.float32), B.astype(np.float32), dtype=np.float32)
    assert np.array_equal(kernels.khatri_rao(torch.from_numpy(A).cuda(), torch.from_numpy(B)).cpu().numpy(), kr_ref)
    kc = kernels.khatri_rao(torch.from_numpy(A), torch.from_numpy(B))
    assert not kc.is_cuda and np.array_equal(kc.numpy(), kr_ref)
    w = rng.random(6)
    rec = kernels.cp_reconstruct([torch.from_numpy(f).cuda() for f in F], torch.from_numpy(w))
    assert rel(rec.cpu().numpy(), oracle.cp_reconstruct(F, w)) <= 1e-5
    rec_c = kernels.cp_reconstruct([torch.from_numpy(f) for f in F])
    assert not rec_c.is_cuda and rel(rec_c.numpy(), oracle.cp_reconstruct(F)) <= 1e-5
    with pytest.raises(TypeError):
        kernels.mttkrp(Xd.double(), [f.cuda() for f in variants["f64 cuda"]], 0)
    with pytest.raises(TypeError):
        kernels.to_bf16(torch.from_numpy(X))


def test_mttkrp_reference_kats_on_gpu():
    # mttkrp.rs:1000-1043 and property_tests.rs:208-229 evaluated by the HIP path
    X = np.arange(1, 13, dtype=np.float32).reshape(2, 3, 2)
    u1 = np.array([[1.0, 0.0], [2.0, 1.0]], np.float32)
    u2 = np.array([[1.0, 0.0], [0.5, 1.0], [0.0, 1.0]], np.float32)
    u3 = np.array([[1.0, 0.0], [0.0, 1.0]], np.float32)
    manual = np.zeros((3, 2))
    for i1 in range(3):
        for r in range(2):
            manual[i1, r] = sum(X[i0, i1, i2] * u1[i0, r] * u3[i2, r] for i0 in range(2) for i2 in range(2))
    assert np.abs(kernels.mttkrp(dev(X), [dev(u1), dev(u2), dev(u3)], 1).cpu().numpy() - manual).max() < 1e-5
    for size, rank in [(3, 2), (7, 4), (32, 3), (64, 16)]:
        ones = np.ones((size,) * 3, np.float32)
        F = [np.ones((size, rank), np.float32)] * 3
        for mode in range(3):
            got = kernels.mttkrp(dev(ones), [dev(f) for f in F], mode).cpu().numpy()
            assert np.abs(got - size * size).max() < 1e-3


def test_mttkrp_c2_shape_fp32():
    print("[case] 256^3 rank-64 model tensor, all modes", flush=True)
    X, F, w, init = synth.dense_256_r64()
    Xd = dev(X)
    Fd = [dev(f) for f in init]
    X64 = X.astype(np.float64)
    for mode in range(3):
        got = kernels.mttkrp(Xd, Fd, mode).cpu().numpy()
        ref = oracle.mttkrp_fast(X64, f64(init), mode)
        e = rel(got, ref)
        print(f"  mode {mode}: rel frob err {e:.3e}", flush=True)
        assert e <= 1e-5


def bf16_round(a):
    """Round float32 -> bfloat16 (RNE) -> float32 in numpy."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.parametrize("shape,rank", [((128, 128, 64), 32), ((64, 96, 128), 64), ((32, 32, 32, 32), 16),
                                        ((64, 128), 5), ((96, 64), 8), ((6, 5, 8, 7, 64), 3),
                                        ((4, 5, 3, 6, 4, 64), 12)])
def test_mttkrp_bf16_error_bound(shape, rank):
    print(f"[case] bf16 mttkrp {shape} rank {rank}", flush=True)
    rng = np.random.default_rng(1)
    X = rng.standard_normal(shape).astype(np.float32)
    F = [rng.standard_normal((n, rank)).astype(np.float32) for n in X.shape]
    Xd = dev(X)
    Xb = kernels.to_bf16(Xd)
    # conversion is RNE, identical to torch's and to the numpy rounding
    assert torch.equal(Xb.view(torch.int16), Xd.to(torch.bfloat16).view(torch.int16))
    assert np.array_equal(Xb.float().cpu().numpy(), bf16_round(X))
    Fd = [dev(f) for f in F]
    Xr64 = bf16_round(X).astype(np.float64)
    for mode in range(len(shape)):
        got = kernels.mttkrp_bf16(Xb, Fd, mode).cpu().numpy()
        ref = oracle.mttkrp_fast(X.astype(np.float64), f64(F), mode)
        # the factors enter as an exact bf16 hi+lo pair, so the kernel computes the
        # MTTKRP of the bf16-rounded tensor to FP32-accumulation accuracy
        ref_b = oracle.mttkrp_fast(Xr64, f64(F), mode)
        e, eb = rel(got, ref), rel(got, ref_b)
        print(f"  mode {mode}: rel frob err vs fp64 {e:.3e}, vs fp64 of bf16(X) {eb:.3e}", flush=True)
        assert e < 5e-3
        assert eb < 2e-5


def test_khatri_rao_bit_exact():
    rng = np.random.default_rng(2)
    for (I, J, K) in [(2, 2, 2), (3, 2, 2), (37, 19, 5), (500, 64, 32), (128, 128, 64)]:
        A = rng.standard_normal((I, K)).astype(np.float32)
        B = rng.standard_normal((J, K)).astype(np.float32)
        ref = oracle.khatri_rao(A, B, dtype=np.float32)
        assert np.array_equal(kernels.khatri_rao(dev(A), dev(B)).cpu().numpy(), ref)
        assert np.array_equal(kernels.khatri_rao(A, B), ref)
    # reference KAT khatri_rao.rs:41-55
    c = kernels.khatri_rao(np.array([[1.0, 2.0], [3.0, 4.0]]), np.array([[5.0, 6.0], [7.0, 8.0]]))
    assert c[:, 0].tolist() == [5, 7, 15, 21] and c[:, 1].tolist() == [12, 16, 24, 32]


def test_cp_reconstruct_vs_oracle():
    rng = np.random.default_rng(3)
    for shape, R in [((3, 4, 5), 3), ((32, 16, 8, 4), 5), ((64, 64, 64), 16)]:
        F = [rng.standard_normal((n, R)).astype(np.float32) for n in shape]
        w = rng.random(R).astype(np.float32)
        got = kernels.cp_reconstruct([dev(f) for f in F], dev(w)).cpu().numpy()
        ref = oracle.cp_reconstruct(f64(F), w.astype(np.float64))
        assert rel(got, ref) < 1e-5


def test_cp_als_small_vs_oracle():
    print("[case] cp_als 64^3 rank 8 (20 iters, tol 1e-4)", flush=True)
    X, _, init = synth.dense_64_r8()
    cp = decomp.cp_als_with_init(dev(X), [dev(f) for f in init], 20, 1e-4)
    ref_F, info = oracle.cp_als(X, init, 20, 1e-4)
    print(f"  gpu fit {cp.fit:.8f} iters {cp.iters} {cp.convergence.reason}; "
          f"oracle fit {info['fit']:.8f} iters {info['iters']} {info['reason']}", flush=True)
    assert abs(cp.fit - info["fit"]) <= 1e-4
    assert cp.iters == info["iters"]
    assert cp.convergence.reason.name == info["reason"]
    hist = np.array(cp.convergence.fit_history)
    assert np.abs(hist - info["fit_history"]).max() <= 1e-4
    for g, r in zip(cp.factors, ref_F):
        assert rel(g.cpu().numpy(), r) < 1e-3


def test_cp_als_c2_vs_oracle():
    print("[case] cp_als 256^3 rank 64, 10 iters", flush=True)
    X, _, _, init = synth.dense_256_r64()
    cp = decomp.cp_als_with_init(dev(X), [dev(f) for f in init], 10, 0.0)
    ref_F, info = oracle.cp_als(X, init, 10, 0.0)
    errs = [rel(g.cpu().numpy(), r) for g, r in zip(cp.factors, ref_F)]
    print(f"  gpu fit {cp.fit:.8f} oracle {info['fit']:.8f} |dfit| {abs(cp.fit - info['fit']):.2e} "
          f"factor rel errs {['%.2e' % e for e in errs]} gpu {cp.seconds:.4f}s oracle {info['seconds']:.1f}s",
          flush=True)
    assert abs(cp.fit - info["fit"]) <= 1e-4
    assert max(errs) < 1e-3


def test_cp_als_host_path_and_errors():
    X, _, init = synth.dense_64_r8()
    cp = decomp.cp_als_with_init(X, init, 3, 0.0)
    _, info = oracle.cp_als(X, init, 3, 0.0)
    assert abs(cp.fit - info["fit"]) <= 1e-4 and isinstance(cp.factors[0], np.ndarray)
    with pytest.raises(decomp.InvalidRank, match="Invalid rank: 0"):
        decomp.cp_als(X, 0, 5, 1e-4)
    with pytest.raises(decomp.InvalidRank):
        decomp.cp_als(X, 65, 5, 1e-4)
    with pytest.raises(decomp.InvalidTolerance):
        decomp.cp_als(X, 4, 5, 1.5)
    cp0 = decomp.cp_als_with_init(X, init, 5, 1e-4, time_limit=0.0)
    assert cp0.iters == 0 and cp0.convergence.reason == decomp.ConvergenceReason.TimeLimit


CONSTRAINTS = [
    ("nonnegative", dict(nonnegative=True), lambda: decomp.CpConstraints.nonnegative_()),
    ("l2", dict(l2=0.5), lambda: decomp.CpConstraints.l2_regularized(0.5)),
    ("l1", dict(l1=0.05), lambda: decomp.CpConstraints.l1_regularized(0.05)),
    ("tikhonov1", dict(tikhonov_order=1, tikhonov_lambda=0.3), lambda: decomp.CpConstraints.tikhonov(0.3, 1)),
    ("tikhonov2", dict(tikhonov_order=2, tikhonov_lambda=0.3), lambda: decomp.CpConstraints.tikhonov(0.3, 2)),
    ("orthogonal", dict(orthogonal=True), lambda: decomp.CpConstraints.orthogonal_()),
    ("nonneg+orthogonal", dict(nonnegative=True, orthogonal=True),
     lambda: decomp.CpConstraints(nonnegative=True, orthogonal=True)),
]


@pytest.mark.parametrize("name,orc_kw,make", CONSTRAINTS, ids=[c[0] for c in CONSTRAINTS])
def test_cp_als_constrained_vs_oracle(name, orc_kw, make):
    # cp_als_constrained (core.rs:225-408) on 48x40x36 rank 6 from the same init:
    # fit history within 1e-4 of the FP64 oracle, same iteration count and reason,
    # factors within 1e-3.  The data is nonnegative so the input check passes.
    # Orthogonal: Q of the positive-diagonal QR on both sides (the third-party
    # routine's sign convention is not visible); the factors must come out with
    # orthonormal columns.
    print(f"[case] cp_als constrained {name} 48x40x36 rank 6", flush=True)
    rng = np.random.default_rng(31)
    shape, R = (48, 40, 36), 6
    A = [rng.random((n, R)) for n in shape]
    X = (np.einsum("ir,jr,kr->ijk", *A) + 0.01 * rng.random(shape)).astype(np.float32)
    init = [rng.random((n, R)).astype(np.float32) + 0.1 for n in shape]
    cp = decomp.cp_als_with_init(dev(X), [dev(f) for f in init], 15, 1e-6, constraints=make())
    ref_F, info = oracle.cp_als(X, init, 15, 1e-6, **orc_kw)
    errs = [rel(g.cpu().numpy(), r) for g, r in zip(cp.factors, ref_F)]
    print(f"  gpu fit {cp.fit:.8f} iters {cp.iters} {cp.convergence.reason.name}; oracle fit {info['fit']:.8f} "
          f"iters {info['iters']} {info['reason']}; factor rel errs {['%.1e' % e for e in errs]}", flush=True)
    assert abs(cp.fit - info["fit"]) <= 1e-4
    assert cp.iters == info["iters"] and cp.convergence.reason.name == info["reason"]
    assert np.abs(np.array(cp.convergence.fit_history) - info["fit_history"]).max() <= 1e-4
    assert max(errs) < 1e-3
    if orc_kw.get("orthogonal"):
        for g in cp.factors:
            g = g.cpu().numpy().astype(np.float64)
            assert np.allclose(g.T @ g, np.eye(R), atol=1e-5)
    if orc_kw.get("nonnegative") and not orc_kw.get("orthogonal"):
        assert all(float(g.min()) >= 0.0 for g in cp.factors)


def test_cp_als_bf16_fit_close():
    """BF16 ALS at 256^3 r64, 10 iterations.  Its leading modes contract bf16(X),
    its last mode the fp32 tensor exactly, so (a) the reported fit is the fit of
    the returned factors to X (FP64 oracle.fit, within 1e-6); (b) the run follows
    the FP64 oracle of that same computation (x_lead = bf16(X)): factors within
    1e-3, fit history within 1e-5 (measured 1.3e-5 and 2.0e-6: FP32 accumulation
    and bf16 Khatri-Rao rows against the FP64 model); (c) the fit is within 1e-4
    of the FP64 ALS on X (measured 1.7e-5)."""
    print("[case] cp_als bf16 256^3 rank 64, 10 iters", flush=True)
    X, _, _, init = synth.dense_256_r64()
    cp = decomp.cp_als_with_init(dev(X), [dev(f) for f in init], 10, 0.0, precision="bf16")
    _, info = oracle.cp_als(X, init, 10, 0.0)
    ref_F, info_b = oracle.cp_als(X, init, 10, 0.0, x_lead=bf16_round(X))
    got = [g.cpu().numpy() for g in cp.factors]
    fit_exact = oracle.fit(X, got)
    errs = [rel(g, r) for g, r in zip(got, ref_F)]
    dh = np.abs(np.array(cp.convergence.fit_history) - info_b["fit_history"]).max()
    print(f"  bf16 fit {cp.fit:.10f} fit(X, factors) {fit_exact:.10f} |d| {abs(cp.fit - fit_exact):.2e}; "
          f"fp64 ALS {info['fit']:.8f} |dfit| {abs(cp.fit - info['fit']):.2e}; vs bf16-lead oracle: "
          f"factors {['%.1e' % e for e in errs]} hist {dh:.1e}", flush=True)
    assert abs(cp.fit - fit_exact) <= 1e-6
    assert max(errs) <= 1e-3 and dh <= 1e-5
    assert abs(cp.fit - info["fit"]) <= 1e-4


def test_cp_als_bf16_tolerance_stop_matches_fp64():
    """With tol > 0 the BF16 run stops where the FP64 ALS on X does (same
    iterations and reason), since its fit is the fit to X: 64^3 rank 8, 20
    iterations, tol 1e-4 (the CPU-path workload)."""
    print("[case] cp_als bf16 64^3 rank 8, 20 iters tol 1e-4", flush=True)
    X, _, init = synth.dense_64_r8()
    cp = decomp.cp_als_with_init(dev(X), [dev(f) for f in init], 20, 1e-4, precision="bf16")
    _, info = oracle.cp_als(X, init, 20, 1e-4)
    got = [g.cpu().numpy() for g in cp.factors]
    fit_exact = oracle.fit(X, got)
    print(f"  bf16 fit {cp.fit:.10f} iters {cp.iters} {cp.convergence.reason.name}; fp64 fit {info['fit']:.10f} "
          f"iters {info['iters']} {info['reason']}; fit(X, factors) {fit_exact:.10f}", flush=True)
    assert cp.iters == info["iters"] and cp.convergence.reason.name == info["reason"]
    assert abs(cp.fit - fit_exact) <= 1e-6
    assert abs(cp.fit - info["fit"]) <= 1e-4


def test_cp_als_oscillation_vs_oracle():
    """The Oscillation stop (core.rs:149-164): an orthogonally constrained fit of a
    noise tensor whose fit drops six times; the FP32 device run must stop at the
    same iteration with the same count (the smallest fit change of the run,
    5.9e-5, is far above the FP32 fit error)."""
    print("[case] cp_als oscillation (12, 8, 9) rank 5 orthogonal", flush=True)
    rng = np.random.default_rng(126)
    shp = tuple(int(x) for x in rng.integers(8, 24, size=3))
    R = int(rng.integers(2, 6))
    X = rng.standard_normal(shp).astype(np.float32)
    init = [rng.standard_normal((n, R)).astype(np.float32) for n in shp]
    cp = decomp.cp_als_with_init(dev(X), [dev(f) for f in init], 40, 0.0,
                                 constraints=decomp.CpConstraints.orthogonal_())
    _, info = oracle.cp_als(X, init, 40, 0.0, orthogonal=True)
    print(f"  gpu iters {cp.iters} {cp.convergence.reason.name} osc {cp.convergence.oscillation_count}; oracle "
          f"iters {info['iters']} {info['reason']} osc {info['oscillation_count']}", flush=True)
    assert info["reason"] == "Oscillation"
    assert cp.convergence.reason.name == "Oscillation"
    assert cp.iters == info["iters"] and cp.convergence.oscillation_count == info["oscillation_count"]
    assert cp.convergence.oscillated
    assert np.abs(np.array(cp.convergence.fit_history) - info["fit_history"]).max() <= 1e-5


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_cp_als_linalg_error_after_ridge(precision):
    """A Gamma that stays singular after the ridge retry is CpError::LinalgError
    (helpers.rs:477-490), on the device as in the oracle: an infinite entry in an
    initial factor makes every Gamma_n non-finite.  The context stays usable."""
    print(f"[case] cp_als LinalgError {precision}", flush=True)
    rng = np.random.default_rng(12)
    X = rng.standard_normal((40, 36, 32)).astype(np.float32)
    init = [rng.standard_normal((n, 4)).astype(np.float32) for n in X.shape]
    bad = [f.copy() for f in init]
    bad[2][1, 0] = np.inf
    with pytest.raises(oracle.LinalgError):
        oracle.cp_als(X, bad, 5, 0.0)
    with pytest.raises(decomp.LinalgError, match="Linear algebra error"):
        decomp.cp_als_with_init(dev(X), [dev(f) for f in bad], 5, 0.0, precision=precision)
    cp = decomp.cp_als_with_init(dev(X), [dev(f) for f in init], 3, 0.0, precision=precision)
    assert np.isfinite(cp.fit)


@pytest.mark.parametrize("shape", [(40, 36, 32), (35, 33, 36), (24, 20, 30)])
def test_cp_als_bf16_setup_paths(shape):
    """The BF16 ALS input (bf16 copy, last-to-front bf16 copy, norm) comes from one
    fused pass when the last mode is a multiple of 4 and X is 16-byte aligned, from
    three separate passes otherwise.  A copy of X at an odd element offset takes
    the separate passes; both runs must give the same factors bit for bit (the
    norms differ only in FP64 summation order), and the fit must stay within the
    bf16 input rounding of the FP64 oracle's."""
    import torch
    print(f"[case] cp_als bf16 setup paths {shape}", flush=True)
    rng = np.random.default_rng(5)
    X = synth.cp_tensor(synth.normal_factors(shape, 6, rng), noise=1e-2, rng=rng)
    init = synth.init_factors(shape, 6, seed=2)
    Xa = dev(X)
    buf = torch.zeros(X.size + 1, dtype=torch.float32, device=Xa.device)
    buf[1:] = Xa.reshape(-1)
    Xm = buf[1:].view(shape)  # 4-byte offset: not 16-byte aligned
    assert Xm.data_ptr() % 16 != 0
    cpa = decomp.cp_als_with_init(Xa, [dev(f) for f in init], 8, 0.0, precision="bf16")
    cpm = decomp.cp_als_with_init(Xm, [dev(f) for f in init], 8, 0.0, precision="bf16")
    fit_exact = oracle.fit(X, [g.cpu().numpy() for g in cpa.factors])
    print(f"  fit aligned {cpa.fit:.10f} offset {cpm.fit:.10f} fit(X, factors) {fit_exact:.10f}", flush=True)
    for a, m in zip(cpa.factors, cpm.factors):
        assert torch.equal(a, m)
    assert abs(cpa.fit - cpm.fit) <= 1e-12
    assert abs(cpa.fit - fit_exact) <= 1e-6


@pytest.mark.parametrize("shape", [(256, 256, 256), (512, 512, 128)])
def test_c3_congruent_precision_vs_fp64(shape):
    """Ill-conditioned Grams (pairwise column congruence 0.9, noise 1e-4), rank 64,
    10 ALS iterations from the same initial factors.  The normal equations amplify
    rounding by the condition number of Gamma.
    FP32 against the FP64 oracle on X: fit within 1e-5, factors within 1e-2
    (measured 7e-7 / 1.3e-3).
    BF16 against two FP64 oracle runs:
      - the FP64 model of the BF16 computation (leading modes on bf16(X), the last
        mode, norm and fit on X: x_lead = bf16(X)): fit within 1e-5 and factors
        within 1e-2, the FP32 bars, so that what separates BF16 from the FP64 ALS
        on X is the bf16 rounding of the input and nothing else;
      - its reported fit is the fit of its factors to X (oracle.fit, within 1e-6).
    Against the FP64 ALS on X itself BF16 differs by the amplified input rounding
    (measured fit 3e-4, factors 0.2), recorded and bounded loosely (1e-3 / 0.5).
    First-sweep MTTKRP within 1e-5 (FP32) and 5e-3 (BF16).  (The reference's own
    f32 path reports fit 1.0 here: its sequential f32 fit identity cancels.)"""
    print(f"[case] C3 congruent {shape} rank 64, 10 iters fp32/bf16 vs fp64", flush=True)
    X, _, _, init = synth.congruent(shape=shape)
    ref_F, info = oracle.cp_als(X, init, 10, 0.0)
    lead_F, info_l = oracle.cp_als(X, init, 10, 0.0, x_lead=bf16_round(X))
    Xd = dev(X)
    Fd = [dev(f) for f in init]
    m_ref = oracle.mttkrp_fast(X, f64(init), 0)
    m32 = kernels.mttkrp(Xd, Fd, 0).cpu().numpy()
    m16 = kernels.mttkrp_bf16(kernels.to_bf16(Xd), Fd, 0).cpu().numpy()
    e32, e16 = rel(m32, m_ref), rel(m16, m_ref)
    print(f"  mttkrp mode 0 rel err fp32 {e32:.2e} bf16 {e16:.2e}", flush=True)
    results = {}
    for prec in ("fp32", "bf16"):
        cp = decomp.cp_als_with_init(Xd, Fd, 10, 0.0, precision=prec)
        got = [g.cpu().numpy() for g in cp.factors]
        errs = [rel(g, r) for g, r in zip(got, ref_F)]
        errs_l = [rel(g, r) for g, r in zip(got, lead_F)]
        fit_exact = oracle.fit(X, got)
        results[prec] = (abs(cp.fit - info["fit"]), max(errs), abs(cp.fit - info_l["fit"]), max(errs_l),
                         abs(cp.fit - fit_exact))
        print(f"  {prec}: fit {cp.fit:.8f} fp64 ALS {info['fit']:.8f} |dfit| {results[prec][0]:.2e} factors "
              f"{['%.2e' % e for e in errs]}; bf16-lead model fit {info_l['fit']:.8f} |dfit| {results[prec][2]:.2e} "
              f"factors {['%.2e' % e for e in errs_l]}; |fit - fit(X, factors)| {results[prec][4]:.1e}", flush=True)
    assert e32 <= 1e-5 and e16 <= 5e-3
    assert results["fp32"][0] <= 1e-5 and results["fp32"][1] <= 1e-2
    assert results["fp32"][4] <= 1e-6
    assert results["bf16"][2] <= 1e-5 and results["bf16"][3] <= 1e-2
    assert results["bf16"][4] <= 1e-6
    assert results["bf16"][0] <= 1e-3 and results["bf16"][1] <= 0.5


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("shape", [(64, 64, 64), (33, 50, 70)])
def test_cp_als_ridge_retry_vs_oracle(shape, precision):
    """A zero column in the initial factors makes every Gamma_n singular (a zero
    row and column), so every solve fails its first factorisation and takes the
    ridge retry of helpers.rs:477-490 (Gamma + eps 10R I).  The zero column
    decouples from the others, so the solution does not depend on eps: column 3
    stays exactly zero and the rest must match the FP64 oracle, which takes the
    same retry (solve path 1 on both sides)."""
    print(f"[case] cp_als ridge retry {shape} {precision}", flush=True)
    rng = np.random.default_rng(11)
    X = synth.cp_tensor(synth.normal_factors(shape, 6, rng), noise=1e-2, rng=rng)
    init = synth.init_factors(shape, 6, seed=4)
    for f in init:
        f[:, 3] = 0.0
    cp = decomp.cp_als_with_init(dev(X), [dev(f) for f in init], 6, 0.0, precision=precision)
    # BF16: the FP64 model of its computation (leading modes on bf16(X))
    lead = None if precision == "fp32" else bf16_round(X)
    ref_F, info = oracle.cp_als(X, init, 6, 0.0, x_lead=lead)
    print(f"  fit {cp.fit:.8f} oracle {info['fit']:.8f} paths {cp.solve_path_max} / {info['solve_path_max']}",
          flush=True)
    assert info["solve_path_max"] == 1
    assert cp.solve_path_max == 1
    assert abs(cp.fit - info["fit"]) <= 1e-4
    keep = [c for c in range(6) if c != 3]
    for g, r in zip(cp.factors, ref_F):
        g = g.cpu().numpy()
        assert not g[:, 3].any()
        assert rel(g[:, keep], r[:, keep]) < 1e-3


# Ranks above 64 take the wide-rank job (Gamma_n^{-1} by scalar-pivot Gauss-Jordan
# with the matrix in registers, as its own kernel), the wide row solve and the wide
# Gram: rank 72 and the bound 128 on 3-way tensors, rank 100 on a 2-way one (Gamma
# is a single Gram), and rank 66 beside a mode of 1100 rows (the chunked Gram, in
# 64-row chunks at these ranks).  The 2-way rank stays below both sizes: at a rank
# equal to the smaller size the model is exact after one sweep and the square
# factor is fixed only up to the init, so rounding is not damped at all.
ALS_WIDE = [((96, 80, 72), 72), ((130, 129, 128), 128), ((400, 300), 100), ((1100, 70, 66), 66)]


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("shape,rank", ALS_WIDE)
def test_cp_als_wide_rank_vs_oracle(shape, rank, precision):
    """The bars and the oracle of test_cp_als_nway_vs_oracle, at ranks 66 to 128."""
    test_cp_als_nway_vs_oracle(shape, rank, precision)


@pytest.mark.parametrize("name,orc_kw,make", CONSTRAINTS[:6], ids=[c[0] for c in CONSTRAINTS[:6]])
def test_cp_als_wide_rank_constrained_vs_oracle(name, orc_kw, make):
    """The regularisation band, the row epilogue and the orthogonal constraint at
    rank 70 (the wide-rank job forms the same Gamma_n + lambda D^T D or + l2 I, the
    wide row solve applies the same soft threshold and clamp, the 128-column
    Householder QR the same positive-diagonal Q): 90x80x72 from the same init, 6
    sweeps, fit history within 1e-4 of the FP64 oracle, factors within 1e-3;
    orthogonal factors with orthonormal columns.  The orthogonal case fits a model of
    normalised N(0,1) factor columns: with orthonormal factors Gamma_n = I and each
    factor is Q of the MTTKRP M, whose 70 columns are nearly collinear on the
    nonnegative data (one dominant mean direction), which leaves Q's trailing
    columns to rounding (fit history 8e-9 apart, factors 0.3)."""
    print(f"[case] cp_als wide-rank constrained {name} 90x80x72 rank 70", flush=True)
    rng = np.random.default_rng(37)
    shape, R = (90, 80, 72), 70
    if orc_kw.get("orthogonal"):
        X, _, _, init = synth.dense_256_r64(shape=shape, rank=R, noise=1e-2, seed=37, init_seed=38)
    else:
        A = [rng.random((n, R)) for n in shape]
        X = (np.einsum("ir,jr,kr->ijk", *A) + 0.01 * rng.random(shape)).astype(np.float32)
        init = [rng.random((n, R)).astype(np.float32) + 0.1 for n in shape]
    cp = decomp.cp_als_with_init(dev(X), [dev(f) for f in init], 6, 0.0, constraints=make())
    ref_F, info = oracle.cp_als(X, init, 6, 0.0, **orc_kw)
    errs = [rel(g.cpu().numpy(), r) for g, r in zip(cp.factors, ref_F)]
    dh = np.abs(np.array(cp.convergence.fit_history) - info["fit_history"]).max()
    print(f"  gpu fit {cp.fit:.8f} oracle {info['fit']:.8f} paths {cp.solve_path_max} / {info['solve_path_max']} "
          f"hist {dh:.1e} factor rel errs {['%.1e' % e for e in errs]}", flush=True)
    assert cp.iters == info["iters"] and cp.solve_path_max == info["solve_path_max"]
    assert dh <= 1e-4
    assert max(errs) < 1e-3
    if orc_kw.get("orthogonal"):
        for g in cp.factors:
            g = g.cpu().numpy().astype(np.float64)
            assert np.allclose(g.T @ g, np.eye(R), atol=1e-5)
    elif orc_kw.get("nonnegative"):
        assert all(float(g.min()) >= 0.0 for g in cp.factors)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_cp_als_wide_rank_ridge_retry_vs_oracle(precision):
    """The ridge retry in the wide-rank job: a zero column makes every Gamma_n
    singular, the first elimination meets a zero pivot and the retry (Gamma + eps
    10R I) succeeds, on the device and in the oracle (solve path 1 on both).  The
    zero column stays exactly zero; the others match the oracle."""
    shape, R = (80, 72, 70), 68
    print(f"[case] cp_als wide-rank ridge retry {shape} rank {R} {precision}", flush=True)
    rng = np.random.default_rng(41)
    X = synth.cp_tensor(synth.normal_factors(shape, R, rng), noise=1e-2, rng=rng)
    init = synth.init_factors(shape, R, seed=5)
    for f in init:
        f[:, 3] = 0.0
    cp = decomp.cp_als_with_init(dev(X), [dev(f) for f in init], 5, 0.0, precision=precision)
    lead = None if precision == "fp32" else bf16_round(X)
    ref_F, info = oracle.cp_als(X, init, 5, 0.0, x_lead=lead)
    keep = [c for c in range(R) if c != 3]
    errs = [rel(g.cpu().numpy()[:, keep], r[:, keep]) for g, r in zip(cp.factors, ref_F)]
    print(f"  fit {cp.fit:.8f} oracle {info['fit']:.8f} paths {cp.solve_path_max} / {info['solve_path_max']} "
          f"factor rel errs {['%.1e' % e for e in errs]}", flush=True)
    assert info["solve_path_max"] == 1 and cp.solve_path_max == 1
    assert abs(cp.fit - info["fit"]) <= 1e-4
    for g in cp.factors:
        assert not g.cpu().numpy()[:, 3].any()
    assert max(errs) < 1e-3


def test_cp_als_rank_bounds():
    """Rank 129 exceeds the device solve's bound of 128: refused with Unsupported
    before any launch, with and without the orthogonal constraint."""
    print("[case] cp_als rank bounds 130^3", flush=True)
    import torch
    X = torch.zeros((130, 130, 130), dtype=torch.float32, device="cuda")
    init = [torch.zeros((130, 129), dtype=torch.float32, device="cuda") for _ in range(3)]
    with pytest.raises(Exception, match="bound 128"):
        decomp.cp_als_with_init(X, init, 2, 0.0)
    with pytest.raises(Exception, match="bound 128"):
        decomp.cp_als_with_init(X, init, 2, 0.0, constraints=decomp.CpConstraints.orthogonal_())


def test_cp_als_bf16_start_gate_changes_timing_only(monkeypatch):
    """The start gate of the fused BF16 launch (the MTTKRP workgroups wait for the
    FP64 job's Gram loads) only delays work: factors and fit are the same bit for
    bit with TENRSO_JOB_GATE=1 and =0.  The variable is read when a context is
    created, so each run gets a new stream (and with it a new context)."""
    import torch
    print("[case] cp_als bf16 start gate on/off (96, 64, 64) rank 16", flush=True)
    shape = (96, 64, 64)
    rng = np.random.default_rng(21)
    X = synth.cp_tensor(synth.normal_factors(shape, 16, rng), noise=1e-2, rng=rng)
    init = synth.init_factors(shape, 16, seed=6)
    Xd = dev(X)
    Fd = [dev(f) for f in init]
    runs = []
    for gate in ("1", "0"):
        monkeypatch.setenv("TENRSO_JOB_GATE", gate)
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            cp = decomp.cp_als_with_init(Xd, Fd, 8, 0.0, precision="bf16")
            facs = [f.clone() for f in cp.factors]
        s.synchronize()
        runs.append((cp.fit, facs))
    print(f"  fit gate on {runs[0][0]:.10f} off {runs[1][0]:.10f}", flush=True)
    assert runs[0][0] == runs[1][0]
    for a, b in zip(runs[0][1], runs[1][1]):
        assert torch.equal(a, b)


# (2500, 24, 20), (1500, 40, 36): a mode of more than 1024 rows takes the chunked
# Gram (128-row chunks, their FP64 partials summed in order), at rank 8 and at
# rank 36 (R^2 > 1024: threads form two entries each).  (40, 64, 250),
# (96, 250, 250), (12, 10, 64, 250), (300, 250): a ragged last mode, which FP32
# runs on copies padded to a multiple of 32 (and BF16 on its own padded copies),
# in 3-, 4- and 2-way tensors (the 2-way one pads its transposed copy too).  The
# 4-way case runs at rank 4, where eight sweeps fit to 0.994: at rank 6 that shape
# fits only 0.77 after eight sweeps, and BF16's rounding of X then moves the factors
# by 3e-4 relative whether the last mode is ragged (250) or not (248), which says
# more about the conditioning than about the padded copies.
ALS_NWAY = [((64, 48), 5), ((96, 32), 8), ((16, 20, 12, 32), 6), ((6, 5, 8, 7, 32), 3), ((4, 5, 3, 6, 4, 32), 3),
            ((30, 20, 25), 1), ((2500, 24, 20), 8), ((1500, 40, 36), 36), ((40, 64, 250), 8), ((96, 250, 250), 8),
            ((12, 10, 64, 250), 4), ((300, 250), 7)]


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("shape,rank", ALS_NWAY)
def test_cp_als_nway_vs_oracle(shape, rank, precision):
    """cp_als on 2-, 4-, 5- and 6-way tensors and at rank 1 (the ALS job loads the Grams
    of up to four modes per round trip, so five and more modes take a second group;
    a 2-way Gamma is a single Gram).  FP32: fit history within 1e-4 of the FP64
    oracle and factors within 1e-3.  BF16 contracts bf16(X) in its leading modes and
    X itself in the last one, so it is checked against the FP64 oracle of that
    computation (x_lead = bf16(X)): factors within 1e-3, fit history within 1e-4.
    Both precisions: the reported fit is the fit of the returned factors to X
    (FP64 oracle.fit) up to the fit identity's cancellation.  The identity
    (cp/helpers.rs:513-530) takes the residual from ||X||^2 + ||Xhat||^2 - 2<X, Xhat>;
    an error e ||X||^2 in <X, Xhat> moves the fit by e / (1 - fit).  <X, Xhat> comes
    from the fp32 MFMA MTTKRP of the last mode (split partials summed in FP64),
    whose e is measured at 0.8e-8 to 3.1e-8 here, so the bar is
    1e-6 + 1e-7 / (1 - fit): 2.6e-5 at these tensors' fit of 0.996 (measured up to
    7.6e-6), 1.3e-6 at the 256^3 workload's 0.60 (measured 3.8e-10)."""
    print(f"[case] cp_als {len(shape)}-way {shape} rank {rank} {precision}", flush=True)
    rng = np.random.default_rng(len(shape) * 7 + rank)
    X = synth.cp_tensor(synth.normal_factors(shape, rank, rng), noise=1e-2, rng=rng)
    init = synth.init_factors(shape, rank, seed=3)
    cp = decomp.cp_als_with_init(dev(X), [dev(f) for f in init], 8, 0.0, precision=precision)
    lead = None if precision == "fp32" else bf16_round(X)
    ref_F, info = oracle.cp_als(X, init, 8, 0.0, x_lead=lead)
    got = [g.cpu().numpy() for g in cp.factors]
    errs = [rel(g, r) for g, r in zip(got, ref_F)]
    fit_exact = oracle.fit(X, got)
    dh = np.abs(np.array(cp.convergence.fit_history) - info["fit_history"]).max()
    print(f"  fit {cp.fit:.10f} oracle {info['fit']:.10f} fit(X, factors) {fit_exact:.10f} iters "
          f"{cp.iters}/{info['iters']} factor rel errs {['%.1e' % e for e in errs]} hist {dh:.1e}", flush=True)
    assert cp.iters == info["iters"]
    assert max(errs) < 1e-3
    assert abs(cp.fit - fit_exact) <= 1e-6 + 1e-7 / (1.0 - fit_exact)
    assert dh <= 1e-4
