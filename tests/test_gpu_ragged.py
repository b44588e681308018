"""GPU parity of the MTTKRP and CP-ALS on ragged shapes: fastest other modes
that no multiple of 32 divides (100, 250, 70, 65, 31), which the MFMA kernels
cover with a partly filled last window whose loads past the mode's end are
masked (vector loads where the mode is a multiple of the load vector,
element-wise otherwise).  Bars as for the aligned shapes: FP32 MTTKRP within
1e-5 (relative Frobenius) of the FP64 oracle in every mode; BF16 within 2e-5
of the FP64 MTTKRP of the bf16-rounded tensor; CP-ALS fit history within 1e-4
of the FP64 oracle run from the same initial factors (BF16: of the FP64 model
of its computation, leading modes on bf16(X)), same iterations.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import oracle  # noqa: E402
from tenrso_amd import decomp, kernels, synth  # noqa: E402

RAGGED = [((100, 100, 100), 16), ((250, 250, 250), 64), ((33, 50, 70), 7), ((17, 31, 65), 5),
          ((100, 36, 70), 32), ((40, 30, 20, 10), 6)]


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def bf16_round(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.mark.parametrize("shape,rank", RAGGED)
def test_ragged_mttkrp_fp32_and_bf16(shape, rank):
    print(f"[case] ragged mttkrp {shape} rank {rank}", flush=True)
    rng = np.random.default_rng(sum(shape) * 3 + rank)
    X = rng.standard_normal(shape).astype(np.float32)
    F = [rng.standard_normal((n, rank)).astype(np.float32) for n in shape]
    Xd, Fd = dev(X), [dev(f) for f in F]
    Xb = kernels.to_bf16(Xd)
    X64, Xr64 = X.astype(np.float64), bf16_round(X).astype(np.float64)
    F64 = [f.astype(np.float64) for f in F]
    for mode in range(len(shape)):
        ref = oracle.mttkrp_fast(X64, F64, mode)
        got = kernels.mttkrp(Xd, Fd, mode).cpu().numpy()
        e = rel(got, ref)
        gb = kernels.mttkrp_bf16(Xb, Fd, mode).cpu().numpy()
        eb = rel(gb, oracle.mttkrp_fast(Xr64, F64, mode))
        print(f"  mode {mode}: fp32 {e:.2e}, bf16 vs fp64 of bf16(X) {eb:.2e}", flush=True)
        assert e <= 1e-5
        assert eb <= 2e-5
    # a view at a 4-byte offset (no 16-byte alignment) gives the same values
    buf = torch.zeros(X.size + 1, dtype=torch.float32, device="cuda")
    buf[1:] = Xd.reshape(-1)
    Xo = buf[1:].view(shape)
    for mode in range(len(shape)):
        assert rel(kernels.mttkrp(Xo, Fd, mode).cpu().numpy(), oracle.mttkrp_fast(X64, F64, mode)) <= 1e-5


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("shape,rank", [((100, 100, 100), 16), ((33, 50, 70), 7), ((17, 31, 65), 5)])
def test_ragged_cp_als(shape, rank, precision):
    print(f"[case] ragged cp_als {shape} rank {rank} {precision}", flush=True)
    rng = np.random.default_rng(rank * 11)
    X = synth.cp_tensor(synth.normal_factors(shape, rank, rng), noise=1e-2, rng=rng)
    init = synth.init_factors(shape, rank, seed=4)
    cp = decomp.cp_als_with_init(dev(X), [dev(f) for f in init], 8, 0.0, precision=precision)
    lead = None if precision == "fp32" else bf16_round(X)
    ref_F, info = oracle.cp_als(X, init, 8, 0.0, x_lead=lead)
    got = [g.cpu().numpy() for g in cp.factors]
    errs = [rel(g, r) for g, r in zip(got, ref_F)]
    dh = np.abs(np.array(cp.convergence.fit_history) - info["fit_history"]).max()
    fit_exact = oracle.fit(X, got)
    print(f"  fit {cp.fit:.8f} oracle {info['fit']:.8f} factors {['%.1e' % e for e in errs]} hist {dh:.1e}",
          flush=True)
    assert cp.iters == info["iters"]
    assert max(errs) < 1e-3
    assert dh <= 1e-4
    assert abs(cp.fit - fit_exact) <= 1e-6 + 1e-7 / (1.0 - fit_exact)
