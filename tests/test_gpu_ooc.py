"""GPU tests of the out-of-core (streaming) MTTKRP: a host tensor streamed through
a device-memory budget in mode-0 slabs (tenrso_mttkrp_f32_streaming; the
reference's tenrso-ooc example mttkrp_streaming.rs:66-110 describes the
algorithm and calls the in-memory mttkrp).  Every mode against the in-core
device MTTKRP (relative Frobenius <= 1e-5) and the FP64 oracle (<= 1e-5), for
pinned and pageable inputs, with budgets that force many slabs, one slab, and
a ragged last slab; a budget too small for two slabs is refused."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import oracle  # noqa: E402
from tenrso_amd import kernels  # noqa: E402

MiB = 1 << 20


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.mark.parametrize("shape,rank,budget_mib", [((100, 64, 96), 16, 4), ((37, 50, 70), 8, 2),
                                                   ((20, 24, 16, 32), 6, 2), ((64, 32, 32), 32, 256)])
def test_streaming_matches_in_core_and_oracle(shape, rank, budget_mib):
    print(f"[case] streaming mttkrp {shape} rank {rank} budget {budget_mib} MiB", flush=True)
    rng = np.random.default_rng(sum(shape))
    X = rng.standard_normal(shape).astype(np.float32)
    F = [rng.standard_normal((n, rank)).astype(np.float32) for n in shape]
    Xd = torch.from_numpy(X).cuda()
    Fd = [torch.from_numpy(f).cuda() for f in F]
    Xp = torch.from_numpy(X).pin_memory()
    for mode in range(len(shape)):
        ref = oracle.mttkrp_fast(X.astype(np.float64), [f.astype(np.float64) for f in F], mode)
        core = kernels.mttkrp(Xd, Fd, mode).cpu().numpy()
        got, st = kernels.mttkrp_streaming(X, F, mode, budget_mib * MiB, return_stats=True)
        gotp, stp = kernels.mttkrp_streaming(Xp, F, mode, budget_mib * MiB, return_stats=True)
        print(f"  mode {mode}: slabs {st['slabs']} x {st['slab_rows']} rows, vs core {rel(got, core):.1e}, "
              f"vs fp64 {rel(got, ref):.1e}; pinned {stp['pinned']}", flush=True)
        assert not st["pinned"] and stp["pinned"]
        assert rel(got, ref) <= 1e-5 and rel(got, core) <= 1e-5
        assert np.array_equal(got, gotp.numpy())


def test_streaming_512cube_256mib_budget():
    """512^3 rank 32 (512 MiB tensor) through a 256 MiB budget, all modes, against
    the in-core device MTTKRP; reports the host-to-device rate."""
    print("[case] streaming mttkrp 512^3 rank 32, 256 MiB budget", flush=True)
    shape, R = (512, 512, 512), 32
    g = torch.Generator().manual_seed(7)
    Xp = torch.randn(shape, generator=g, dtype=torch.float32).pin_memory()
    F = [torch.randn((n, R), generator=g, dtype=torch.float32).numpy() for n in shape]
    Xd = Xp.cuda()
    Fd = [torch.from_numpy(f).cuda() for f in F]
    for mode in range(3):
        core = kernels.mttkrp(Xd, Fd, mode).cpu().numpy()
        got, st = kernels.mttkrp_streaming(Xp, F, mode, 256 * MiB, return_stats=True)
        print(f"  mode {mode}: {st['slabs']} slabs of {st['slab_rows']} rows, {st['seconds'] * 1e3:.1f} ms, "
              f"H2D {st['h2d_GBs']:.1f} GB/s, vs in-core {rel(got, core):.1e}", flush=True)
        assert st["slabs"] >= 3
        assert rel(got, core) <= 1e-5


def test_streaming_refuses_tiny_budget():
    X = np.ones((8, 256, 256), np.float32)
    F = [np.ones((n, 4), np.float32) for n in X.shape]
    with pytest.raises(kernels._lib.TenrsoError, match="device budget"):
        kernels.mttkrp_streaming(X, F, 0, 64 * 1024)
    with pytest.raises(TypeError):
        kernels.mttkrp_streaming(torch.from_numpy(X).cuda(), F, 0, 64 * MiB)
