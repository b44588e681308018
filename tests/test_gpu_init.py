"""GPU parity of the SVD-based initialisations (InitStrategy::Svd / Nnsvd /
LeverageScore, tenrso-decomp cp/helpers.rs:232-373) against the FP64 oracle, whose
SVD is numpy's.  Both sides turn every u_r so its first entry of largest magnitude
is positive.  Bar: each column within 1e-5 of the oracle relative to the column's
largest entry (the device uses the eigenpairs of X_(k) X_(k)^T in FP64; the data
have well-separated leading singular values).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import oracle  # noqa: E402
from tenrso_amd import decomp, synth  # noqa: E402
from tenrso_amd.decomp import InitStrategy  # noqa: E402


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


KINDS = [(InitStrategy.Svd, "svd"), (InitStrategy.Nnsvd, "nnsvd"), (InitStrategy.LeverageScore, "leverage_score")]


def model(shape, rank, seed=0):
    """CP tensor with distinct column weights (well-separated singular values) + noise 1e-3."""
    rng = np.random.default_rng(seed)
    F = synth.normal_factors(shape, rank, rng, normalize=True)
    w = np.linspace(3.0, 1.0, rank)
    return synth.cp_tensor(F, weights=w, noise=1e-3, rng=rng)


@pytest.mark.parametrize("init,kind", KINDS)
@pytest.mark.parametrize("shape,rank", [((40, 36, 30), 5), ((24, 20, 18, 16), 4), ((96, 64, 33), 8),
                                        ((80, 76, 74), 72)])
def test_svd_init_vs_oracle(init, kind, shape, rank):
    print(f"[case] {kind} init {shape} rank {rank}", flush=True)
    X = model(shape, rank)
    got = decomp.initialize_factors(shape, rank, init, seed=3, tensor=dev(X))
    want = oracle.svd_init(X.astype(np.float64), rank, kind)
    for k, (g, w) in enumerate(zip(got, want)):
        assert g.shape == (shape[k], rank)
        for r in range(rank):
            scale = max(np.abs(w[:, r]).max(), 1e-30)
            err = np.abs(g[:, r] - w[:, r]).max() / scale
            assert err <= 1e-5, (k, r, err)


@pytest.mark.parametrize("init,kind", KINDS)
def test_svd_init_columns_past_the_rank_of_the_unfolding(init, kind):
    """Mode 0 has 3 rows, so min(R, I, J) = 3 of the 5 columns come from the SVD and
    the rest from N(0, 0.01) draws: as they are for Svd, |.| for Nnsvd, times the
    leverage scores for LeverageScore (helpers.rs:254-262, 301-308, 354-366)."""
    print(f"[case] {kind} init (3, 40, 36) rank 5", flush=True)
    shape, rank = (3, 40, 36), 5
    X = model(shape, 5, seed=1)
    got = decomp.initialize_factors(shape, rank, init, seed=7, tensor=dev(X))
    want = oracle.svd_init(X.astype(np.float64), rank, kind)
    g, w = got[0], want[0]
    for r in range(3):
        assert np.abs(g[:, r] - w[:, r]).max() <= 1e-5 * max(np.abs(w[:, r]).max(), 1e-30)
    tail = g[:, 3:]
    assert np.isfinite(tail).all() and np.abs(tail).max() < 0.1
    if kind == "nnsvd":
        assert (tail >= 0).all()
    # the other modes have I_k > rank: every column from the SVD
    for k in (1, 2):
        assert np.abs(got[k] - want[k]).max() <= 1e-5 * np.abs(want[k]).max()


def test_cp_als_from_svd_init_vs_oracle():
    """cp_als(init=Svd) on the device against the oracle's ALS from the oracle's
    Svd init: same fit history within 1e-4 over 10 iterations."""
    print("[case] cp_als svd init (40, 36, 30) rank 5", flush=True)
    shape, rank = (40, 36, 30), 5
    X = model(shape, rank, seed=2)
    cp = decomp.cp_als(dev(X), rank, 10, 0.0, init=InitStrategy.Svd, seed=0)
    init = [f.astype(np.float32) for f in oracle.svd_init(X.astype(np.float64), rank, "svd")]
    _, info = oracle.cp_als(X, init, 10, 0.0)
    print(f"  gpu fit {cp.fit:.8f} oracle fit {info['fit']:.8f}", flush=True)
    assert np.abs(np.array(cp.convergence.fit_history) - info["fit_history"]).max() <= 1e-4
