"""Deterministic synthetic workloads for the CP-ALS / MTTKRP path.

All draws use numpy's PCG64 (``numpy.random.default_rng(seed)``), so the CPU
oracle and the GPU run see byte-identical inputs.  The reference draws its
tensors and initial factors from an unseeded ``thread_rng``
(tenrso-decomp/src/cp/helpers.rs:212-231, tenrso-core/src/dense/creation.rs:33-91),
so no reference stream exists to reproduce; the shapes and distributions follow
the workloads named below.

Workloads
---------
dense_64_r8      64^3 f32, rank-8 CP model (factors N(0,1), seed 0) + N(0, 0.01^2) noise;
                 init N(0,1) from seed 1; cp_als rank 8, 20 iters, tol 1e-4.
dense_256_r64    256^3 f32 (64 MiB), rank-64 model: factor columns N(0,1) normalised,
                 weights U(0.5, 2), noise sigma 1e-3; init N(0,1) (seed 1).
congruent_*      256^3 and 512x512x128, rank 64, factor columns with pairwise
                 congruence 0.9 (a_r = sqrt(.9) u + sqrt(.1) v_r, u, v_r orthonormal),
                 noise sigma 1e-4.
dense4_96_r32    96^4 f32 (~340 MB), rank-32 model + noise sigma 1e-3.
sparse_zipf      3-way COO 10000x8000x6000, Zipf(s=1.2) per-mode indices over a seeded
                 permutation of each mode, duplicates removed, values U(0,1).
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "cp_tensor", "normal_factors", "congruent_factors", "dense_64_r8", "dense_256_r64", "congruent",
    "dense4_96_r32", "sparse_zipf", "init_factors",
]


def cp_tensor(factors, weights=None, noise=0.0, rng=None, dtype=np.float32):
    """X = sum_r w_r a_r o b_r o ... + noise * N(0,1), built in float64, returned as dtype."""
    F = [np.asarray(f, dtype=np.float64) for f in factors]
    R = F[0].shape[1]
    w = np.ones(R) if weights is None else np.asarray(weights, dtype=np.float64)
    shape = [f.shape[0] for f in F]
    # X_(0) = (A0 * w) @ KR(A1, ..., A_{N-1})^T, KR in forward order
    kr = F[1] if len(F) > 1 else np.ones((1, R))
    for f in F[2:]:
        kr = (kr[:, None, :] * f[None, :, :]).reshape(-1, R)
    X = ((F[0] * w) @ kr.T).reshape(shape)
    if noise:
        if rng is None:
            raise ValueError("noise needs an rng")
        X += noise * rng.standard_normal(shape)
    return np.ascontiguousarray(X.astype(dtype))


def normal_factors(shape, rank, rng, normalize=False):
    F = [rng.standard_normal((n, rank)) for n in shape]
    if normalize:
        F = [f / np.linalg.norm(f, axis=0, keepdims=True) for f in F]
    return F


def congruent_factors(shape, rank, rng, congruence=0.9):
    """Columns with pairwise inner product `congruence` and unit norm (needs I_k > rank)."""
    out = []
    for n in shape:
        if n < rank + 1:
            raise ValueError(f"mode size {n} must exceed rank {rank}")
        Qm, _ = np.linalg.qr(rng.standard_normal((n, rank + 1)))
        u, V = Qm[:, :1], Qm[:, 1:]
        out.append(np.sqrt(congruence) * u + np.sqrt(1.0 - congruence) * V)
    return out


def init_factors(shape, rank, seed=1, dtype=np.float32):
    """Initial factors N(0,1), one (I_k x R) block per mode, drawn in mode order."""
    rng = np.random.default_rng(seed)
    return [rng.standard_normal((n, rank)).astype(dtype) for n in shape]


def dense_64_r8():
    rng = np.random.default_rng(0)
    shape = (64, 64, 64)
    F = normal_factors(shape, 8, rng)
    X = cp_tensor(F, noise=0.01, rng=rng)
    return X, F, init_factors(shape, 8, seed=1)


def dense_256_r64(shape=(256, 256, 256), rank=64, noise=1e-3, seed=0, init_seed=1):
    rng = np.random.default_rng(seed)
    F = normal_factors(shape, rank, rng, normalize=True)
    w = rng.uniform(0.5, 2.0, size=rank)
    X = cp_tensor(F, w, noise=noise, rng=rng)
    return X, F, w, init_factors(shape, rank, seed=init_seed)


def congruent(shape=(256, 256, 256), rank=64, noise=1e-4, seed=0, init_seed=1, congruence=0.9):
    rng = np.random.default_rng(seed)
    F = congruent_factors(shape, rank, rng, congruence)
    w = rng.uniform(0.5, 2.0, size=rank)
    X = cp_tensor(F, w, noise=noise, rng=rng)
    return X, F, w, init_factors(shape, rank, seed=init_seed)


def dense4_96_r32(shape=(96, 96, 96, 96), rank=32, noise=1e-3, seed=0, init_seed=1):
    rng = np.random.default_rng(seed)
    F = normal_factors(shape, rank, rng, normalize=True)
    w = rng.uniform(0.5, 2.0, size=rank)
    X = cp_tensor(F, w, noise=noise, rng=rng)
    return X, F, w, init_factors(shape, rank, seed=init_seed)


def _zipf_index(rng, n, size, s):
    """Zipf(s) ranks truncated to [0, n) mapped through a fixed permutation of the mode."""
    # inverse-CDF sampling of a truncated Zipf over ranks 1..n (exact, vectorised)
    ranks = np.arange(1, n + 1, dtype=np.float64)
    cdf = np.cumsum(ranks ** (-s))
    cdf /= cdf[-1]
    u = rng.random(size)
    return np.searchsorted(cdf, u, side="right").clip(0, n - 1)


def sparse_zipf(shape=(10000, 8000, 6000), nnz=20_000_000, s=1.2, seed=0):
    """Unique COO tensor: returns (indices (ndim, nnz) int32, values (nnz,) float32).

    Draws Zipf(s) indices per mode over a seeded permutation of that mode in
    batches until `nnz` unique coordinates exist (duplicates are dropped, first
    occurrence kept), then draws values U(0,1).  Output order is by linear index.
    """
    rng = np.random.default_rng(seed)
    perms = [rng.permutation(n) for n in shape]
    strides = np.ones(len(shape), dtype=np.int64)
    for k in range(len(shape) - 2, -1, -1):
        strides[k] = strides[k + 1] * shape[k + 1]
    keys = np.empty(0, dtype=np.int64)
    batch = max(nnz, 1024)
    while keys.size < nnz:
        lin = np.zeros(batch, dtype=np.int64)
        for k, n in enumerate(shape):
            lin += perms[k][_zipf_index(rng, n, batch, s)].astype(np.int64) * strides[k]
        keys = np.unique(np.concatenate([keys, lin]))
    if keys.size > nnz:
        keep = np.sort(rng.choice(keys.size, size=nnz, replace=False))
        keys = keys[keep]
    idx = np.empty((len(shape), nnz), dtype=np.int32)
    rem = keys.copy()
    for k in range(len(shape) - 1, -1, -1):
        idx[k] = (rem % shape[k]).astype(np.int32)
        rem //= shape[k]
    vals = rng.random(nnz).astype(np.float32)
    return idx, vals
