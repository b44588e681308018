"""CPU-side checks of the C ABI: the library loads, exports every declared symbol,
and reproduces the reference's argument validation (no device work)."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "tenrso_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tenrso_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from tenrso_amd import _lib
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_mttkrp_validation_messages():
    # same checks and messages as mttkrp.rs:85-122 (tests at mttkrp.rs:311-340)
    from tenrso_amd import kernels
    X = np.ones((2, 3), np.float32)
    u1, u2 = np.ones((2, 1), np.float32), np.ones((3, 1), np.float32)
    with pytest.raises(ValueError, match="Mode 5 out of bounds for tensor with rank 2"):
        kernels.mttkrp(X, [u1, u2], 5)
    with pytest.raises(ValueError, match=r"Number of factor matrices \(2\) must match tensor rank \(3\)"):
        kernels.mttkrp(np.ones((2, 3, 4), np.float32), [u1, u2], 0)
    with pytest.raises(ValueError, match="Factor matrix 1 has 1 columns, expected 2"):
        kernels.mttkrp(X, [np.ones((2, 2), np.float32), u2], 0)
    with pytest.raises(ValueError, match=r"Factor matrix 1 has 4 rows, expected 3 \(tensor mode-1 size\)"):
        kernels.mttkrp(X, [u1, np.ones((4, 1), np.float32)], 0)
    with pytest.raises(ValueError, match="Need at least 2 factor matrices for MTTKRP"):
        kernels.mttkrp(np.ones((3,), np.float32), [np.ones((3, 1), np.float32)], 0)
    with pytest.raises(ValueError, match="Tile size must be greater than 0"):
        kernels.mttkrp_blocked(X, [u1, u2], 0, 0)


def test_khatri_rao_and_cp_validation():
    from tenrso_amd import decomp, kernels
    with pytest.raises(ValueError, match="Number of columns must match: A has 3 columns, B has 2 columns"):
        kernels.khatri_rao(np.ones((1, 3)), np.ones((1, 2)))
    X = np.ones((4, 5, 6), np.float32)
    with pytest.raises(decomp.InvalidRank, match="Invalid rank: 0"):
        decomp.cp_als(X, 0, 5, 1e-4)
    with pytest.raises(decomp.InvalidRank, match="Invalid rank: 5"):
        decomp.cp_als(X, 5, 5, 1e-4)
    with pytest.raises(decomp.InvalidTolerance, match="Invalid tolerance: 1.0"):
        decomp.cp_als(X, 2, 5, 1.0)
    with pytest.raises(decomp.InvalidRegularization, match="L1 lambda must be >= 0"):
        decomp.cp_als_constrained(X, 2, 5, 1e-4, constraints=decomp.CpConstraints.l1_regularized(-1.0))
    with pytest.raises(decomp.InvalidRegularization, match=r"ElasticNet alpha must be in \[0, 1\]"):
        decomp.cp_als_constrained(X, 2, 5, 1e-4, constraints=decomp.CpConstraints.elastic_net(0.1, 2.0))


def test_constraints_effective_params():
    # cp/tests.rs:493-519
    from tenrso_amd.decomp import CpConstraints
    assert CpConstraints.l2_regularized(0.1).effective_l2() == 0.1
    assert CpConstraints.l1_regularized(0.2).effective_l1() == 0.2
    en = CpConstraints.elastic_net(1.0, 0.3)
    assert abs(en.effective_l1() - 0.3) < 1e-12 and abs(en.effective_l2() - 0.7) < 1e-12
    assert CpConstraints.tikhonov(0.5, 0).effective_l2() == 0.5
    assert CpConstraints.tikhonov(0.5, 1).effective_l2() == 0.0
    assert CpConstraints(l2_reg=0.25).effective_l2() == 0.25
