"""CPU tests of the COO / CSF format mirror (tenrso_amd.sparse) against the
reference's own format tests (tenrso-sparse/src/coo.rs:350-470, csf.rs:501-682)
and brute force.  The HIP library is not needed here."""
import numpy as np
import pytest

from tenrso_amd.sparse import CooError, CooTensor, CsfTensor


def test_coo_new_validation():
    # coo.rs:101-149
    with pytest.raises(CooError, match="Length mismatch"):
        CooTensor([[0], [0]], [1.0, 2.0], [2, 2])
    with pytest.raises(CooError, match="Shape cannot contain zeros"):
        CooTensor(np.zeros((2, 0)), [], [2, 0])
    with pytest.raises(CooError, match="Index out of bounds"):
        CooTensor([[0, 2]], [1.0], [2, 2])
    c = CooTensor([[0, 1, 2], [1, 0, 1], [2, 2, 0]], [1.0, 2.0, 3.0], [3, 3, 3])  # coo.rs test_coo_3d
    assert c.rank() == 3 and c.nnz() == 3


def test_coo_to_dense_from_dense_kat():
    # coo.rs to_dense / from_dense tests: [0,1]=1, [1,0]=2, [2,2]=3
    c = CooTensor([[0, 1], [1, 0], [2, 2]], [1.0, 2.0, 3.0], [3, 3])
    d = c.to_dense()
    assert d[0, 1] == 1.0 and d[1, 0] == 2.0 and d[2, 2] == 3.0 and d[0, 0] == 0.0
    dense = np.array([0.0, 1.0, 0.0, 2.0, 0.0, 0.0, 0.0, 0.0, 3.0]).reshape(3, 3)
    assert CooTensor.from_dense(dense, 1e-10).nnz() == 3
    # to_dense keeps the last of duplicate coordinates (coo.rs:279-283)
    dup = CooTensor([[0, 0], [0, 0]], [1.0, 2.0], [2, 2])
    assert dup.to_dense()[0, 0] == 2.0


def test_coo_deduplicate_and_sort_kat():
    # coo.rs test_coo_deduplicate: values summed -> [3, 7, 5]
    c = CooTensor([[0, 0], [0, 0], [1, 1], [1, 1], [2, 2]], [1.0, 2.0, 3.0, 4.0, 5.0], [3, 3])
    c.deduplicate()
    assert c.nnz() == 3 and c.values().tolist() == [3.0, 7.0, 5.0]
    # coo.rs test_coo_sort: [0,1], [1,0], [2,0]
    s = CooTensor([[2, 0], [0, 1], [1, 0]], [1.0, 2.0, 3.0], [3, 3])
    s.sort()
    assert s.indices().T.tolist() == [[0, 1], [1, 0], [2, 0]]
    assert s.values().tolist() == [2.0, 3.0, 1.0]


def test_csf_reference_kats():
    # csf.rs:502-560
    coo = CooTensor.zeros([3, 4, 5])
    coo.push([0, 1, 2], 5.0)
    coo.push([0, 1, 3], 6.0)
    coo.push([1, 2, 3], 7.0)
    csf = CsfTensor.from_coo(coo, [0, 1, 2])
    assert csf.shape() == [3, 4, 5] and csf.nnz() == 3 and csf.mode_order() == [0, 1, 2]
    with pytest.raises(ValueError):
        CsfTensor.from_coo(CooTensor.zeros([3, 3, 3]), [0, 1])
    with pytest.raises(ValueError):
        CsfTensor.from_coo(CooTensor.zeros([3, 3, 3]), [0, 0, 1])
    with pytest.raises(ValueError):
        CsfTensor.from_coo(CooTensor.zeros([3, 3, 3]), [0, 1, 3])
    assert CsfTensor.from_coo(CooTensor.zeros([3, 3, 3]), [0, 1, 2]).nnz() == 0
    # standard fiber pointers: root (0) has one child fiber (0,1) with two leaves
    assert csf.fids[0].tolist() == [0, 1]
    assert csf.fptr[0].tolist() == [0, 1, 2]
    assert csf.fids[1].tolist() == [1, 2]
    assert csf.fptr[1].tolist() == [0, 2, 3]
    assert csf.fids[2].tolist() == [2, 3, 3]


def test_csf_roundtrip_multi_nonzero_fibers():
    # the case the reference iterator mis-decodes (fibers with >1 nonzero; in the
    # csf.rs doc example, nonzero (1,2,3) came back as (0,2,3)); here it round-trips
    rng = np.random.default_rng(0)
    shape = (6, 5, 7, 3)
    lin = rng.choice(np.prod(shape), size=150, replace=False)
    idx = np.array(np.unravel_index(lin, shape))
    vals = rng.random(150)
    coo = CooTensor(idx, vals, shape)
    for order in ([0, 1, 2, 3], [3, 2, 1, 0], [1, 0, 3, 2], [2, 0, 1, 3]):
        csf = CsfTensor.from_coo(coo, order)
        back = csf.to_coo()
        assert np.array_equal(back.to_dense(), coo.to_dense())
        # fptr is monotone and the leaf level covers every nonzero exactly once
        for l in range(len(shape) - 1):
            assert (np.diff(csf.fptr[l]) >= 1).all()
            assert csf.fptr[l][-1] == csf.fids[l + 1].size
        assert csf.fids[-1].size == 150
