"""tenrso-sparse mirror for the CP path: COO / CSF formats and sparse MTTKRP on gfx950.

``CooTensor`` mirrors tenrso-sparse/src/coo.rs (validation of ``new`` 101-149,
``push`` 213-233, ``sort`` 239-258, ``to_dense`` 267-286 with last-wins on
duplicates, ``from_dense`` 292-317, ``deduplicate`` 320-348 summing).  Indices
are held as an (ndim, nnz) int array instead of a Vec of Vecs.

``CsfTensor.from_coo`` mirrors csf.rs:137-264 (validation and lexicographic sort
in ``mode_order``) but uses standard CSF fiber pointers: ``fptr[l][f]`` ..
``fptr[l][f+1]`` index the children of fiber f at level l+1.  The reference
fills ``fptr`` with nonzero offsets (csf.rs:209,246) while its iterator reads
them as fiber indices, which mis-decodes fibers with more than one nonzero.

``SparseTensor`` is the device form (``tenrso_sp_create``): the nonzeros sorted
per output mode on the GPU, the sparse MTTKRP and sparse CP-ALS.  The reference
has no sparse MTTKRP; its contract is the dense ``mttkrp`` of the tensor the COO
represents with duplicates summed.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import CpOptions, CpResult, check, lib

__all__ = ["CooError", "CooTensor", "CsfTensor", "HiCooError", "HiCooTensor", "SparseTensor"]


class CooError(ValueError):
    pass


class CooTensor:
    """COO sparse tensor (coo.rs:73-81)."""

    def __init__(self, indices, values, shape):
        shape = [int(s) for s in shape]
        idx = np.asarray(indices, dtype=np.int64)
        vals = np.asarray(values)
        if idx.ndim == 1 and idx.size == 0:
            idx = idx.reshape(len(shape), 0)
        if idx.ndim == 2 and idx.shape[0] != len(shape) and idx.shape[1] == len(shape):
            idx = idx.T  # accept (nnz, ndim) as the reference's Vec<Vec<usize>>
        if idx.shape[1] != vals.shape[0]:
            raise CooError(f"Length mismatch: indices {idx.shape[1]}, values {vals.shape[0]}")
        if len(shape) == 0:
            raise CooError("Invalid shape: Shape cannot be empty")
        if 0 in shape:
            raise CooError("Invalid shape: Shape cannot contain zeros")
        if idx.shape[0] != len(shape):
            raise CooError(f"Shape mismatch: indices have {idx.shape[0]} dims, shape has {len(shape)}")
        if idx.size and ((idx < 0).any() or (idx >= np.asarray(shape)[:, None]).any()):
            raise CooError("Index out of bounds")
        self._idx = idx
        self._vals = vals.copy()
        self._shape = shape

    @staticmethod
    def zeros(shape):
        return CooTensor(np.zeros((len(shape), 0), np.int64), np.zeros(0, np.float32), shape)

    def nnz(self):
        return int(self._vals.size)

    def shape(self):
        return list(self._shape)

    def rank(self):
        return len(self._shape)

    def indices(self):
        return self._idx

    def values(self):
        return self._vals

    def density(self):
        return self.nnz() / float(np.prod(self._shape))

    def push(self, index, value):
        """coo.rs:213-233 (no duplicate check)."""
        index = [int(i) for i in index]
        if len(index) != len(self._shape):
            raise CooError(f"Shape mismatch: index has {len(index)} dims, shape has {len(self._shape)}")
        for c, s in zip(index, self._shape):
            if c < 0 or c >= s:
                raise CooError("Index out of bounds")
        self._idx = np.concatenate([self._idx, np.asarray(index, np.int64)[:, None]], axis=1)
        self._vals = np.concatenate([self._vals, np.asarray([value], dtype=self._vals.dtype)])

    def sort(self):
        """coo.rs:239-258: row-major (lexicographic) order, stable."""
        order = np.lexsort(self._idx[::-1]) if self.nnz() else np.zeros(0, np.int64)
        self._idx = self._idx[:, order]
        self._vals = self._vals[order]

    def to_dense(self):
        """coo.rs:267-286: duplicates -> the last one wins."""
        out = np.zeros(self._shape, dtype=self._vals.dtype if self._vals.size else np.float32)
        out[tuple(self._idx)] = self._vals  # numpy assigns in order: last occurrence wins
        return out

    @staticmethod
    def from_dense(dense, threshold=0.0):
        """coo.rs:292-317: entries with |x| > threshold, row-major order."""
        dense = np.asarray(dense)
        idx = np.array(np.nonzero(np.abs(dense) > threshold), dtype=np.int64)
        return CooTensor(idx, dense[tuple(idx)], dense.shape)

    def deduplicate(self):
        """coo.rs:320-348: sort, then sum equal coordinates."""
        if self.nnz() <= 1:
            return
        self.sort()
        keep = np.ones(self.nnz(), dtype=bool)
        keep[1:] = (self._idx[:, 1:] != self._idx[:, :-1]).any(axis=0)
        starts = np.nonzero(keep)[0]
        self._vals = np.add.reduceat(self._vals, starts).astype(self._vals.dtype)
        self._idx = self._idx[:, starts]


class CsfTensor:
    """CSF tensor (csf.rs:92-106) with standard fiber pointers (see module doc)."""

    def __init__(self, shape, mode_order, nnz, fptr, fids, vals):
        self._shape, self._mode_order, self._nnz = list(shape), list(mode_order), int(nnz)
        self.fptr, self.fids, self.vals = fptr, fids, vals

    @staticmethod
    def from_coo(coo: CooTensor, mode_order):
        """csf.rs:137-264 construction (validation, lexicographic sort in mode_order)."""
        nd = coo.rank()
        mode_order = [int(m) for m in mode_order]
        if len(mode_order) != nd:
            raise ValueError(f"Invalid mode order: mode_order length {len(mode_order)} != ndim {nd}")
        if sorted(mode_order) != list(range(nd)):
            raise ValueError(f"Invalid mode order: mode_order {mode_order} is not a permutation of 0..{nd}")
        idx, vals = coo.indices(), coo.values()
        if coo.nnz() == 0:
            return CsfTensor(coo.shape(), mode_order, 0, [np.zeros(1, np.int64)] * nd,
                             [np.zeros(0, np.int32)] * nd, vals.copy())
        keys = [idx[m] for m in mode_order]
        order = np.lexsort(keys[::-1], axis=-1)
        sidx = np.stack([idx[m][order] for m in mode_order])
        svals = vals[order]
        nnz = svals.size
        starts = []
        for l in range(nd):
            if l == nd - 1:
                st = np.arange(nnz)
            else:
                ch = np.ones(nnz, bool)
                ch[1:] = (sidx[: l + 1, 1:] != sidx[: l + 1, :-1]).any(axis=0)
                st = np.nonzero(ch)[0]
            starts.append(st)
        fids = [sidx[l][starts[l]].astype(np.int32) for l in range(nd)]
        fptr = []
        for l in range(nd):
            if l == nd - 1:
                fptr.append(np.zeros(1, np.int64))
            else:
                ptr = np.searchsorted(starts[l + 1], starts[l], side="left")
                fptr.append(np.append(ptr, starts[l + 1].size).astype(np.int64))
        return CsfTensor(coo.shape(), mode_order, nnz, fptr, fids, svals)

    def shape(self):
        return list(self._shape)

    def mode_order(self):
        return list(self._mode_order)

    def nnz(self):
        return self._nnz

    def to_coo(self):
        """Expand the tree back to COO coordinates (original mode positions)."""
        nd = len(self._mode_order)
        if self._nnz == 0:
            return CooTensor.zeros(self._shape)
        # leaf-level coordinates: walk parents bottom-up
        coords = [None] * nd
        owner = np.arange(self._nnz)
        coords[nd - 1] = self.fids[nd - 1]
        for l in range(nd - 2, -1, -1):
            parent = np.searchsorted(self.fptr[l], owner, side="right") - 1
            coords[l] = self.fids[l][parent]
            owner = parent
        idx = np.zeros((nd, self._nnz), np.int64)
        for l, m in enumerate(self._mode_order):
            idx[m] = coords[l]
        return CooTensor(idx, self.vals, self._shape)


class HiCooError(ValueError):
    """HiCooError (hicoo.rs): an invalid block shape."""


class HiCooTensor:
    """HiCOO tensor (hicoo.rs:79-95), built on the device by SparseTensor.hicoo.

    block_coords: (ndim, num_blocks) int64; block_ptrs: num_blocks + 1 (block b
    holds nonzeros block_ptrs[b]..block_ptrs[b+1]); local_coords: (ndim, nnz)
    int32; values: nnz float32.  Nonzeros are in the order of
    HiCooTensor::from_coo (stable sort by block coordinate, duplicates kept)."""

    def __init__(self, shape, block_shape, block_coords, block_ptrs, local_coords, values):
        self._shape, self._block_shape = list(shape), list(block_shape)
        self.block_coords_, self.block_ptrs_ = block_coords, block_ptrs
        self.local_coords_, self.values_ = local_coords, values

    def shape(self):
        return list(self._shape)

    def block_shape(self):
        return list(self._block_shape)

    def ndim(self):
        return len(self._shape)

    def nnz(self):
        return int(self.values_.size)

    def num_blocks(self):
        return int(self.block_ptrs_.size) - 1

    def block_coords(self):
        return self.block_coords_

    def block_ptrs(self):
        return self.block_ptrs_

    def local_coords(self):
        return self.local_coords_

    def values(self):
        return self.values_

    def density(self):
        """hicoo.rs density: nnz over the number of elements of the shape."""
        return self.nnz() / float(np.prod(self._shape))

    def global_indices(self):
        """(ndim, nnz) int64: block coordinate * block_shape + local coordinate."""
        nb = self.num_blocks()
        if nb == 0:
            return np.zeros((self.ndim(), 0), np.int64)
        owner = np.repeat(np.arange(nb), np.diff(self.block_ptrs_))
        bs = np.asarray(self._block_shape, np.int64)[:, None]
        return self.block_coords_[:, owner] * bs + self.local_coords_.astype(np.int64)

    def iter(self):
        """HiCooIterator (hicoo.rs:290-292): (global indices, value) in HiCOO order."""
        g = self.global_indices()
        for e in range(self.nnz()):
            yield [int(v) for v in g[:, e]], float(self.values_[e])

    def to_coo(self):
        """hicoo.rs:313-330: back to COO, in HiCOO order."""
        if self.nnz() == 0:
            return CooTensor.zeros(self._shape)
        return CooTensor(self.global_indices(), self.values_, self._shape)

    def to_dense(self):
        return self.to_coo().to_dense()


class SparseTensor:
    """A COO tensor uploaded to the GPU and sorted per output mode (tenrso_sp_create)."""

    def __init__(self, indices, values, shape, device=0):
        self.shape = tuple(int(s) for s in shape)
        idx = np.ascontiguousarray(np.asarray(indices).reshape(len(self.shape), -1), dtype=np.int32)
        vals = np.ascontiguousarray(values, dtype=np.float32)
        self.nnz = int(vals.size)
        self.ctx = _lib.context_for(device)
        self.ptr = C.c_void_p()
        shp = (C.c_int64 * len(self.shape))(*self.shape)
        rc = lib().tenrso_sp_create(self.ctx.ptr, len(self.shape), shp, self.nnz, idx.ctypes.data,
                                    vals.ctypes.data, 1, C.byref(self.ptr))
        if rc == _lib.ERR_INVALID_ARGUMENT:
            raise CooError(_lib.last_error())
        check(rc)

    @staticmethod
    def from_coo(coo: CooTensor, device=0):
        return SparseTensor(coo.indices(), coo.values(), coo.shape(), device)

    def close(self):
        if self.ptr:
            lib().tenrso_sp_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mode_arrays(self, mode):
        """(rowptr, row, others, values) of the mode-sorted nonzeros (host copies)."""
        nd = len(self.shape)
        rowptr = np.zeros(self.shape[mode] + 1, np.int64)
        row = np.zeros(self.nnz, np.int32)
        oth = np.zeros((nd - 1, self.nnz), np.int32)
        vals = np.zeros(self.nnz, np.float32)
        check(lib().tenrso_sp_mode_arrays(self.ctx.ptr, self.ptr, mode, rowptr.ctypes.data, row.ctypes.data,
                                          oth.ctypes.data, vals.ctypes.data))
        return rowptr, row, oth, vals

    def hicoo(self, block_shape):
        """HiCooTensor::from_coo (hicoo.rs:121-205) of the uploaded COO, on the device."""
        nd = len(self.shape)
        block_shape = [int(b) for b in block_shape]
        if len(block_shape) != nd:
            raise HiCooError(f"Invalid block shape: block_shape length {len(block_shape)} != ndim {nd}")
        bsa = (C.c_int64 * nd)(*block_shape)
        nb = C.c_int64(0)
        rc = lib().tenrso_sp_hicoo(self.ctx.ptr, self.ptr, bsa, C.byref(nb), None, None, None, None)
        if rc == _lib.ERR_INVALID_ARGUMENT:
            raise HiCooError(_lib.last_error())
        check(rc)
        B = int(nb.value)
        bc = np.zeros((nd, B), np.int64)
        bp = np.zeros(B + 1, np.int64)
        loc = np.zeros((nd, self.nnz), np.int32)
        vals = np.zeros(self.nnz, np.float32)
        check(lib().tenrso_sp_hicoo(self.ctx.ptr, self.ptr, bsa, C.byref(nb), bc.ctypes.data, bp.ctypes.data,
                                    loc.ctypes.data, vals.ctypes.data))
        return HiCooTensor(self.shape, block_shape, bc, bp, loc, vals)

    def fibres(self, mode):
        """Number of compressed fibres of the mode (distinct (i_mode, i_o1, .., i_o(N-2))
        prefixes), the level above the leaves that the sparse MTTKRP reads."""
        n = C.c_int64()
        check(lib().tenrso_sp_fibres(self.ptr, int(mode), C.byref(n)))
        return int(n.value)

    def csf(self, mode):
        """CsfTensor of mode order [mode, others ascending], built from the device sort."""
        nd = len(self.shape)
        nfib = np.zeros(nd, np.int64)
        check(lib().tenrso_sp_csf(self.ctx.ptr, self.ptr, mode, nfib.ctypes.data, None, None, None))
        fids = [np.zeros(int(n), np.int32) for n in nfib]
        fptr = [np.zeros(int(nfib[l + 1]) and int(nfib[l]) + 1 or 1, np.int64) if l < nd - 1 else np.zeros(1, np.int64)
                for l in range(nd)]
        for l in range(nd - 1):
            fptr[l] = np.zeros(int(nfib[l]) + 1, np.int64)
        vals = np.zeros(self.nnz, np.float32)
        fp = (C.c_void_p * nd)(*[a.ctypes.data for a in fptr])
        fi = (C.c_void_p * nd)(*[a.ctypes.data for a in fids])
        check(lib().tenrso_sp_csf(self.ctx.ptr, self.ptr, mode, nfib.ctypes.data, fp, fi, vals.ctypes.data))
        order = [mode] + [m for m in range(nd) if m != mode]
        return CsfTensor(self.shape, order, self.nnz, fptr, fids, vals)

    def mttkrp(self, factors, mode):
        """Sparse MTTKRP; factors are float32 CUDA tensors, returns a CUDA tensor."""
        import torch
        from .kernels import _dev_f32
        # factors of any float dtype or device are converted to float32 on the
        # device the tensor lives on (never reinterpreted)
        dev = torch.device("cuda", self.ctx.device)
        F = [_dev_f32(f, dev) for f in factors]
        if len(F) != len(self.shape):
            raise ValueError(f"Number of factor matrices ({len(F)}) must match tensor rank ({len(self.shape)})")
        R = int(F[0].shape[1])
        for i, f in enumerate(F):
            if int(f.shape[1]) != R:
                raise ValueError(f"Factor matrix {i} has {f.shape[1]} columns, expected {R}")
            if int(f.shape[0]) != self.shape[i]:
                raise ValueError(f"Factor matrix {i} has {f.shape[0]} rows, expected {self.shape[i]} "
                                 f"(tensor mode-{i} size)")
        out = torch.empty((self.shape[mode], R), dtype=torch.float32, device=F[0].device)
        ptrs = (C.c_void_p * len(F))(*[f.data_ptr() for f in F])
        torch.cuda.synchronize()  # factors come from torch's stream; the tensor lives on the context stream
        check(lib().tenrso_sp_mttkrp_f32(self.ctx.ptr, self.ptr, ptrs, R, mode, out.data_ptr()))
        self.ctx.synchronize()
        return out

    def cp_als_with_init(self, init_factors, max_iters, tol, time_limit=None):
        """CP-ALS (cp/core.rs:59-182 rules) with the sparse MTTKRP; returns decomp.CpDecomp."""
        import torch
        from . import decomp
        F = [torch.as_tensor(f).float().cuda().contiguous().clone() for f in init_factors]
        R = int(F[0].shape[1])
        decomp._validate(self.shape, R, tol)
        opts = CpOptions()
        opts.rank, opts.max_iters, opts.precision = R, int(max_iters), _lib.PREC_FP32
        opts.tol = float(tol)
        opts.time_limit_s = -1.0 if time_limit is None else float(time_limit)
        opts.validate_nonneg_input = 1
        res = CpResult()
        hist = np.zeros(max(int(max_iters), 1), np.float64)
        ptrs = (C.c_void_p * len(F))(*[f.data_ptr() for f in F])
        torch.cuda.synchronize()
        rc = lib().tenrso_sp_cp_als_f32(self.ctx.ptr, self.ptr, ptrs, C.byref(opts), C.byref(res), hist.ctypes.data)
        if rc != _lib.OK:
            msg = _lib.last_error()
            mk = decomp._ERRORS.get(rc)
            raise mk(msg) if mk else _lib.TenrsoError(rc, msg)
        info = decomp.ConvergenceInfo(fit_history=hist[:res.iters].tolist(),
                                      reason=decomp.ConvergenceReason(res.reason), oscillated=bool(res.oscillated),
                                      oscillation_count=int(res.oscillation_count),
                                      final_fit_change=float(res.final_fit_change))
        return decomp.CpDecomp(factors=F, weights=None, fit=float(res.fit), iters=int(res.iters), convergence=info,
                               solve_path_max=int(res.solve_path_max), seconds=float(res.seconds))
