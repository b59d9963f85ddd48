"""MatrixMarket coordinate files and COO -> CSR on the GPU.

`read_coo` parses a file with the library's host reader (the reference's
loadMM, examples/load_mm_and_compare.cpp:287-428: 0-based indices, pattern
values 1, mirrored symmetric / skew-symmetric entries).  `coo_to_csr` is the
reference's COO_unsorted_to_CSR / COO_sorted_to_CSR (src/spc5.hpp:77-114) on
the GPU: a stable radix sort by (row, column), duplicates kept in input
order.  `load_csr` chains the two into device CSR arrays ready for
DeviceMatrix.from_csr.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import Spc5Error, check, lib
from .device import stream_ptr


def read_coo(path: str):
    """(nrows, ncols, i, j, v): host int32 / float64 arrays of the file's
    entries (mirrored ones included), in file order."""
    nr, nc, ne = C.c_int64(), C.c_int64(), C.c_int64()
    check(lib().spc5_hip_mm_info(path.encode(), C.byref(nr), C.byref(nc), C.byref(ne)), "spc5_hip_mm_info")
    n = ne.value
    i = np.empty(max(n, 1), np.int32)
    j = np.empty(max(n, 1), np.int32)
    v = np.empty(max(n, 1), np.float64)
    got = C.c_int64()
    check(lib().spc5_hip_mm_read(path.encode(), n, i.ctypes.data, j.ctypes.data, v.ctypes.data, C.byref(got)),
          "spc5_hip_mm_read")
    return nr.value, nc.value, i[:n], j[:n], v[:n]


def coo_to_csr(i, j, v, nrows: int, ncols: int, sorted: bool = False, stream=None):
    """Device COO (int32 i, j; float64 or float32 v) -> device CSR
    (rowptr int32[nrows+1], col int32[nnz], val like v)."""
    for t, nm in ((i, "i"), (j, "j")):
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.int32 and t.is_contiguous()):
            raise Spc5Error(f"{nm} must be a contiguous int32 CUDA tensor")
    if not (isinstance(v, torch.Tensor) and v.is_cuda and v.dtype in (torch.float64, torch.float32)
            and v.is_contiguous()):
        raise Spc5Error("v must be a contiguous float64 or float32 CUDA tensor")
    nnz = v.numel()
    if i.numel() != nnz or j.numel() != nnz:
        raise Spc5Error("i, j and v need the same length")
    rowptr = torch.empty(nrows + 1, dtype=torch.int32, device=v.device)
    col = torch.empty(max(nnz, 1), dtype=torch.int32, device=v.device)
    val = torch.empty(max(nnz, 1), dtype=v.dtype, device=v.device)
    check(lib().spc5_hip_coo_to_csr(v.element_size(), nrows, ncols, nnz, 1 if sorted else 0, i.data_ptr(),
                                    j.data_ptr(), v.data_ptr(), rowptr.data_ptr(), col.data_ptr(), val.data_ptr(),
                                    stream_ptr(stream)), "spc5_hip_coo_to_csr")
    return rowptr, col[:nnz], val[:nnz]


def load_csr(path: str, dtype=torch.float64, stream=None):
    """MatrixMarket file -> (rowptr, col, val, ncols) on the GPU."""
    nrows, ncols, i, j, v = read_coo(path)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    rowptr, col, val = coo_to_csr(dev(i), dev(j), dev(v).to(dtype), nrows, ncols, stream=stream)
    return rowptr, col, val, ncols
