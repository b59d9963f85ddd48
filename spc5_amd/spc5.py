"""Python mirror of the reference's ``spc5.hpp`` API (same names, argument
meaning and errors), backed by the GPU library.

* ``SPC5Mat`` holds host (numpy) arrays with the reference's field names
  (src/spc5.hpp:44-62).
* ``CSR_to_SPC5*`` convert in place, on the GPU (byte-identical output;
  rowsSize keeps its N+1 length with stale tail entries, values are
  replaced, valuesColumnIndexes is left as it was -- src/spc5.hpp:218-227).
* ``SPC5_*_Spmv`` run y += A x through the reference C ABI exported by
  libspc5_hip.so (host pointers in, GPU kernel, y copied back).
* unknown formats raise ``ValueError``, the Python counterpart of the
  reference's ``std::invalid_argument`` (src/spc5.hpp:1014, 1115).
"""
from __future__ import annotations

import ctypes as C
import enum

import numpy as np

from ._lib import lib

SPC5_VEC_PADDING_X = 15  # src/spc5.hpp:26-29
SPC5_VEC_PADDING_Y = 7


class SPC5_MATRIX_TYPE(enum.IntEnum):  # src/spc5.hpp:17-24
    UNDEFINED_FORMAT = 0
    FORMAT_CSR = 1
    FORMAT_1rVc_WT = 2
    FORMAT_2rVc = 3
    FORMAT_4rVc = 4
    FORMAT_8rVc = 5


_R_OF = {SPC5_MATRIX_TYPE.FORMAT_1rVc_WT: 1, SPC5_MATRIX_TYPE.FORMAT_2rVc: 2,
         SPC5_MATRIX_TYPE.FORMAT_4rVc: 4, SPC5_MATRIX_TYPE.FORMAT_8rVc: 8}
_FMT_OF = {v: k for k, v in _R_OF.items()}


def ValPerVec(dtype) -> int:
    """512 bits / value size (src/spc5.hpp:145-148)."""
    return 512 // (8 * np.dtype(dtype).itemsize)


def mask_dtype(dtype):
    """SPC5Mat_Mask (src/spc5.hpp:31-42)."""
    return np.uint8 if np.dtype(dtype) == np.float64 else np.uint16


class SPC5Mat:
    """Host matrix with the reference's fields (src/spc5.hpp:44-62)."""

    def __init__(self, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.format = SPC5_MATRIX_TYPE.UNDEFINED_FORMAT
        self.numberOfRows = 0
        self.numberOfColumns = 0
        self.numberOfNNZ = 0
        self.values = np.zeros(0, self.dtype)
        self.rowsSize = np.zeros(1, np.int32)
        self.valuesColumnIndexes = np.zeros(0, np.int32)
        self.blocksColumnIndexesWithMasks = np.zeros(0, np.uint8)
        self.numberOfBlocks = 0


def COO_sorted_to_CSR(nbRows: int, nbCols: int, i, j, v, dtype=np.float64) -> SPC5Mat:
    """Entries sorted by (row, column) -> CSR (src/spc5.hpp:77-104)."""
    i = np.asarray(i, np.int64)
    csr = SPC5Mat(dtype)
    csr.format = SPC5_MATRIX_TYPE.FORMAT_CSR
    csr.numberOfRows = int(nbRows)
    csr.numberOfColumns = int(nbCols)
    csr.numberOfNNZ = len(i)
    csr.values = np.ascontiguousarray(v, dtype=csr.dtype)
    csr.valuesColumnIndexes = np.ascontiguousarray(j, dtype=np.int32)
    counts = np.bincount(i, minlength=nbRows) if len(i) else np.zeros(nbRows, np.int64)
    csr.rowsSize = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    return csr


def COO_unsorted_to_CSR(nbRows: int, nbCols: int, i, j, v, dtype=np.float64) -> SPC5Mat:
    """Sort by (row, column), then COO_sorted_to_CSR (src/spc5.hpp:107-114)."""
    i = np.asarray(i)
    j = np.asarray(j)
    order = np.lexsort((j, i))
    return COO_sorted_to_CSR(nbRows, nbCols, i[order], j[order], np.asarray(v)[order], dtype)


def _to_device_csr(csr: SPC5Mat):
    import torch
    dev = torch.device("cuda")
    rowptr = torch.from_numpy(np.ascontiguousarray(csr.rowsSize[: csr.numberOfRows + 1], np.int32)).to(dev)
    col = torch.from_numpy(np.ascontiguousarray(csr.valuesColumnIndexes, np.int32)).to(dev)
    val = torch.from_numpy(np.ascontiguousarray(csr.values)).to(dev)
    return rowptr, col, val


def _csr_to_spc5(csr: SPC5Mat, r: int) -> None:
    from .device import convert_csr
    if csr.format != SPC5_MATRIX_TYPE.FORMAT_CSR:
        raise ValueError("CSR_to_SPC5 : matrix is not in CSR format")
    rowptr, col, val = _to_device_csr(csr)
    rows_size, hdr, vout, nb = convert_csr(rowptr, col, val, r)
    nbr = (csr.numberOfRows + r - 1) // r
    rs = csr.rowsSize.copy()
    rs[: nbr + 1] = rows_size.cpu().numpy()
    csr.rowsSize = rs
    csr.blocksColumnIndexesWithMasks = hdr.cpu().numpy()
    csr.values = vout.cpu().numpy()
    csr.numberOfBlocks = int(nb)
    csr.format = _FMT_OF[r]


def CSR_to_SPC5_1rVc(csr: SPC5Mat) -> None:
    _csr_to_spc5(csr, 1)


def CSR_to_SPC5_2rVc(csr: SPC5Mat) -> None:
    _csr_to_spc5(csr, 2)


def CSR_to_SPC5_4rVc(csr: SPC5Mat) -> None:
    _csr_to_spc5(csr, 4)


def CSR_to_SPC5_8rVc(csr: SPC5Mat) -> None:
    _csr_to_spc5(csr, 8)


def CSR_to_SPC5(mat: SPC5Mat, matType) -> None:
    """Format switch (src/spc5.hpp:989-1018)."""
    try:
        r = _R_OF[SPC5_MATRIX_TYPE(matType)]
    except (KeyError, ValueError):
        raise ValueError("CSR_to_SPC5 : Unknown format type") from None
    _csr_to_spc5(mat, r)


def _coo_to(r):
    def f(nbRows, nbCols, i, j, v, dtype=np.float64):
        m = COO_unsorted_to_CSR(nbRows, nbCols, i, j, v, dtype)
        _csr_to_spc5(m, r)
        return m
    f.__name__ = f"COO_to_SPC5_{r}rVc"
    f.__doc__ = f"COO -> CSR -> beta({r},VS) (src/spc5.hpp:490-496, 615-621, 744-750, 870-876)."
    return f


COO_to_SPC5_1rVc = _coo_to(1)
COO_to_SPC5_2rVc = _coo_to(2)
COO_to_SPC5_4rVc = _coo_to(4)
COO_to_SPC5_8rVc = _coo_to(8)


def _block_count(csr: SPC5Mat, r: int) -> int:
    from .device import block_count
    rowptr, col, _ = _to_device_csr(csr)
    return block_count(rowptr, col, r, ValPerVec(csr.dtype))


def SPC5_1rVc_block_count(csr: SPC5Mat) -> int:
    return _block_count(csr, 1)


def SPC5_2rVc_block_count(csr: SPC5Mat) -> int:
    return _block_count(csr, 2)


def SPC5_4rVc_block_count(csr: SPC5Mat) -> int:
    return _block_count(csr, 4)


def SPC5_8rVc_block_count(csr: SPC5Mat) -> int:
    return _block_count(csr, 8)


def SPC5_block_count(csr: SPC5Mat, matType) -> int:
    """Format switch of the block counters (src/spc5.hpp:1120-1148).  Unlike
    the reference, which discards the count, this returns it."""
    try:
        r = _R_OF[SPC5_MATRIX_TYPE(matType)]
    except (KeyError, ValueError):
        raise ValueError("SPC5_Spmv : Unknown format type") from None
    return _block_count(csr, r)


def _spmv(mat: SPC5Mat, r: int, x, y):
    if mat.format != _FMT_OF[r]:
        raise ValueError(f"matrix is not in {SPC5_type_to_string(_FMT_OF[r])} format")
    if mat.dtype == np.float64:
        f = getattr(lib(), f"core_SPC5_{r}rVc_Spmv_double")
    else:
        f = getattr(lib(), f"core_SPC5_{r}rVc_Spmv_float")
    if x.dtype != mat.dtype or y.dtype != mat.dtype or not (x.flags.c_contiguous and y.flags.c_contiguous):
        raise ValueError("x and y must be contiguous arrays of the matrix dtype")
    f(mat.numberOfRows, mat.rowsSize.ctypes.data_as(C.c_void_p),
      mat.blocksColumnIndexesWithMasks.ctypes.data_as(C.c_void_p), mat.values.ctypes.data_as(C.c_void_p),
      x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p))
    return y


def SPC5_1rVc_Spmv(mat, x, y):
    return _spmv(mat, 1, x, y)


def SPC5_2rVc_Spmv(mat, x, y):
    return _spmv(mat, 2, x, y)


def SPC5_4rVc_Spmv(mat, x, y):
    return _spmv(mat, 4, x, y)


def SPC5_8rVc_Spmv(mat, x, y):
    return _spmv(mat, 8, x, y)


def SPC5_Spmv(mat: SPC5Mat, x, y):
    """y += A x, format switch (src/spc5.hpp:1090-1118)."""
    if mat.format not in _R_OF:
        raise ValueError("SPC5_Spmv : Unknown format type")
    return _spmv(mat, _R_OF[mat.format], x, y)


def SPC5_type_to_string(matType) -> str:
    """src/spc5.hpp:1229-1252."""
    return {SPC5_MATRIX_TYPE.FORMAT_1rVc_WT: "1rVc_WT", SPC5_MATRIX_TYPE.FORMAT_2rVc: "2rVc",
            SPC5_MATRIX_TYPE.FORMAT_4rVc: "4rVc", SPC5_MATRIX_TYPE.FORMAT_8rVc: "8rVc"}.get(matType, "undefined")


# Format-selection model of the reference (src/spc5.hpp:1151-1223): a cubic
# in the mean number of values per block for r = 1, 2, 4 with the
# reference's CPU-fitted coefficients; block counts come from the GPU.
_FIND_BEST_SEQ = {
    True: [(1, (2.186490e-03, -7.931533e-02, 8.719444e-01, 3.712380e-01)),
           (2, (9.822706e-04, -4.500002e-02, 6.714119e-01, 2.413985e-01)),
           (4, (1.320496e-04, -1.222130e-02, 3.520409e-01, 3.603195e-01))],
    False: [(1, (-4.039852e-03, 7.997606e-02, 6.892119e-02, 1.508645e+00)),
            (2, (-9.195940e-04, 4.442050e-03, 5.273045e-01, 1.101610e+00)),
            (4, (-4.464933e-05, 1.867266e-03, 1.520910e-01, 1.186131e+00))],
}
# src/spc5.hpp:1264-1304 (double only in the reference)
_FIND_BEST_OMP = [(1, (4.3764, 0.0981, 1.8728, 0.0721)),
                  (2, (5.2172, 0.1435, 0.8331, 0.0556)),
                  (4, (5.3212, 0.1921, 0.5201, 0.0263))]


def SPC5_find_best(csr: SPC5Mat):
    """(format, estimated GFlop/s) from the reference's sequential model."""
    best, est = SPC5_MATRIX_TYPE.UNDEFINED_FORMAT, 0.0
    for r, c in _FIND_BEST_SEQ[csr.dtype == np.float64]:
        a = csr.numberOfNNZ / _block_count(csr, r)
        s = c[0] * a ** 3 + c[1] * a ** 2 + c[2] * a + c[3]
        if s > est:
            best, est = _FMT_OF[r], s
    return best, est


def SPC5_find_best_omp(csr: SPC5Mat, nbThreads: int):
    """(format, estimate) from the reference's OpenMP model (double only)."""
    if csr.dtype != np.float64:
        raise ValueError("SPC5_find_best_omp is defined for double only (src/spc5.hpp:1259-1264)")
    best, est = SPC5_MATRIX_TYPE.UNDEFINED_FORMAT, 0.0
    for r, c in _FIND_BEST_OMP:
        a = csr.numberOfNNZ / _block_count(csr, r)
        s = c[0] + c[1] * nbThreads + c[2] * a + c[3] * nbThreads * a
        if s > est:
            best, est = _FMT_OF[r], s
    return best, est
