"""Device-resident beta(r,VS) matrices (PyTorch tensors for HBM and streams).

``DeviceMatrix`` holds the reference layout -- block-row pointers, packed
headers, block-ordered values (include/spc5_hip.h) -- in device memory and the
library handle with its tile plan.  All work runs in the HIP library; torch
only allocates memory and supplies streams.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import (SPC5_HIP_BF16, SPC5_HIP_F32, SPC5_HIP_F64, SPC5_HIP_OPT_ROW_SORT, SPC5_HIP_OPT_SPMM_COLUMNS,
                   SPC5_HIP_OPT_SPMM_X16, SPC5_HIP_OPT_TILE_VALUES, Spc5Error, check, lib)

_VT = {torch.float64: SPC5_HIP_F64, torch.float32: SPC5_HIP_F32, torch.bfloat16: SPC5_HIP_BF16}
_VEC_DTYPE = {SPC5_HIP_F64: torch.float64, SPC5_HIP_F32: torch.float32, SPC5_HIP_BF16: torch.float32}
FORMAT_NAMES = {1: "1rVc_WT", 2: "2rVc", 4: "4rVc", 8: "8rVc"}


def default_vs(dtype: torch.dtype) -> int:
    """VS of the reference formats: 512 bits / value size (src/spc5.hpp:145-148)."""
    return 8 if dtype in (torch.float64, torch.bfloat16) else 16


def mask_bytes(vs: int) -> int:
    return 1 if vs == 8 else 2


def stream_ptr(stream=None):
    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else C.c_void_p(0)


def _need_cuda(t, name, dtype=None):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise Spc5Error(f"{name} must be a CUDA (HIP) tensor")
    if dtype is not None and t.dtype != dtype:
        raise Spc5Error(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise Spc5Error(f"{name} must be contiguous")


def convert_csr(rowptr, col, values, r: int, vs: int | None = None, stream=None):
    """GPU converter: CSR device tensors -> (rows_size[ceil(N/r)+1], headers, values).

    Byte-identical to the reference CSR_to_SPC5 (src/spc5.hpp:165-228)."""
    _need_cuda(rowptr, "rowptr", torch.int32)
    _need_cuda(col, "col", torch.int32)
    _need_cuda(values, "values")
    vs = default_vs(values.dtype) if vs is None else vs
    n = rowptr.numel() - 1
    nbr = (n + r - 1) // r
    dev = rowptr.device
    rows_size = torch.empty(nbr + 1, dtype=torch.int32, device=dev)
    nb = C.c_int64(0)
    s = stream_ptr(stream)
    check(lib().spc5_hip_convert_count(vs, r, n, _ptr(rowptr), _ptr(col), _ptr(rows_size), C.byref(nb), s),
          "spc5_hip_convert_count")
    nblocks = nb.value
    hdr = torch.empty(nblocks * (4 + r * mask_bytes(vs)), dtype=torch.uint8, device=dev)
    vout = torch.empty_like(values)
    check(lib().spc5_hip_convert_fill(values.element_size(), vs, mask_bytes(vs), r, n, _ptr(rowptr), _ptr(col),
                                      _ptr(values), _ptr(rows_size), _ptr(hdr), _ptr(vout), s),
          "spc5_hip_convert_fill")
    return rows_size, hdr, vout, nblocks


def block_count(rowptr, col, r: int, vs: int = 8, stream=None) -> int:
    """Number of beta(r,VS) blocks of a device CSR matrix (GPU count pass)."""
    _need_cuda(rowptr, "rowptr", torch.int32)
    _need_cuda(col, "col", torch.int32)
    n = rowptr.numel() - 1
    tmp = torch.empty((n + r - 1) // r + 1, dtype=torch.int32, device=rowptr.device)
    nb = C.c_int64(0)
    check(lib().spc5_hip_convert_count(vs, r, n, _ptr(rowptr), _ptr(col), _ptr(tmp), C.byref(nb),
                                       stream_ptr(stream)), "spc5_hip_convert_count")
    return nb.value


class DeviceMatrix:
    """A beta(r,VS) matrix in HBM with its GPU work plan (library handle)."""

    def __init__(self, *, r, vs, nrows, ncols, nnz, nblocks, rows_size, headers, values, stream=None,
                 _base=None):
        self.r = int(r)
        self.vs = int(vs)
        self.nrows = int(nrows)
        self.ncols = int(ncols)
        self.nnz = int(nnz)
        self.nblocks = int(nblocks)
        self.rows_size = rows_size
        self.headers = headers
        self.values = values
        self.value_type = _VT[values.dtype]
        self.vec_dtype = _VEC_DTYPE[self.value_type]
        self._base = _base
        self._h = C.c_void_p(0)
        for t, nm in ((rows_size, "rows_size"), (headers, "headers"), (values, "values")):
            _need_cuda(t, nm)
        if _base is None:
            check(lib().spc5_hip_mat_create(self.value_type, self.vs, self.r, self.nrows, self.ncols, self.nnz,
                                            self.nblocks, _ptr(rows_size), _ptr(headers), _ptr(values),
                                            stream_ptr(stream), C.byref(self._h)), "spc5_hip_mat_create")
        else:
            check(lib().spc5_hip_mat_create_values_view(_base._h, self.value_type, _ptr(values), C.byref(self._h)),
                  "spc5_hip_mat_create_values_view")

    @classmethod
    def from_csr(cls, rowptr, col, values, r: int, ncols: int | None = None, vs: int | None = None, stream=None):
        """Convert a device CSR matrix on the GPU and build the plan."""
        rows_size, hdr, vout, nblocks = convert_csr(rowptr, col, values, r, vs, stream)
        n = rowptr.numel() - 1
        vs = default_vs(values.dtype) if vs is None else vs
        return cls(r=r, vs=vs, nrows=n, ncols=n if ncols is None else ncols, nnz=values.numel(), nblocks=nblocks,
                   rows_size=rows_size, headers=hdr, values=vout, stream=stream)

    def save(self, path: str):
        """Write the matrix (beta(r,VS) arrays) to `path`; see load."""
        check(lib().spc5_hip_mat_save(self._h, str(path).encode()), "spc5_hip_mat_save")

    @classmethod
    def load(cls, path: str, stream=None):
        """Read a matrix written by save (or spc5_hip_mat_save) into HBM and
        build its plan.  The layout is the one include/spc5_hip.h gives for
        spc5_hip_mat_save."""
        with open(path, "rb") as f:
            blob = f.read()
        if blob[:8] != b"SPC5BRVC" or int(np.frombuffer(blob, np.uint32, 1, 8)[0]) != 1:
            raise Spc5Error(f"not a beta(r,VS) matrix file: {path}")
        vt, vs, mb, r, _ = (int(a) for a in np.frombuffer(blob, np.int32, 5, 12))
        nrows, ncols, nnz, nblocks = (int(a) for a in np.frombuffer(blob, np.int64, 4, 32))
        if vt not in _VT.values() or r not in FORMAT_NAMES or mb != mask_bytes(vs):
            raise Spc5Error(f"unsupported matrix file: {path}")
        nbr = (nrows + r - 1) // r
        off = 64
        rs = np.frombuffer(blob, np.int32, nbr + 1, off)
        off += 4 * (nbr + 1)
        hb = nblocks * (4 + r * mb)
        hdr = np.frombuffer(blob, np.uint8, hb, off)
        off += hb
        vdt = {SPC5_HIP_F64: np.float64, SPC5_HIP_F32: np.float32, SPC5_HIP_BF16: np.int16}[vt]
        vals = np.frombuffer(blob, vdt, nnz, off)
        if off + vals.nbytes != len(blob):
            raise Spc5Error(f"matrix file size does not match its header: {path}")
        dev = lambda a: torch.from_numpy(np.array(a)).cuda()
        v = dev(vals)
        if vt == SPC5_HIP_BF16:
            v = v.view(torch.bfloat16)
        return cls(r=r, vs=vs, nrows=nrows, ncols=ncols, nnz=nnz, nblocks=nblocks, rows_size=dev(rs),
                   headers=dev(hdr) if hb else torch.zeros(1, dtype=torch.uint8, device="cuda"), values=v,
                   stream=stream)

    def values_view(self, values):
        """Same block structure and plan, other values (e.g. float32/bfloat16
        copies of a float64 matrix stored with VS=8)."""
        _need_cuda(values, "values")
        if values.numel() != self.nnz:
            raise Spc5Error("values_view needs nnz values")
        return DeviceMatrix(r=self.r, vs=self.vs, nrows=self.nrows, ncols=self.ncols, nnz=self.nnz,
                            nblocks=self.nblocks, rows_size=self.rows_size, headers=self.headers, values=values,
                            _base=self)

    def info(self):
        nt, nbr, nb, nnz = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().spc5_hip_mat_info(self._h, C.byref(nt), C.byref(nbr), C.byref(nb), C.byref(nnz)), "info")
        return {"ntiles": nt.value, "nblockrows": nbr.value, "nblocks": nb.value, "nnz": nnz.value}

    @property
    def row_sort(self) -> bool:
        """Sum each batch's products in row order (r > 1; see
        SPC5_HIP_OPT_ROW_SORT in include/spc5_hip.h)."""
        return getattr(self, "_row_sort", False)

    @row_sort.setter
    def row_sort(self, on: bool):
        check(lib().spc5_hip_mat_set_option(self._h, SPC5_HIP_OPT_ROW_SORT, 1 if on else 0), "set_option")
        self._row_sort = bool(on)

    @property
    def spmm_columns(self) -> bool:
        """Multi-right-hand-side products request each X row once per set
        column of a block (r > 1; see SPC5_HIP_OPT_SPMM_COLUMNS in
        include/spc5_hip.h)."""
        return getattr(self, "_spmm_columns", False)

    @spmm_columns.setter
    def spmm_columns(self, on: bool):
        check(lib().spc5_hip_mat_set_option(self._h, SPC5_HIP_OPT_SPMM_COLUMNS, 1 if on else 0), "set_option")
        self._spmm_columns = bool(on)

    @property
    def spmm_x16(self) -> bool:
        """Multi-right-hand-side products request 16 bytes of an X row per
        lane (two doubles or four floats) where X's layout allows it; on by
        default for double vectors (see SPC5_HIP_OPT_SPMM_X16 in
        include/spc5_hip.h)."""
        return getattr(self, "_spmm_x16", self.vec_dtype == torch.float64)

    @spmm_x16.setter
    def spmm_x16(self, on: bool):
        check(lib().spc5_hip_mat_set_option(self._h, SPC5_HIP_OPT_SPMM_X16, 1 if on else 0), "set_option")
        self._spmm_x16 = bool(on)

    @property
    def tile_values(self) -> int:
        """Values one wave's tile aims at (0: the default; see
        SPC5_HIP_OPT_TILE_VALUES in include/spc5_hip.h)."""
        return getattr(self, "_tile_values", 0)

    @tile_values.setter
    def tile_values(self, target: int):
        check(lib().spc5_hip_mat_set_option(self._h, SPC5_HIP_OPT_TILE_VALUES, int(target)), "set_option")
        self._tile_values = int(target)

    def tile_target(self) -> int:
        """Values per tile the current plan aims at (the set tile size, or
        the default rule nnz / 8192 clamped to [512, 8192])."""
        return self.tile_values or min(8192, max(512, self.nnz // 8192))

    @property
    def format_name(self):
        return FORMAT_NAMES[self.r]

    def header_bytes(self):
        return 4 + self.r * mask_bytes(self.vs)

    def algorithmic_bytes(self, vec_bytes: int | None = None):
        """Bytes one y += A x must move: values, headers, block-row pointers,
        x once and y read + written."""
        sv = self.values.element_size()
        sx = self.vec_dtype.itemsize if vec_bytes is None else vec_bytes
        nbr = (self.nrows + self.r - 1) // self.r
        return (self.nnz * sv + self.nblocks * self.header_bytes() + (nbr + 1) * 4 + self.ncols * sx
                + 2 * self.nrows * sx)

    def spmv(self, x, y, stream=None):
        """y += A x on the GPU (x, y device tensors of the vector dtype)."""
        _need_cuda(x, "x", self.vec_dtype)
        _need_cuda(y, "y", self.vec_dtype)
        if x.numel() < self.ncols or y.numel() < self.nrows:
            raise Spc5Error(f"x needs >= {self.ncols} and y >= {self.nrows} entries")
        check(lib().spc5_hip_spmv(self._h, _ptr(x), _ptr(y), stream_ptr(stream)), "spc5_hip_spmv")
        return y

    def spmm(self, X, Y, alpha=1.0, beta=1.0, kt: int = 0, stream=None):
        """Y = alpha * (A X) + beta * Y for row-major blocks X (ncols x k) and
        Y (nrows x k); kt = columns per pass (0: all k, one read of A)."""
        for t, nm, rows in ((X, "X", self.ncols), (Y, "Y", self.nrows)):
            if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dim() != 2 or t.dtype != self.vec_dtype:
                raise Spc5Error(f"{nm} must be a 2-D {self.vec_dtype} CUDA tensor")
            if t.stride(1) != 1 or t.shape[0] < rows:
                raise Spc5Error(f"{nm} must be row-major with >= {rows} rows")
        k = X.shape[1]
        if Y.shape[1] != k:
            raise Spc5Error("X and Y need the same number of columns")
        check(lib().spc5_hip_spmm(self._h, k, int(kt), float(alpha), _ptr(X), X.stride(0), float(beta), _ptr(Y),
                                  Y.stride(0), stream_ptr(stream)), "spc5_hip_spmm")
        return Y

    def close(self):
        if self._h and self._h.value:
            lib().spc5_hip_mat_destroy(self._h)
            self._h = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
