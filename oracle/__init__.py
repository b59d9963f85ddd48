"""CPU checker for the SPC5 beta(r,VS) hot path.  TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker: the product
library (``spc5_amd``) never imports, links or calls anything here.

Two back ends:

* ``libspc5_oracle.so`` -- plain-C restatement of the reference algorithms
  (``spc5_oracle.c``; each function cites the ``src/spc5.hpp`` lines it
  follows).
* ``_ref/libspc5_ref.so`` -- the reference header itself, compiled from the
  reference tree by ``oracle/Makefile`` (present only where that tree was
  available at build time).  Used to pin the restatement and as the timed
  CPU baseline.

Parity pins of every oracle function are listed in tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libspc5_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libspc5_ref.so")

VS_OF = {np.dtype(np.float64): 8, np.dtype(np.float32): 16}
MSZ_OF = {np.dtype(np.float64): 1, np.dtype(np.float32): 2}

_P = C.c_void_p
_I = C.c_int
_L = C.c_long
_I64 = C.c_int64


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


_lib = None
_ref = None


def lib():
    """Load the C restatement (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (run __graft_entry__.build())")
        L = C.CDLL(ORACLE_SO)
        L.spc5o_block_count.restype = _I64
        L.spc5o_block_count.argtypes = [_I, _I, _I, _P, _P]
        L.spc5o_convert.restype = _I64
        L.spc5o_convert.argtypes = [_I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P]
        for n in ("spc5o_csr_spmv_f64", "spc5o_csr_spmv_f32"):
            getattr(L, n).restype = None
            getattr(L, n).argtypes = [_I, _P, _P, _P, _P, _P]
        for n in ("spc5o_beta_spmv_f64", "spc5o_beta_spmv_f32"):
            getattr(L, n).restype = None
            getattr(L, n).argtypes = [_I, _I, _I, _L, _P, _P, _P, _P, _P]
        L.spc5o_decode.restype = _I64
        L.spc5o_decode.argtypes = [_I, _I, _I, _I, _L, _P, _P, _P, _P, _P, _P]
        L.spc5o_thread_split.restype = None
        L.spc5o_thread_split.argtypes = [_I, _I, _I, _I, _P, _P, _I, _P, _P, _P]
        L.spc5o_find_best.restype = _I
        L.spc5o_find_best.argtypes = [_I, _I, _P, _P, _I, _I, C.POINTER(C.c_double)]
        L.spc5o_coo_sorted_to_csr.restype = None
        L.spc5o_coo_sorted_to_csr.argtypes = [_I, _I, _P, _P, _P, _P, _P, _P]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    """Load the compiled reference (raises if it was not built)."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"reference build missing: {REF_SO}")
        L = C.CDLL(REF_SO)
        for n in ("ref_convert_f64", "ref_convert_f32"):
            getattr(L, n).restype = _I64
            getattr(L, n).argtypes = [_I, _I, _I, _I, _P, _P, _P, _P, _P, _I64, _P]
        for n in ("ref_spmv_f64", "ref_spmv_f32"):
            getattr(L, n).restype = None
            getattr(L, n).argtypes = [_I, _L, _P, _P, _P, _P, _P]
        for n in ("ref_csr_spmv_f64", "ref_csr_spmv_f32"):
            getattr(L, n).restype = None
            getattr(L, n).argtypes = [_I, _I, _I, _P, _P, _P, _P, _P]
        for n in ("ref_block_count_f64", "ref_block_count_f32"):
            getattr(L, n).restype = _I
            getattr(L, n).argtypes = [_I, _I, _I, _I, _P, _P, _P]
        L.ref_find_best.restype = _I
        L.ref_find_best.argtypes = [_I, _I, _I, _I, _I, _P, _P, _P, _P, C.POINTER(C.c_double)]
        L.ref_thread_split_f64.restype = None
        L.ref_thread_split_f64.argtypes = [_I, _I, _I, _I, _P, _P, _P, _I, _P, _P, _P]
        for n in ("ref_omp_bench_f64", "ref_omp_bench_f32"):
            getattr(L, n).restype = _I
            getattr(L, n).argtypes = [_I, _I, _I, _I, _P, _P, _P, _P, _P, _I, _I, _I,
                                      C.POINTER(C.c_double), _P]
        _ref = L
    return _ref


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# --------------------------------------------------------------------------
# Restatement (C) wrappers
# --------------------------------------------------------------------------

def block_count(r: int, vs: int, rowptr, col) -> int:
    rowptr, col = _i32(rowptr), _i32(col)
    return int(lib().spc5o_block_count(r, vs, len(rowptr) - 1, _ptr(rowptr), _ptr(col)))


def convert(r: int, rowptr, col, values, vs: int | None = None, msz: int | None = None):
    """CSR -> beta(r,VS).  Returns (bptr[ceil(n/r)+1], headers uint8, values permuted, nblocks)."""
    rowptr, col = _i32(rowptr), _i32(col)
    values = np.ascontiguousarray(values)
    vs = VS_OF[values.dtype] if vs is None else vs
    msz = MSZ_OF[values.dtype] if msz is None else msz
    n = len(rowptr) - 1
    nb = block_count(r, vs, rowptr, col)
    nbr = (n + r - 1) // r
    bptr = np.zeros(nbr + 1, dtype=np.int32)
    hdr = np.zeros(max(1, nb * (4 + r * msz)), dtype=np.uint8)
    vout = np.zeros_like(values)
    got = lib().spc5o_convert(r, vs, msz, values.dtype.itemsize, n, _ptr(rowptr), _ptr(col),
                              _ptr(values), _ptr(bptr), _ptr(hdr), _ptr(vout))
    assert got == nb
    return bptr, hdr[: nb * (4 + r * msz)], vout, nb


def csr_spmv(rowptr, col, val, x, y):
    """y += A x (in place) with CSR_Spmv_scalar's order."""
    rowptr, col = _i32(rowptr), _i32(col)
    f = lib().spc5o_csr_spmv_f64 if val.dtype == np.float64 else lib().spc5o_csr_spmv_f32
    f(len(rowptr) - 1, _ptr(rowptr), _ptr(col), _ptr(val), _ptr(x), _ptr(y))
    return y


def beta_spmv(r: int, nrows: int, bptr, hdr, val, x, y, vs=None, msz=None):
    """y += A x (in place) with the reference kernel's arithmetic.

    x must be padded by VS-1 and y by r-1 entries, as for the reference."""
    vs = VS_OF[val.dtype] if vs is None else vs
    msz = MSZ_OF[val.dtype] if msz is None else msz
    f = lib().spc5o_beta_spmv_f64 if val.dtype == np.float64 else lib().spc5o_beta_spmv_f32
    bptr = _i32(bptr)
    f(r, vs, msz, nrows, _ptr(bptr), _ptr(hdr), _ptr(val), _ptr(x), _ptr(y))
    return y


def decode(r: int, nrows: int, bptr, hdr, val, vs=None, msz=None):
    """beta(r,VS) -> (row, col, value) triplets in stream order."""
    vs = VS_OF[val.dtype] if vs is None else vs
    msz = MSZ_OF[val.dtype] if msz is None else msz
    n = len(val)
    rows = np.zeros(n, np.int32)
    cols = np.zeros(n, np.int32)
    vals = np.zeros_like(val)
    bptr = _i32(bptr)
    got = lib().spc5o_decode(r, vs, msz, val.dtype.itemsize, nrows, _ptr(bptr), _ptr(hdr), _ptr(val),
                             _ptr(rows), _ptr(cols), _ptr(vals))
    return rows[:got], cols[:got], vals[:got]


def thread_split(r: int, msz: int, nrows: int, bptr, hdr, nthreads: int):
    bptr = _i32(bptr)
    start = np.zeros(nthreads, np.int32)
    cnt = np.zeros(nthreads, np.int32)
    voff = np.zeros(nthreads, np.int32)
    nb = int(bptr[(nrows + r - 1) // r])
    lib().spc5o_thread_split(r, msz, nrows, nb, _ptr(bptr), _ptr(hdr), nthreads,
                             _ptr(start), _ptr(cnt), _ptr(voff))
    return start, cnt, voff


def find_best(is_double: bool, rowptr, col, nthreads: int = 0):
    rowptr, col = _i32(rowptr), _i32(col)
    est = C.c_double(0.0)
    r = lib().spc5o_find_best(int(is_double), len(rowptr) - 1, _ptr(rowptr), _ptr(col),
                              int(rowptr[-1]), nthreads, C.byref(est))
    return r, est.value


def coo_sorted_to_csr(nrows: int, ii, jj, vv):
    ii, jj = _i32(ii), _i32(jj)
    vv = np.ascontiguousarray(vv, dtype=np.float64)
    nnz = len(vv)
    rowptr = np.zeros(nrows + 1, np.int32)
    col = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1), np.float64)
    lib().spc5o_coo_sorted_to_csr(nrows, nnz, _ptr(ii), _ptr(jj), _ptr(vv), _ptr(rowptr), _ptr(col), _ptr(val))
    return rowptr, col[:nnz], val[:nnz]


# --------------------------------------------------------------------------
# Compiled reference wrappers
# --------------------------------------------------------------------------

def ref_convert(r: int, rowptr, col, values, ncols: int | None = None):
    """Reference CSR_to_SPC5 on a copy.  Returns (rowsSize[N+1] as left by the
    reference, headers, permuted values, nblocks)."""
    rowptr, col = _i32(rowptr), _i32(col)
    values = np.ascontiguousarray(values)
    n = len(rowptr) - 1
    ncols = n if ncols is None else ncols
    msz = MSZ_OF[values.dtype]
    cap = max(16, len(values) * (4 + r * msz) + 16)
    rs = np.zeros(n + 1, np.int32)
    hdr = np.zeros(cap, np.uint8)
    vout = np.zeros(max(1, len(values)), values.dtype)
    f = ref().ref_convert_f64 if values.dtype == np.float64 else ref().ref_convert_f32
    nb = f(r, n, ncols, len(values), _ptr(rowptr), _ptr(col), _ptr(values), _ptr(rs), _ptr(hdr), cap, _ptr(vout))
    if nb < 0:
        raise RuntimeError("header capacity")
    return rs, hdr[: nb * (4 + r * msz)], vout[: len(values)], int(nb)


def ref_spmv(r: int, nrows: int, bptr, hdr, val, x, y):
    f = ref().ref_spmv_f64 if val.dtype == np.float64 else ref().ref_spmv_f32
    bptr = _i32(bptr)
    hdr = np.ascontiguousarray(hdr, np.uint8)
    f(r, nrows, _ptr(bptr), _ptr(hdr), _ptr(val), _ptr(x), _ptr(y))
    return y


def ref_csr_spmv(rowptr, col, val, x, y, ncols=None):
    rowptr, col = _i32(rowptr), _i32(col)
    n = len(rowptr) - 1
    f = ref().ref_csr_spmv_f64 if val.dtype == np.float64 else ref().ref_csr_spmv_f32
    f(n, n if ncols is None else ncols, len(val), _ptr(rowptr), _ptr(col), _ptr(val), _ptr(x), _ptr(y))
    return y


def ref_block_count(r: int, rowptr, col, values):
    rowptr, col = _i32(rowptr), _i32(col)
    n = len(rowptr) - 1
    f = ref().ref_block_count_f64 if values.dtype == np.float64 else ref().ref_block_count_f32
    return int(f(r, n, n, len(values), _ptr(rowptr), _ptr(col), _ptr(values)))


def ref_find_best(is_double: bool, rowptr, col, values, nthreads: int = 0):
    rowptr, col = _i32(rowptr), _i32(col)
    n = len(rowptr) - 1
    est = C.c_double(0.0)
    v64 = values if is_double else None
    v32 = None if is_double else values
    r = ref().ref_find_best(int(is_double), nthreads, n, n, len(values), _ptr(rowptr), _ptr(col),
                            _ptr(v64), _ptr(v32), C.byref(est))
    return r, est.value


def ref_thread_split(r: int, rowptr, col, values, nthreads: int):
    rowptr, col = _i32(rowptr), _i32(col)
    n = len(rowptr) - 1
    start = np.zeros(nthreads, np.int32)
    cnt = np.zeros(nthreads, np.int32)
    voff = np.zeros(nthreads, np.int32)
    ref().ref_thread_split_f64(r, n, n, len(values), _ptr(rowptr), _ptr(col), _ptr(values), nthreads,
                               _ptr(start), _ptr(cnt), _ptr(voff))
    return start, cnt, voff


def ref_omp_bench(r: int, rowptr, col, values, x, y, nloops: int, nthreads: int, ncols=None, nsamples: int = 1):
    """Reference OpenMP path: convert, split (untimed), then nsamples times
    nloops x SPC5_Spmv_omp.

    Returns (threads used, conversion seconds, [SpMV seconds of each sample of nloops])."""
    rowptr, col = _i32(rowptr), _i32(col)
    n = len(rowptr) - 1
    tc = C.c_double(0)
    ts = np.zeros(max(1, nsamples), dtype=np.float64)
    f = ref().ref_omp_bench_f64 if values.dtype == np.float64 else ref().ref_omp_bench_f32
    th = f(r, n, n if ncols is None else ncols, len(values), _ptr(rowptr), _ptr(col), _ptr(values),
           _ptr(x), _ptr(y), nloops, max(1, nsamples), nthreads, C.byref(tc), _ptr(ts))
    return th, tc.value, [float(t) for t in ts]
