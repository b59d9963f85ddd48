"""Seeded benchmark workloads (numpy CSR on the host), via libspc5_synth.so.

Workloads (sizes as used by the benchmark and the parity tests):

* ``laplacian2d(512)``       5-point Laplacian, N=262,144, nnz=1,308,672,
                             diagonal 4, off-diagonal -1
* ``clustered_runs(N)``      random rows: lognormal lengths (mean 32,
                             sigma 0.8, clipped [1,512]), runs of 1..16
                             contiguous columns within +-N/64 of the
                             diagonal, values U[-1,1]; N=4,194,304 at scale
* ``stencil27(160)``         27-point stencil, N=4,096,000, nnz=109,215,352,
                             SPD with diagonal 26+1e-3, off-diagonal -1
* ``banded(n, bw, fill)``    banded matrices for the uneven batch
* ``uniform(n, seed)``       vectors U[-1,1]
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SYNTH_SO = os.path.join(HERE, "libspc5_synth.so")
_lib = None

_P = C.c_void_p
_I64 = C.c_int64
_U64 = C.c_uint64
_D = C.c_double


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SYNTH_SO):
            raise RuntimeError(f"{SYNTH_SO} missing: build with `python -m spc5_amd.build`")
        L = C.CDLL(SYNTH_SO)
        L.spc5_synth_uniform_f64.argtypes = [_I64, _U64, _D, _D, _P]
        L.spc5_synth_uniform_f32.argtypes = [_I64, _U64, _D, _D, _P]
        L.spc5_synth_runs_count.argtypes = [_I64, _U64, _P]
        L.spc5_synth_runs_count.restype = _I64
        L.spc5_synth_runs_fill.argtypes = [_I64, _U64, _U64, _P, _P, _P]
        L.spc5_synth_lap2d.argtypes = [_I64, _D, _D, _P, _P, _P]
        L.spc5_synth_lap2d.restype = _I64
        L.spc5_synth_stencil27.argtypes = [_I64, _D, _D, _P, _P, _P]
        L.spc5_synth_stencil27.restype = _I64
        L.spc5_synth_banded_count.argtypes = [_I64, _I64, _D, _U64, _P]
        L.spc5_synth_banded_count.restype = _I64
        L.spc5_synth_banded_fill.argtypes = [_I64, _I64, _D, _U64, _P, _P, _P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def uniform(n: int, seed: int, lo=-1.0, hi=1.0, dtype=np.float64):
    out = np.empty(n, dtype)
    f = lib().spc5_synth_uniform_f64 if np.dtype(dtype) == np.float64 else lib().spc5_synth_uniform_f32
    f(n, seed, lo, hi, _p(out))
    return out


def laplacian2d(n: int, diag=4.0, off=-1.0):
    N = n * n
    rowptr = np.empty(N + 1, np.int32)
    nnz = 5 * N - 4 * n
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    got = lib().spc5_synth_lap2d(n, diag, off, _p(rowptr), _p(col), _p(val))
    assert got == nnz
    return rowptr, col, val


def stencil27(n: int, diag=26.001, off=-1.0):
    N = n ** 3
    nnz = (3 * n - 2) ** 3
    rowptr = np.empty(N + 1, np.int32)
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    got = lib().spc5_synth_stencil27(n, diag, off, _p(rowptr), _p(col), _p(val))
    assert got == nnz
    return rowptr, col, val


def clustered_runs(N: int, seed: int = 1, value_seed: int = 2):
    rowptr = np.empty(N + 1, np.int32)
    nnz = lib().spc5_synth_runs_count(N, seed, _p(rowptr))
    if nnz < 0:
        raise ValueError("nnz does not fit int32")
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    lib().spc5_synth_runs_fill(N, seed, value_seed, _p(rowptr), _p(col), _p(val))
    return rowptr, col, val


def banded_threshold(fill: float) -> int:
    """Keep-threshold of the banded generators: floor(fill * 2^64), saturated."""
    t = fill * 18446744073709551616.0
    return 0xFFFFFFFFFFFFFFFF if t >= 18446744073709551615.0 else int(t)


def banded_gpu(n: int, bw: int, fill: float, seed: int):
    """The banded matrix of `banded`, generated on the GPU (device tensors
    rowptr int32, col int32, val float64); byte-identical to the host one."""
    import torch
    from ._lib import check
    from ._lib import lib as hip
    thr = banded_threshold(fill)
    rowptr = torch.empty(n + 1, dtype=torch.int32, device="cuda")
    nnz = C.c_int64(0)
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(hip().spc5_hip_synth_banded_count(n, bw, thr, seed, C.c_void_p(rowptr.data_ptr()), C.byref(nnz), s),
          "spc5_hip_synth_banded_count")
    col = torch.empty(max(1, nnz.value), dtype=torch.int32, device="cuda")
    val = torch.empty(max(1, nnz.value), dtype=torch.float64, device="cuda")
    check(hip().spc5_hip_synth_banded_fill(n, bw, thr, seed, C.c_void_p(rowptr.data_ptr()),
                                           C.c_void_p(col.data_ptr()), C.c_void_p(val.data_ptr()), s),
          "spc5_hip_synth_banded_fill")
    return rowptr, col[: nnz.value], val[: nnz.value]


def uneven_batch_params(count: int = 2000, seed: int = 11):
    """Per-matrix (rows, bandwidth, seed) of the uneven batch: rows
    log-uniform in [1,000, 200,000], bandwidth uniform in {4..256}."""
    u = uniform(2 * count, seed, 0.0, 1.0)
    out = []
    for i in range(count):
        rows = int(round(1000.0 * (200.0 ** u[2 * i])))
        bw = 4 + int(u[2 * i + 1] * 253.0)
        out.append((rows, min(bw, 256), 1000 + i))
    return out


def banded(n: int, bw: int, fill: float, seed: int):
    rowptr = np.empty(n + 1, np.int32)
    nnz = lib().spc5_synth_banded_count(n, bw, fill, seed, _p(rowptr))
    if nnz < 0:
        raise ValueError("nnz does not fit int32")
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    lib().spc5_synth_banded_fill(n, bw, fill, seed, _p(rowptr), _p(col), _p(val))
    return rowptr, col, val
