"""Batching form: many independent beta(r,VS) matrices multiplied in one
persistent, hipGraph-captured launch (libspc5_hip.so spc5_hip_batch_*)."""
from __future__ import annotations

import ctypes as C

from ._lib import Spc5Error, check, lib
from .device import DeviceMatrix, _need_cuda, _ptr, stream_ptr


# `waves` for one wave per tile (any count above the tile count is capped there)
ONE_WAVE_PER_TILE = 2**31 - 1


class Batch:
    """y_i += A_i x_i for all i in one launch.  Matrices may use different r
    but share value type and VS; the x_i / y_i tensors are kept alive here.
    `waves`: persistent waves walking the tile list (0: the kernel's resident
    count; ONE_WAVE_PER_TILE: one wave per tile)."""

    def __init__(self, mats: list[DeviceMatrix], xs, ys, waves: int = 0):
        if not mats or not (len(mats) == len(xs) == len(ys)):
            raise Spc5Error("Batch needs equal, non-empty lists of matrices, x and y")
        for m, x, y in zip(mats, xs, ys):
            _need_cuda(x, "x", m.vec_dtype)
            _need_cuda(y, "y", m.vec_dtype)
            if x.numel() < m.ncols or y.numel() < m.nrows:
                raise Spc5Error("x or y too short for its matrix")
        self.mats, self.xs, self.ys = list(mats), list(xs), list(ys)
        n = len(mats)
        H = (C.c_void_p * n)(*[m._h.value for m in mats])
        X = (C.c_void_p * n)(*[_ptr(x).value for x in xs])
        Y = (C.c_void_p * n)(*[_ptr(y).value for y in ys])
        self._h = C.c_void_p(0)
        check(lib().spc5_hip_batch_create(n, H, X, Y, int(waves), C.byref(self._h)), "spc5_hip_batch_create")

    def info(self):
        nt, wv = C.c_int64(), C.c_int()
        check(lib().spc5_hip_batch_info(self._h, C.byref(nt), C.byref(wv)), "spc5_hip_batch_info")
        return {"ntiles": nt.value, "waves": wv.value, "nmats": len(self.mats)}

    def spmv(self, stream=None, graph: bool = True):
        f = lib().spc5_hip_batch_spmv if graph else lib().spc5_hip_batch_spmv_direct
        check(f(self._h, stream_ptr(stream)), "spc5_hip_batch_spmv")

    def algorithmic_bytes(self):
        return sum(m.algorithmic_bytes() for m in self.mats)

    def close(self):
        if self._h and self._h.value:
            lib().spc5_hip_batch_destroy(self._h)
            self._h = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
