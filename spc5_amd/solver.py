"""Precision form: reduced-precision copies of a double matrix and the
CG / mixed-precision iterative-refinement solvers of libspc5_hip.so.

    A64 = DeviceMatrix.from_csr(rowptr, col, values_f64, r=4)     # beta(4,8)
    A16 = lowp(A64, "bf16")      # same headers, bfloat16 values
    x, st = ir_cg(A64, A16, b, rtol=1e-10)       # restarted refinement
    x, st = mixed_cg(A64, A16, b, rtol=1e-10)    # reliable updates
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import SPC5_HIP_BF16, SPC5_HIP_F32, SolveStats, Spc5Error, check, lib
from .device import DeviceMatrix, _need_cuda, _ptr, stream_ptr

_TO = {"f32": (SPC5_HIP_F32, torch.float32), "bf16": (SPC5_HIP_BF16, torch.bfloat16)}


def lowp(A64: DeviceMatrix, kind: str, stream=None) -> DeviceMatrix:
    """float ("f32") or bfloat16 ("bf16") copy of a double matrix stored with
    VS=8: the block structure and plan are shared, only values differ
    (each value rounded to nearest even; bfloat16 through float)."""
    if A64.values.dtype != torch.float64 or A64.vs != 8:
        raise Spc5Error("lowp needs a double beta(r,8) matrix")
    code, dt = _TO[kind]
    out = torch.empty(A64.nnz, dtype=dt, device=A64.values.device)
    check(lib().spc5_hip_values_convert(code, _ptr(A64.values), _ptr(out), A64.nnz, stream_ptr(stream)),
          "spc5_hip_values_convert")
    return A64.values_view(out)


def spmv_ex(A: DeviceMatrix, x, y, alpha=1.0, beta=1.0, stream=None):
    """y = alpha * (A x) + beta * y (y not read when beta == 0)."""
    _need_cuda(x, "x", A.vec_dtype)
    _need_cuda(y, "y", A.vec_dtype)
    check(lib().spc5_hip_spmv_ex(A._h, float(alpha), _ptr(x), float(beta), _ptr(y), stream_ptr(stream)),
          "spc5_hip_spmv_ex")
    return y


def cg(A: DeviceMatrix, b, x=None, rtol=1e-10, maxit=10000, stream=None):
    """Conjugate gradient in the matrix's vector precision.  Returns (x, stats)."""
    _need_cuda(b, "b", A.vec_dtype)
    if x is None:
        x = torch.zeros_like(b)
    _need_cuda(x, "x", A.vec_dtype)
    st = SolveStats()
    check(lib().spc5_hip_cg(A._h, _ptr(b), _ptr(x), float(rtol), int(maxit), C.byref(st), stream_ptr(stream)),
          "spc5_hip_cg")
    return x, {"iterations": st.iterations, "relres": st.relres}


def ir_cg(A64: DeviceMatrix, Alo: DeviceMatrix, b, x=None, rtol=1e-10, inner_rtol=None, max_outer=100,
          max_inner=5000, stream=None):
    """Iterative refinement: double residual and update, float CG on Alo.
    inner_rtol defaults to 3e-6 for float values (two refinement steps reach
    1e-10 on the 27-point stencil; 1e-4 took three) and 2e-2 for bfloat16."""
    _need_cuda(b, "b", torch.float64)
    if x is None:
        x = torch.zeros_like(b)
    _need_cuda(x, "x", torch.float64)
    if inner_rtol is None:
        inner_rtol = 2e-2 if Alo.values.dtype == torch.bfloat16 else 3e-6
    st = SolveStats()
    check(lib().spc5_hip_ir_cg(A64._h, Alo._h, _ptr(b), _ptr(x), float(rtol), float(inner_rtol), int(max_outer),
                               int(max_inner), C.byref(st), stream_ptr(stream)), "spc5_hip_ir_cg")
    return x, {"iterations": st.iterations, "outer": st.outer, "relres": st.relres}


def mixed_cg(A64: DeviceMatrix, Alo: DeviceMatrix, b, x=None, rtol=1e-10, delta=0.1, maxit=20000, vectors="f32",
             tail_rtol=None, stream=None):
    """CG with reliable updates: the iteration multiplies by Alo (float or
    bfloat16 values, float products and sums); its vectors are float
    (vectors="f32", with a float correction folded into x at every update)
    or double ("f64").  b - A64 x is recomputed in double whenever the
    iterated residual has dropped by `delta`.  Below tail_rtol (default: 0 =
    never for float values, whose iteration converges to 1e-10 with the
    reliable updates alone; 1e-3 for bfloat16, whose iteration stalls) the
    solve finishes as double CG with A64 from the same search direction."""
    _need_cuda(b, "b", torch.float64)
    if x is None:
        x = torch.zeros_like(b)
    _need_cuda(x, "x", torch.float64)
    if vectors not in ("f64", "f32"):
        raise Spc5Error("vectors must be 'f64' or 'f32'")
    if tail_rtol is None:
        tail_rtol = 1e-3 if Alo.values.dtype == torch.bfloat16 else 0.0
    st = SolveStats()
    check(lib().spc5_hip_mixed_cg(A64._h, Alo._h, _ptr(b), _ptr(x), float(rtol), float(delta), int(maxit),
                                  1 if vectors == "f64" else 0, float(tail_rtol), C.byref(st), stream_ptr(stream)),
          "spc5_hip_mixed_cg")
    return x, {"iterations": st.iterations, "outer": st.outer, "relres": st.relres,
               "tail_iterations": st.tail_iterations}
