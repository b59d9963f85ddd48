"""Autotuner of the batching form: r (rows per block) per matrix and the
column chunk kt of multi-right-hand-side products.

r is chosen from a cost model evaluated on GPU block counts (one count pass
per candidate, no conversion):

    t(r) = a_r * bytes(r) + b_r * nblocks(r) + c_r * tiles(r)

bytes(r) = nnz*sv + nblocks*(4 + r*MB) + (ceil(N/r)+1)*4 + ncols*sx + 2*N*sx
is what one product must move.  The per-r coefficients are fitted (least
squares) by `calibrate()` from timed products of synthetic matrices on the
device and stored as JSON; without a calibration file the model reduces to
bytes at the device bandwidth plus a fixed per-block decode cost.  The
reference's own CPU model (SPC5_find_best, cubic in values per block) is
kept in spc5_amd.spc5 for API parity; its coefficients describe a CPU and are
not used here.

`mode="measure"` converts the matrix with each candidate r and times it
instead (exact, costs four conversions); `tune_spmm_kt` times the candidate
column chunks of a product with k right-hand sides.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np
import torch

from .device import DeviceMatrix, block_count, default_vs, mask_bytes

RS = (1, 2, 4, 8)
DEFAULT_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_mi355x.json")
# bytes at 5.5 TB/s plus 12 ps per block of decode work: used until calibrate() runs
_DEFAULT = {str(r): {"a": 1.0 / 5.5e12, "b": 1.2e-11, "c": 0.0} for r in RS}


def load_coefficients(path: str = DEFAULT_FILE):
    if os.path.exists(path):
        with open(path) as f:
            d = json.load(f)
        return d.get("coefficients", _DEFAULT)
    return _DEFAULT


def spmv_bytes(nrows, ncols, nnz, nb, r, value_bytes, vec_bytes, vs):
    return (nnz * value_bytes + nb * (4 + r * mask_bytes(vs)) + ((nrows + r - 1) // r + 1) * 4
            + ncols * vec_bytes + 2 * nrows * vec_bytes)


def features(nrows, ncols, nnz, nb, r, value_bytes, vec_bytes, vs, tile_target=None):
    tiles = max(1, nnz // (tile_target or max(512, min(8192, nnz // 8192 or 1))))
    return np.array([spmv_bytes(nrows, ncols, nnz, nb, r, value_bytes, vec_bytes, vs), nb, tiles], dtype=np.float64)


def block_counts(rowptr, col, vs: int = 8):
    """{r: number of beta(r,VS) blocks} from the GPU count pass."""
    return {r: block_count(rowptr, col, r, vs) for r in RS}


def predict(counts, nrows, ncols, nnz, value_bytes=8, vec_bytes=8, vs=8, coeffs=None):
    coeffs = coeffs or load_coefficients()
    out = {}
    for r, nb in counts.items():
        f = features(nrows, ncols, nnz, nb, r, value_bytes, vec_bytes, vs)
        c = coeffs[str(r)]
        out[r] = float(c["a"] * f[0] + c["b"] * f[1] + c["c"] * f[2])
    return out


def _time_spmv(m: DeviceMatrix, reps=20, warm=3):
    x = torch.ones(m.ncols, dtype=m.vec_dtype, device="cuda")
    y = torch.zeros(m.nrows, dtype=m.vec_dtype, device="cuda")
    for _ in range(warm):
        m.spmv(x, y)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        m.spmv(x, y)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps * 1e-3


def choose_r(rowptr, col, values, ncols=None, mode: str = "model", coeffs=None):
    """Best r for a device CSR matrix.  Returns (r, details)."""
    vs = default_vs(values.dtype) if values.dtype != torch.float32 else 8
    n = rowptr.numel() - 1
    ncols = n if ncols is None else ncols
    nnz = values.numel()
    vb = values.element_size()
    xb = 8 if values.dtype == torch.float64 else 4
    counts = block_counts(rowptr, col, vs)
    if mode == "model":
        pred = predict(counts, n, ncols, nnz, vb, xb, vs, coeffs)
        best = min(pred, key=pred.get)
        return best, {"mode": "model", "blocks": counts, "predicted_s": pred}
    times = {}
    for r in RS:
        m = DeviceMatrix.from_csr(rowptr, col, values, r, ncols=ncols, vs=vs)
        times[r] = _time_spmv(m)
        del m
    best = min(times, key=times.get)
    return best, {"mode": "measure", "blocks": counts, "measured_s": times}


def calibrate(matrices, path: str = DEFAULT_FILE, reps=20):
    """Fit the per-r coefficients from timed products.  `matrices` is a list
    of device CSR triples (rowptr, col, values float64).  Writes `path`."""
    rows = {r: [] for r in RS}
    ts = {r: [] for r in RS}
    for rp, col, val in matrices:
        n = rp.numel() - 1
        for r in RS:
            m = DeviceMatrix.from_csr(rp, col, val, r)
            t = _time_spmv(m, reps=reps)
            rows[r].append(features(n, n, val.numel(), m.nblocks, r, 8, 8, 8))
            ts[r].append(t)
            del m
    coeffs = {}
    for r in RS:
        Xf = np.array(rows[r])
        y = np.array(ts[r])
        sol, *_ = np.linalg.lstsq(Xf, y, rcond=None)
        sol = np.maximum(sol, 0.0)
        coeffs[str(r)] = {"a": float(sol[0]), "b": float(sol[1]), "c": float(sol[2])}
    with open(path, "w") as f:
        json.dump({"device": torch.cuda.get_device_name(0), "time": time.time(), "coefficients": coeffs,
                   "samples": len(matrices)}, f, indent=1)
    return coeffs


def tune_spmm_kt(m: DeviceMatrix, k: int, candidates=None, reps=5):
    """Fastest column chunk for Y += A X with k right-hand sides."""
    cands = candidates or sorted({k, 64, 32, 16, 8} & set(range(1, k + 1)) | {k}, reverse=True)
    X = torch.ones((m.ncols, k), dtype=m.vec_dtype, device="cuda")
    Y = torch.zeros((m.nrows, k), dtype=m.vec_dtype, device="cuda")
    res = {}
    for kt in cands:
        m.spmm(X, Y, kt=kt)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            m.spmm(X, Y, kt=kt)
        e1.record()
        torch.cuda.synchronize()
        res[kt] = e0.elapsed_time(e1) / reps * 1e-3
    best = min(res, key=res.get)
    return best, res
