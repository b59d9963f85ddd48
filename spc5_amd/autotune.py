"""Autotuner of the SpMV and its batching form: per matrix, r (rows per
block), the tile size of the product, the persistent wave count of a batch,
and the column chunk and record kind of multi-right-hand-side products.

r is chosen from a cost model evaluated on GPU block counts (one count pass
per candidate, no conversion), one set of non-negative costs per r:

    t(r) = a_r * nnz + b_r * nblocks(r) + c_r * batches(r)

batches(r) = max(nblocks/64, nnz/cap) counts the kernel's batches (64 blocks
or `cap` values each).  a_r prices the value stream and the per-value
expansion and row sums of r's kernel, b_r the block headers and x
requests, c_r the fixed work of a batch.  The costs are fitted by
non-negative least squares (`fit_coefficients`) to timed products of
synthetic matrices on the device and stored, with the raw timings, in
tuning_mi355x.json (`calibrate`).  Left out of the fit one matrix at a time
(`leave_one_out`), the model picks the measured best r for 20 of the 24
calibration matrices, 0.7 % slower than the best on average and 10.5 % at
worst (a single cost set for all r, fitted by plain least squares with
negative costs, managed 18).  Without a calibration file it reduces to
bytes at 5.5 TB/s plus a fixed per-block decode cost.  The reference's own
CPU model (SPC5_find_best, cubic in values per block) is kept in
spc5_amd.spc5 for API parity; its coefficients describe a CPU and are not
used here.

`mode="measure"` converts the matrix with each candidate r and times it
instead (exact, costs four conversions).  `tune_tiles` times candidate tile
sizes of one matrix's product, `tune_batch_waves` candidate persistent wave
counts of a batch, `tune_spmm_kt` the column chunks of a product with k
right-hand sides, and `tune_spmm` picks r, the chunk and the record kind
(per value or per block column) by timing.
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


def spmv_bytes(nrows, ncols, nnz, nb, r, value_bytes, vec_bytes, vs):
    return (nnz * value_bytes + nb * (4 + r * mask_bytes(vs)) + ((nrows + r - 1) // r + 1) * 4
            + ncols * vec_bytes + 2 * nrows * vec_bytes)


def batch_cap(value_bytes):
    """Values per batch of the SpMV tile for values of this width."""
    return 384 if value_bytes == 8 else 512


def features(nrows, ncols, nnz, nb, r, value_bytes, vec_bytes, vs):
    """The model's terms of one (matrix, r): [nnz, nblocks, batches] in the
    slots of r, zeros in the slots of the other r (3 per r, r = 1, 2, 4, 8)."""
    batches = max(nb / 64.0, nnz / float(batch_cap(value_bytes)))
    out = np.zeros(3 * len(RS), dtype=np.float64)
    i = RS.index(r)
    out[3 * i: 3 * i + 3] = (float(nnz), float(nb), batches)
    return out


def _default_coefficients():
    # bytes at 5.5 TB/s (values 8 B, headers ~5 B per block) plus 12 ps per block
    return {str(r): {"value": 8.0 / 5.5e12, "block": (4 + r) / 5.5e12 + 1.2e-11, "batch": 0.0} for r in RS}


def load_coefficients(path: str = DEFAULT_FILE):
    if os.path.exists(path):
        with open(path) as f:
            d = json.load(f)
        c = d.get("coefficients")
        if isinstance(c, dict) and all(str(r) in c for r in RS):
            return c
    return _default_coefficients()


def _model_vector(coeffs):
    return np.array([coeffs[str(r)][k] for r in RS for k in ("value", "block", "batch")], dtype=np.float64)


def _measurement_features(m):
    return features(m["nrows"], m.get("ncols", m["nrows"]), m["nnz"], m["nblocks"], m["r"], 8, 8, 8)


def fit_coefficients(measurements):
    """Non-negative least-squares fit of the per-r costs to timed products:
    `measurements` is a list of {r, nrows, ncols?, nnz, nblocks, seconds}
    (double values).  Columns are scaled to unit maximum for the solve."""
    from scipy.optimize import nnls
    X = np.array([_measurement_features(m) for m in measurements])
    y = np.array([m["seconds"] for m in measurements], dtype=np.float64)
    scale = np.abs(X).max(axis=0)
    scale[scale == 0] = 1.0
    w, _ = nnls(X / scale, y)
    w = w / scale
    return {str(r): {"value": float(w[3 * i]), "block": float(w[3 * i + 1]), "batch": float(w[3 * i + 2])}
            for i, r in enumerate(RS)}


def leave_one_out(measurements):
    """Fit without each matrix in turn and compare the model's pick of r for
    it with the measured best.  `measurements` holds every r of a matrix
    (matrices are told apart by nrows, nnz).  Returns {agree, matrices,
    mean_loss, worst_loss} (loss: the pick's time over the best, minus 1)."""
    groups = {}
    for m in measurements:
        groups.setdefault((m["nrows"], m["nnz"]), []).append(m)
    keys = list(groups)
    agree, losses = 0, []
    for k in keys:
        coeffs = fit_coefficients([m for kk in keys if kk != k for m in groups[kk]])
        w = _model_vector(coeffs)
        pred = {m["r"]: float(np.dot(w, _measurement_features(m))) for m in groups[k]}
        meas = {m["r"]: m["seconds"] for m in groups[k]}
        rp, rm = min(pred, key=pred.get), min(meas, key=meas.get)
        agree += rp == rm
        losses.append(meas[rp] / meas[rm] - 1.0)
    return {"agree": agree, "matrices": len(keys), "mean_loss": float(np.mean(losses)),
            "worst_loss": float(np.max(losses))}


def block_counts(rowptr, col, vs: int = 8):
    """{r: number of beta(r,VS) blocks} from the GPU count pass."""
    return {r: block_count(rowptr, col, r, vs) for r in RS}


def predict(counts, nrows, ncols, nnz, value_bytes=8, vec_bytes=8, vs=8, coeffs=None):
    w = _model_vector(coeffs or load_coefficients())
    out = {}
    for r, nb in counts.items():
        out[r] = float(np.dot(w, features(nrows, ncols, nnz, nb, r, value_bytes, vec_bytes, vs)))
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
    raw = []
    for rp, col, val in matrices:
        n = rp.numel() - 1
        for r in RS:
            m = DeviceMatrix.from_csr(rp, col, val, r)
            t = _time_spmv(m, reps=reps)
            raw.append({"r": r, "nrows": n, "nnz": int(val.numel()), "nblocks": int(m.nblocks),
                        "ntiles": int(m.info()["ntiles"]), "seconds": t})
            del m
    coeffs = fit_coefficients(raw)
    with open(path, "w") as f:
        json.dump({"device": torch.cuda.get_device_name(0), "time": time.time(),
                   "model": "t(r) = value_r*nnz + block_r*nblocks(r) + batch_r*batches(r), non-negative least squares",
                   "coefficients": coeffs, "leave_one_out": leave_one_out(raw), "samples": len(matrices),
                   "measurements": raw}, f, indent=1)
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


def tune_spmm(mats, k: int, kt_candidates=None, reps=5):
    """Fastest schedule for Y += A X with k right-hand sides over candidate
    block shapes of one matrix: `mats` maps r -> DeviceMatrix (the same
    matrix converted with that r).  Tries every r, column chunk kt and
    record kind: one record per stored value, or (r > 1) one per set column
    of a block (SPC5_HIP_OPT_SPMM_COLUMNS), and X requests of one element or
    16 bytes per lane (SPC5_HIP_OPT_SPMM_X16).  Leaves each matrix set to
    its best record kind.  Returns ((r, kt, columns, x16),
    {(r, kt, columns, x16): seconds})."""
    times = {}
    for r, m in mats.items():
        best_kind, best_t = (False, False), None
        kinds = ((False, False), (False, True)) + (((True, False), (True, True)) if r > 1 else ())
        for cols, x16 in kinds:
            m.spmm_columns = cols
            m.spmm_x16 = x16
            _, res = tune_spmm_kt(m, k, kt_candidates, reps)
            for kt, t in res.items():
                times[(r, kt, cols, x16)] = t
                if best_t is None or t < best_t:
                    best_t, best_kind = t, (cols, x16)
        m.spmm_columns, m.spmm_x16 = best_kind
    best = min(times, key=times.get)
    return best, times


def _time_fn(fn, reps=10, warm=2):
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps * 1e-3


def tune_tiles(m: DeviceMatrix, candidates=(0, 1024, 1536, 2048, 3072, 4096, 8192), reps=20):
    """Fastest tile size (values per wave's tile, SPC5_HIP_OPT_TILE_VALUES;
    0 is the default rule) for the matrix's product; leaves the matrix set
    to it.  Returns (best, {candidate: seconds})."""
    x = torch.ones(m.ncols, dtype=m.vec_dtype, device="cuda")
    y = torch.zeros(m.nrows, dtype=m.vec_dtype, device="cuda")
    times = {}
    for t in candidates:
        m.tile_values = t
        times[t] = _time_fn(lambda: m.spmv(x, y), reps)
    best = min(times, key=times.get)
    m.tile_values = best
    return best, times


def tune_batch_tiles(mats, make_batch, candidates=(0, 2048, 4096, 8192, 16384), reps=3):
    """Fastest tile size for a batch: every matrix of `mats` is set to each
    candidate (SPC5_HIP_OPT_TILE_VALUES; 0 is each matrix's default rule),
    `make_batch()` builds a batch.Batch of them, and one graph launch is
    timed.  A batch gathers its parallelism from many matrices, so it can
    use larger tiles than one matrix of its size alone (fewer tiles, fewer x
    windows re-read across tiles): on 2,000 banded matrices of 6.0 G values
    in double, 17.9 ms with the default rule's 512-value tiles and 15.5 ms
    with 8,192.  Leaves every matrix set to the best candidate.  Returns
    (best, {candidate: seconds}); the batches built here are closed."""
    times = {}
    for tv in candidates:
        for m in mats:
            m.tile_values = tv
        b = make_batch()
        times[tv] = _time_fn(lambda: b.spmv(graph=True), reps, warm=1)
        b.close()
    best = min(times, key=times.get)
    for m in mats:
        m.tile_values = best
    return best, times


def tune_batch_waves(make_batch, fractions=(0.5, 1.0, 2.0, 4.0), reps=5):
    """Fastest persistent wave count of a batch: `make_batch(waves)` builds
    a batch.Batch (waves = 0 is the kernel's resident count); candidates are
    fractions of that count and one wave per tile (the tiles then dealt out
    by the hardware's dispatcher, which measured fastest on 2,000 banded
    matrices with large tiles).  Returns (waves, {waves: seconds}); the
    batches built here are closed."""
    b0 = make_batch(0)
    base, ntiles = b0.info()["waves"], b0.info()["ntiles"]
    b0.close()
    times = {}
    for w in sorted({max(1, min(ntiles, int(round(base * f)))) for f in fractions} | {max(1, ntiles)}):
        b = make_batch(w)
        times[b.info()["waves"]] = _time_fn(lambda: b.spmv(graph=True), reps)
        b.close()
    best = min(times, key=times.get)
    return best, times


def calibration_set():
    """Seeded synthetic device CSR matrices for `calibrate`: banded (the
    batch's family) over bandwidths and fills, clustered runs, the 27-point
    stencil and the 5-point Laplacian.  Each is large enough (>= 260K rows)
    that kernel time, not launch overhead, is what gets fitted."""
    from . import synth
    mats = []
    for n in (400_000, 2_000_000):
        for bw, fill in ((4, 1.0), (8, 0.5), (16, 0.3), (32, 1.0), (64, 0.3), (128, 0.3), (128, 0.05),
                         (256, 0.3), (256, 0.1)):
            mats.append(synth.banded_gpu(n, bw, fill, seed=7 + bw + n))
    for rp, col, val in (synth.clustered_runs(1 << 20), synth.clustered_runs(1 << 22), synth.stencil27(64),
                         synth.stencil27(128), synth.laplacian2d(1024), synth.laplacian2d(2048)):
        mats.append(tuple(torch.from_numpy(a).cuda() for a in (rp, col, val)))
    return mats


def main(argv=None):
    """python -m spc5_amd.autotune --calibrate [--out FILE]: fit the model on
    this GPU, then print the model's choice of r against the measured best
    for every matrix of the set."""
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--calibrate", action="store_true")
    ap.add_argument("--out", default=DEFAULT_FILE)
    args = ap.parse_args(argv)
    if not args.calibrate:
        ap.error("nothing to do (use --calibrate)")
    mats = calibration_set()
    coeffs = calibrate(mats, args.out)
    print(json.dumps(coeffs, indent=1), flush=True)
    agree = 0
    for i, (rp, col, val) in enumerate(mats):
        r_model, det = choose_r(rp, col, val, mode="model", coeffs=coeffs)
        r_meas, mdet = choose_r(rp, col, val, mode="measure")
        t = mdet["measured_s"]
        agree += r_model == r_meas
        print(f"matrix {i}: rows {rp.numel() - 1} nnz {val.numel()} model r={r_model} measured r={r_meas} "
              f"(model pick {t[r_model] * 1e3:.4f} ms, best {t[r_meas] * 1e3:.4f} ms)", flush=True)
    print(f"model agrees with measurement on {agree} of {len(mats)}", flush=True)


if __name__ == "__main__":
    main()
