"""Writes the golden vectors of tests/golden/ from the compiled reference.

Run in a container that has the reference tree and a build of
oracle/_ref/libspc5_ref.so (``make -C oracle``):

    python tests/golden/make_golden.py

Outputs (data only: inputs and the reference's outputs):

* ``small_<case>.npz`` -- per small matrix: the CSR input, and for double
  and float and r in {1,2,4,8} the reference converter's rowsSize (all N+1
  entries as left in place), headers, permuted values and block count; the
  reference AVX-512 kernel's y for a fixed x and y0; the reference CSR y;
  block counts; the SPC5_find_best choices; the OpenMP split of a double
  matrix for 3 and 8 threads.
* ``lap512.json`` -- the 512x512 5-point Laplacian (N=262,144): block counts
  and SHA-256 digests of the reference converter's rowsSize / headers /
  values for beta(1,8) and beta(4,8) double, and of the reference kernel's y
  for x = U[-1,1] (splitmix64 seed 0) and y0 = 0.
* ``big_digests.json`` -- block counts and SHA-256 digests of the reference
  converter's output at full size: the clustered-runs matrix (4,194,304
  rows, double beta(r,8) and float beta(r,16), every r) and the 27-point
  stencil 160^3 (double beta(2,8), beta(4,8)).

``python tests/golden/make_golden.py [small] [lap512] [big]`` writes a
subset.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle as O  # noqa: E402
from matgen import clustered_csr, dense_csr, laplacian2d, random_csr  # noqa: E402

RS = (1, 2, 4, 8)
# Seed of each small case's x and y0 (fixed, so that rerunning this script
# rewrites the same vectors).
SEEDS = {"lap8": 1001, "dense16": 1002, "rand37x70": 1003, "clustered101x300": 1004, "rows13": 1005,
         "onerow200": 1006, "empty5": 1007, "dense_runs64": 1008}


def cases():
    yield "lap8", laplacian2d(8), 64
    yield "dense16", dense_csr(16), 16
    rp, c, v = random_csr(37, 70, 0.12, seed=11)
    yield "rand37x70", (rp, c, v), 70
    rp, c, v = clustered_csr(101, 300, seed=5, empty_frac=0.15)
    yield "clustered101x300", (rp, c, v), 300
    rp, c, v = clustered_csr(13, 40, seed=9, mean_len=6)
    yield "rows13", (rp, c, v), 40
    rng = np.random.default_rng(21)
    cols = np.sort(rng.choice(1000, 200, replace=False)).astype(np.int32)
    yield "onerow200", (np.array([0, 200], np.int32), cols, rng.uniform(-1, 1, 200)), 1000
    yield "empty5", (np.zeros(6, np.int32), np.zeros(0, np.int32), np.zeros(0)), 5
    rp, c, v = clustered_csr(64, 64, seed=3, mean_len=30, max_run=16, spread=8)
    yield "dense_runs64", (rp, c, v), 64


def make_small():
    for name, (rp, col, val64), ncols in cases():
        out = {"rowptr": rp, "col": col, "val64": val64.astype(np.float64), "ncols": np.int64(ncols)}
        n = len(rp) - 1
        rng = np.random.default_rng(SEEDS[name])
        x64 = np.zeros(ncols + 15)
        x64[:ncols] = rng.uniform(-1, 1, ncols)
        y064 = np.zeros(n + 7)
        y064[:n] = rng.uniform(-1, 1, n)
        out["x64"], out["y064"] = x64, y064
        out["x32"], out["y032"] = x64.astype(np.float32), y064.astype(np.float32)
        for tag, dt in (("f64", np.float64), ("f32", np.float32)):
            v = val64.astype(dt)
            x = out["x64"] if dt == np.float64 else out["x32"]
            y0 = out["y064"] if dt == np.float64 else out["y032"]
            ycsr = y0.copy()
            O.ref_csr_spmv(rp, col, v, x, ycsr, ncols=ncols)
            out[f"ycsr_{tag}"] = ycsr
            for r in RS:
                rs, hdr, vout, nb = O.ref_convert(r, rp, col, v, ncols=ncols)
                out[f"rs_{tag}_r{r}"] = rs
                out[f"hdr_{tag}_r{r}"] = hdr
                out[f"val_{tag}_r{r}"] = vout
                out[f"nb_{tag}_r{r}"] = np.int64(nb)
                out[f"bc_{tag}_r{r}"] = np.int64(O.ref_block_count(r, rp, col, v))
                y = y0.copy()
                if nb > 0:
                    O.ref_spmv(r, n, rs, hdr, vout, x, y)
                out[f"y_{tag}_r{r}"] = y
            if len(v) > 0:
                fr, fe = O.ref_find_best(dt == np.float64, rp, col, v)
                out[f"findbest_{tag}"] = np.array([fr, fe])
        if len(val64) > 0:
            for nth in (8, 36):
                fr, fe = O.ref_find_best(True, rp, col, val64, nthreads=nth)
                out[f"findbest_omp{nth}_f64"] = np.array([fr, fe])
            for r in RS:
                for nth in (3, 8):
                    st, cnt, voff = O.ref_thread_split(r, rp, col, val64, nth)
                    out[f"split_r{r}_t{nth}"] = np.stack([st, cnt, voff])
        np.savez_compressed(os.path.join(HERE, f"small_{name}.npz"), **out)
        print("wrote", name, flush=True)


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).view(np.uint8).tobytes()).hexdigest()


def make_lap512():
    from spc5_amd import synth
    rp, col, val = synth.laplacian2d(512)
    n = len(rp) - 1
    x = np.zeros(n + 15)
    x[:n] = synth.uniform(n, 0)
    res = {"n": n, "nnz": int(rp[-1]), "x_seed": 0}
    for r in (1, 4):
        rs, hdr, vout, nb = O.ref_convert(r, rp, col, val)
        y = np.zeros(n + 7)
        O.ref_spmv(r, n, rs, hdr, vout, x, y)
        res[f"r{r}"] = {"nblocks": nb, "rowsSize_sha256": sha(rs), "headers_sha256": sha(hdr),
                        "values_sha256": sha(vout), "y_sha256": sha(y[:n]),
                        "y_inf": float(np.abs(y[:n]).max())}
    with open(os.path.join(HERE, "lap512.json"), "w") as f:
        json.dump(res, f, indent=1)
    print("wrote lap512", flush=True)


def make_big():
    """Full-size digests of the reference converter on the two large
    workloads: the clustered-runs matrix (N = 4,194,304, double beta(r,8)
    and float beta(r,16), every r) and the 27-point stencil 160^3 (double
    beta(2,8) and beta(4,8)).  rowsSize is hashed over its first
    ceil(N/r)+1 entries (the block-row pointers; the reference leaves the
    rest of the N+1 array as it was)."""
    from spc5_amd import synth
    res = {}
    rp, col, val = synth.clustered_runs(4194304, seed=1, value_seed=2)
    n = len(rp) - 1
    cl = {"n": n, "nnz": int(rp[-1]), "generator": "clustered_runs(4194304, seed=1, value_seed=2)"}
    for tag, dt in (("f64", np.float64), ("f32", np.float32)):
        v = val.astype(dt)
        for r in RS:
            rs, hdr, vout, nb = O.ref_convert(r, rp, col, v)
            nbr = (n + r - 1) // r
            cl[f"{tag}_r{r}"] = {"nblocks": nb, "rowsSize_sha256": sha(rs[: nbr + 1]), "headers_sha256": sha(hdr),
                                 "values_sha256": sha(vout)}
            print("clustered", tag, r, nb, flush=True)
            del rs, hdr, vout
        del v
    res["clustered"] = cl
    del rp, col, val
    rp, col, val = synth.stencil27(160)
    n = len(rp) - 1
    st = {"n": n, "nnz": int(rp[-1]), "generator": "stencil27(160)"}
    for r in (2, 4):
        rs, hdr, vout, nb = O.ref_convert(r, rp, col, val)
        nbr = (n + r - 1) // r
        st[f"f64_r{r}"] = {"nblocks": nb, "rowsSize_sha256": sha(rs[: nbr + 1]), "headers_sha256": sha(hdr),
                           "values_sha256": sha(vout)}
        print("stencil27", r, nb, flush=True)
        del rs, hdr, vout
    res["stencil27"] = st
    with open(os.path.join(HERE, "big_digests.json"), "w") as f:
        json.dump(res, f, indent=1)
    print("wrote big_digests", flush=True)


if __name__ == "__main__":
    if not O.ref_available():
        sys.exit("oracle/_ref/libspc5_ref.so is missing: run `make -C oracle` where the reference tree exists")
    which = sys.argv[1:] or ["small", "lap512", "big"]
    if "small" in which:
        make_small()
    if "lap512" in which:
        make_lap512()
    if "big" in which:
        make_big()
