"""Small seeded CSR matrices for the parity tests (numpy only)."""
from __future__ import annotations

import numpy as np


def csr_from_sets(n_rows, cols_per_row, rng, dtype=np.float64, lo=-1.0, hi=1.0):
    rowptr = np.zeros(n_rows + 1, np.int32)
    for i, c in enumerate(cols_per_row):
        rowptr[i + 1] = rowptr[i] + len(c)
    col = np.concatenate([np.asarray(c, np.int32) for c in cols_per_row]) if n_rows else np.zeros(0, np.int32)
    col = col.astype(np.int32)
    val = rng.uniform(lo, hi, size=len(col)).astype(dtype)
    return rowptr, col, val


def random_csr(n_rows, n_cols, density, seed, dtype=np.float64):
    """Uniform random pattern (sorted unique columns per row)."""
    rng = np.random.default_rng(seed)
    sets = []
    for _ in range(n_rows):
        k = rng.binomial(n_cols, density)
        sets.append(np.sort(rng.choice(n_cols, size=k, replace=False)))
    return csr_from_sets(n_rows, sets, rng, dtype)


def clustered_csr(n_rows, n_cols, seed, mean_len=12, max_run=16, spread=None, dtype=np.float64,
                  empty_frac=0.0):
    """Runs of 1..max_run contiguous columns near the diagonal, some empty rows."""
    rng = np.random.default_rng(seed)
    spread = max(4, n_cols // 8) if spread is None else spread
    sets = []
    for i in range(n_rows):
        if rng.random() < empty_frac:
            sets.append(np.zeros(0, np.int64))
            continue
        L = int(min(max(1, rng.lognormal(np.log(mean_len), 0.8)), 4 * mean_len, n_cols))
        cols = set()
        tries = 0
        while len(cols) < L and tries < 64 * L:
            tries += 1
            run = int(rng.integers(1, max_run + 1))
            c0 = int(np.clip(i * n_cols // max(1, n_rows) + rng.integers(-spread, spread + 1), 0, n_cols - 1))
            for c in range(c0, min(n_cols, c0 + run)):
                cols.add(c)
                if len(cols) >= L:
                    break
        sets.append(np.array(sorted(cols), np.int64))
    return csr_from_sets(n_rows, sets, rng, dtype)


def dense_csr(n, dtype=np.float64):
    """The reference driver's dense test matrix: value = i + j (builddense,
    examples/load_mm_and_compare.cpp:431-449)."""
    rowptr = (np.arange(n + 1, dtype=np.int64) * n).astype(np.int32)
    col = np.tile(np.arange(n, dtype=np.int32), n)
    i = np.repeat(np.arange(n), n)
    val = (i + col).astype(dtype)
    return rowptr, col, val


def laplacian2d(n, dtype=np.float64):
    """5-point Laplacian on an n x n grid: diagonal 4, off-diagonal -1."""
    N = n * n
    rows = []
    cols = []
    vals = []
    idx = np.arange(N)
    gi, gj = idx // n, idx % n
    for di, dj, v in ((-1, 0, -1.0), (0, -1, -1.0), (0, 0, 4.0), (0, 1, -1.0), (1, 0, -1.0)):
        ok = (gi + di >= 0) & (gi + di < n) & (gj + dj >= 0) & (gj + dj < n)
        rows.append(idx[ok])
        cols.append((idx + di * n + dj)[ok])
        vals.append(np.full(ok.sum(), v))
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = np.concatenate(vals)
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    rowptr = np.zeros(N + 1, np.int64)
    np.add.at(rowptr, r + 1, 1)
    rowptr = np.cumsum(rowptr).astype(np.int32)
    return rowptr, c.astype(np.int32), v.astype(dtype)
