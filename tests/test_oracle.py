"""Pins of the CPU oracle (oracle/spc5_oracle.c).  CPU only.

Every oracle function is checked against something it cannot simply
restate:

* spc5o_convert / spc5o_block_count -- byte-exact against the compiled
  reference converter's golden output (tests/golden/small_*.npz: all N+1
  rowsSize entries, headers, permuted values, block counts), against the
  live reference build when present, and against the closed-form block
  counts of the Laplacian, the 27-point stencil and dense matrices.
* spc5o_beta_spmv -- bit-exact against the reference AVX-512 kernel's y in
  the golden files (summation order: per-lane FMA over blocks, then the
  halving reduction tree), and against closed forms (Laplacian x ones,
  the reference driver's dense i+j matrix).
* spc5o_csr_spmv -- against the reference CSR y (normwise 1e-14) and a
  dense numpy product.
* spc5o_decode -- decode(convert(A)) reproduces the CSR triplets.
* spc5o_thread_split / spc5o_find_best -- equal to the reference's outputs
  in the golden files.
* spc5o_coo_sorted_to_csr -- against numpy's bincount/cumsum.
"""
import numpy as np
import pytest

import oracle as O
from conftest import golden_cases, load_golden, normwise_error
from matgen import clustered_csr, dense_csr, laplacian2d, random_csr

RS = (1, 2, 4, 8)
DT = {"f64": np.float64, "f32": np.float32}


@pytest.mark.parametrize("case", golden_cases())
@pytest.mark.parametrize("tag", ["f64", "f32"])
@pytest.mark.parametrize("r", RS)
def test_convert_matches_reference_golden(case, tag, r):
    g = load_golden(case)
    rp, col = g["rowptr"], g["col"]
    v = g["val64"].astype(DT[tag])
    n = len(rp) - 1
    bptr, hdr, vout, nb = O.convert(r, rp, col, v)
    assert nb == int(g[f"nb_{tag}_r{r}"]) == int(g[f"bc_{tag}_r{r}"])
    nbr = (n + r - 1) // r
    rs_ref = g[f"rs_{tag}_r{r}"]
    assert np.array_equal(bptr, rs_ref[: nbr + 1])
    # entries past ceil(N/r) are the stale CSR row pointer
    assert np.array_equal(rs_ref[nbr + 1:], rp[nbr + 1:])
    assert np.array_equal(hdr, g[f"hdr_{tag}_r{r}"])
    assert np.array_equal(vout.view(np.uint8), g[f"val_{tag}_r{r}"].view(np.uint8))
    assert O.block_count(r, O.VS_OF[np.dtype(DT[tag])], rp, col) == nb


@pytest.mark.parametrize("case", golden_cases())
@pytest.mark.parametrize("tag", ["f64", "f32"])
@pytest.mark.parametrize("r", RS)
def test_beta_spmv_bitexact_vs_reference_kernel(case, tag, r):
    g = load_golden(case)
    if int(g[f"nb_{tag}_r{r}"]) == 0:
        pytest.skip("no blocks")
    n = len(g["rowptr"]) - 1
    x = g["x64"] if tag == "f64" else g["x32"]
    y = (g["y064"] if tag == "f64" else g["y032"]).copy()
    O.beta_spmv(r, n, g[f"rs_{tag}_r{r}"], g[f"hdr_{tag}_r{r}"], g[f"val_{tag}_r{r}"], x, y)
    assert np.array_equal(y.view(np.uint8), g[f"y_{tag}_r{r}"].view(np.uint8))


@pytest.mark.parametrize("case", golden_cases())
@pytest.mark.parametrize("tag", ["f64", "f32"])
def test_csr_spmv_vs_reference_and_dense(case, tag):
    g = load_golden(case)
    rp, col = g["rowptr"], g["col"]
    v = g["val64"].astype(DT[tag])
    x = g["x64"] if tag == "f64" else g["x32"]
    y0 = g["y064"] if tag == "f64" else g["y032"]
    n = len(rp) - 1
    y = y0.copy()
    O.csr_spmv(rp, col, v, x, y)
    tol = 1e-14 if tag == "f64" else 1e-6
    assert normwise_error(y[:n], g[f"ycsr_{tag}"][:n], rp, v, x, y0) <= tol
    dense = np.zeros((n, int(g["ncols"])))
    for i in range(n):
        dense[i, col[rp[i]:rp[i + 1]]] = v[rp[i]:rp[i + 1]]
    yd = y0[:n].astype(np.float64) + dense @ x[: int(g["ncols"])].astype(np.float64)
    assert normwise_error(y[:n], yd, rp, v, x, y0) <= tol


@pytest.mark.parametrize("case", golden_cases())
@pytest.mark.parametrize("r", RS)
def test_thread_split_matches_reference(case, r):
    g = load_golden(case)
    if f"split_r{r}_t3" not in g:
        pytest.skip("empty matrix")
    rp, col, v = g["rowptr"], g["col"], g["val64"]
    bptr, hdr, _, _ = O.convert(r, rp, col, v)
    n = len(rp) - 1
    for nth in (3, 8):
        st, cnt, voff = O.thread_split(r, 1, n, bptr, hdr, nth)
        assert np.array_equal(np.stack([st, cnt, voff]), g[f"split_r{r}_t{nth}"])


@pytest.mark.parametrize("case", golden_cases())
def test_find_best_matches_reference(case):
    g = load_golden(case)
    rp, col = g["rowptr"], g["col"]
    if "findbest_f64" not in g:
        pytest.skip("empty matrix")
    for tag, key, nth in (("f64", "findbest_f64", 0), ("f32", "findbest_f32", 0),
                          ("f64", "findbest_omp8_f64", 8), ("f64", "findbest_omp36_f64", 36)):
        r, est = O.find_best(tag == "f64", rp, col, nthreads=nth)
        ref_r, ref_est = g[key]
        assert r == int(ref_r)
        assert est == pytest.approx(float(ref_est), rel=1e-12)


@pytest.mark.parametrize("tag", ["f64", "f32"])
@pytest.mark.parametrize("r", RS)
def test_decode_roundtrip(tag, r):
    rp, col, v = clustered_csr(150, 400, seed=r, empty_frac=0.1, dtype=DT[tag])
    n = len(rp) - 1
    bptr, hdr, vout, nb = O.convert(r, rp, col, v)
    rows, cols, vals = O.decode(r, n, bptr, hdr, vout)
    order = np.lexsort((cols, rows))
    rr = np.repeat(np.arange(n), np.diff(rp))
    assert np.array_equal(rows[order], rr)
    assert np.array_equal(cols[order], col)
    assert np.array_equal(vals[order], v)


def test_block_count_closed_forms():
    # 5-point Laplacian n x n: beta(1,8) = 3N - 2n, beta(4,8) = 3N/4 - 2n/4
    n = 32
    rp, col, v = laplacian2d(n)
    N = n * n
    assert O.block_count(1, 8, rp, col) == 3 * N - 2 * n
    assert O.block_count(4, 8, rp, col) == 3 * N // 4 - 2 * n // 4
    # 27-point stencil n^3 (n = 16): n(3n-2)^2, (n/2)(3n-2)^2, (n/4)(3n-2)^2, 2(n/8)(3n-2)^2
    from spc5_amd import synth
    n = 16
    rp, col, v = synth.stencil27(n)
    L2 = (3 * n - 2) ** 2
    assert O.block_count(1, 8, rp, col) == n * L2
    assert O.block_count(2, 8, rp, col) == n // 2 * L2
    assert O.block_count(4, 8, rp, col) == n // 4 * L2
    assert O.block_count(8, 8, rp, col) == 2 * (n // 8) * L2
    # dense n x n with VS | n: n/VS blocks per block-row, every mask full
    rp, col, v = dense_csr(48)
    for r in RS:
        assert O.block_count(r, 8, rp, col) == (48 // r) * (48 // 8)
        assert O.block_count(r, 16, rp, col) == (48 // r) * (48 // 16)


def test_beta_spmv_closed_forms():
    # Laplacian times ones: 0 inside, 1 on edges, 2 at corners (exact)
    n = 24
    rp, col, v = laplacian2d(n)
    N = n * n
    x = np.zeros(N + 15)
    x[:N] = 1.0
    expect = np.zeros((n, n))
    expect[0, :] += 1
    expect[-1, :] += 1
    expect[:, 0] += 1
    expect[:, -1] += 1
    for r in RS:
        bptr, hdr, vout, nb = O.convert(r, rp, col, v)
        y = np.zeros(N + 7)
        O.beta_spmv(r, N, bptr, hdr, vout, x, y)
        assert np.array_equal(y[:N], expect.ravel())
    # reference driver's dense matrix: A[i,j] = i+j, x[j] = j % 98
    n = 40
    rp, col, v = dense_csr(n)
    x = np.zeros(n + 15)
    x[:n] = np.arange(n) % 98
    ii = np.arange(n)[:, None]
    jj = np.arange(n)[None, :]
    yexp = ((ii + jj) * (jj % 98)).sum(axis=1).astype(np.float64)
    for r in RS:
        bptr, hdr, vout, nb = O.convert(r, rp, col, v)
        y = np.zeros(n + 7)
        O.beta_spmv(r, n, bptr, hdr, vout, x, y)
        assert np.array_equal(y[:n], yexp)


def test_coo_sorted_to_csr_vs_numpy():
    rng = np.random.default_rng(4)
    n = 50
    i = np.sort(rng.integers(0, n, 400))
    j = rng.integers(0, 70, 400)
    order = np.lexsort((j, i))
    i, j = i[order], j[order]
    v = rng.uniform(-1, 1, 400)
    rp, c, vv = O.coo_sorted_to_csr(n, i, j, v)
    assert np.array_equal(rp, np.concatenate([[0], np.cumsum(np.bincount(i, minlength=n))]))
    assert np.array_equal(c, j)
    assert np.array_equal(vv, v)


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("tag", ["f64", "f32"])
def test_oracle_vs_live_reference(seed, tag):
    dt = DT[tag]
    vs = O.VS_OF[np.dtype(dt)]
    for rp, col, v in (random_csr(45 + seed, 90, 0.1, seed, dt),
                       clustered_csr(120 + seed, 333, seed, dtype=dt, empty_frac=0.1)):
        n = len(rp) - 1
        nc = int(col.max()) + 1 if len(col) else 1
        rng = np.random.default_rng(seed)
        x = np.zeros(nc + vs - 1, dt)
        x[:nc] = rng.uniform(-1, 1, nc)
        for r in RS:
            b1, h1, v1, nb1 = O.convert(r, rp, col, v)
            rs, h2, v2, nb2 = O.ref_convert(r, rp, col, v, ncols=nc)
            assert nb1 == nb2
            assert np.array_equal(b1, rs[: len(b1)]) and np.array_equal(h1, h2)
            assert np.array_equal(v1.view(np.uint8), v2.view(np.uint8))
            y1 = np.zeros(n + 7, dt)
            y1[:n] = rng.uniform(-1, 1, n)
            y2 = y1.copy()
            O.beta_spmv(r, n, b1, h1, v1, x, y1)
            O.ref_spmv(r, n, rs, h2, v2, x, y2)
            assert np.array_equal(y1.view(np.uint8), y2.view(np.uint8))
