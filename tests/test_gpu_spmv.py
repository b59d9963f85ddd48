"""GPU parity of the beta(r,VS) converter and SpMV (runs the HIP library).

Bars (written in each test):
* converter: byte-identical to the reference converter (golden files,
  SHA-256 digests at the 512x512 Laplacian) and to the oracle;
* SpMV: max|y - y_ref| <= tol * max(||A||_inf ||x||_inf, ||y0||_inf) with
  tol = 1e-12 (double) and 1e-5 (float), y_ref from the reference kernel
  (golden) or the oracle restatement (pinned bit-exactly to it).
"""
import hashlib
import json
import os

import numpy as np
import pytest
import torch

import oracle as O
from conftest import GOLDEN, golden_cases, load_golden, normwise_error
from matgen import clustered_csr, dense_csr, laplacian2d, random_csr

pytestmark = pytest.mark.gpu

RS = (1, 2, 4, 8)
DT = {"f64": np.float64, "f32": np.float32}
TOL = {"f64": 1e-12, "f32": 1e-5}


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from spc5_amd import lib
    lib()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).view(np.uint8).tobytes()).hexdigest()


def gpu_spmv_from_arrays(r, n, ncols, bptr, hdr, val, x, y0, vs=None, row_sort=False):
    from spc5_amd.device import DeviceMatrix
    nb = int(bptr[(n + r - 1) // r])
    vs = O.VS_OF[np.dtype(val.dtype)] if vs is None else vs
    m = DeviceMatrix(r=r, vs=vs, nrows=n, ncols=ncols, nnz=len(val), nblocks=nb,
                     rows_size=dev(np.ascontiguousarray(bptr[: (n + r - 1) // r + 1], np.int32)),
                     headers=dev(hdr.astype(np.uint8)), values=dev(val))
    m.row_sort = row_sort
    xd = dev(x[:ncols])
    yd = dev(y0[:n])
    m.spmv(xd, yd)
    torch.cuda.synchronize()
    return host(yd)


@pytest.mark.parametrize("case", golden_cases())
@pytest.mark.parametrize("tag", ["f64", "f32"])
@pytest.mark.parametrize("r", RS)
def test_gpu_converter_matches_reference_bytes(case, tag, r):
    from spc5_amd.device import block_count, convert_csr
    g = load_golden(case)
    rp, col = g["rowptr"], g["col"]
    v = g["val64"].astype(DT[tag])
    n = len(rp) - 1
    nbr = (n + r - 1) // r
    rs, hdr, vout, nb = convert_csr(dev(rp), dev(col), dev(v), r)
    assert nb == int(g[f"nb_{tag}_r{r}"])
    assert np.array_equal(host(rs), g[f"rs_{tag}_r{r}"][: nbr + 1])
    assert np.array_equal(host(hdr), g[f"hdr_{tag}_r{r}"])
    assert np.array_equal(host(vout).view(np.uint8), g[f"val_{tag}_r{r}"].view(np.uint8))
    assert block_count(dev(rp), dev(col), r, O.VS_OF[np.dtype(DT[tag])]) == int(g[f"bc_{tag}_r{r}"])


def _stage_overflow_csr(kind, dt, seed):
    """CSR matrices of 40 groups of 64 rows (the converter's groups) in
    which one group outgrows the LDS stage the launch sizes from the mean
    group: "values" gives group 1 far more values than the others; "headers"
    gives group 1 as many values as the others but one block per value
    (scattered columns), so only its header bytes outgrow the stage."""
    rng = np.random.default_rng(seed)
    n, ncols = 64 * 40, 20000
    sets = []
    for i in range(n):
        g = i // 64
        base = (i * 7) % (ncols - 400)
        if kind == "values" and g == 1:
            sets.append(np.arange(base, base + 300, dtype=np.int64))
        elif kind == "headers" and g == 1:
            sets.append(np.arange(16, dtype=np.int64) * 33 + (i % 5))  # 16 scattered columns
        elif kind == "headers":
            sets.append(np.arange(base, base + 16, dtype=np.int64))  # two full runs of 8
        else:
            sets.append(np.sort(rng.choice(np.arange(base, base + 40), 4, replace=False)))
    rp = np.zeros(n + 1, np.int32)
    for i, c in enumerate(sets):
        rp[i + 1] = rp[i] + len(c)
    col = np.concatenate(sets).astype(np.int32)
    v = rng.uniform(-1, 1, len(col)).astype(dt)
    return rp, col, v, ncols


@pytest.mark.parametrize("kind", ["values", "headers"])
@pytest.mark.parametrize("tag", ["f64", "f32"])
@pytest.mark.parametrize("r", RS)
def test_gpu_converter_stage_overflow_bytes(kind, tag, r):
    """Groups too large for the converter's LDS stage (the global-memory
    path) next to staged ones: bytes equal to the oracle's converter, and the
    count pass equal to its block count."""
    from spc5_amd.device import block_count, convert_csr
    rp, col, v, ncols = _stage_overflow_csr(kind, DT[tag], r)
    n = len(rp) - 1
    nbr = (n + r - 1) // r
    bptr, hdr, vout, nb = O.convert(r, rp, col, v)
    rs_g, hdr_g, vout_g, nb_g = convert_csr(dev(rp), dev(col), dev(v), r)
    assert nb_g == nb
    assert np.array_equal(host(rs_g), np.asarray(bptr[: nbr + 1], np.int32))
    assert np.array_equal(host(hdr_g), np.asarray(hdr, np.uint8))
    assert np.array_equal(host(vout_g).view(np.uint8), np.ascontiguousarray(vout).view(np.uint8))
    assert block_count(dev(rp), dev(col), r, O.VS_OF[np.dtype(DT[tag])]) == nb


def _rows_csr(rows):
    rp = np.zeros(len(rows) + 1, np.int32)
    for i, c in enumerate(rows):
        rp[i + 1] = rp[i] + len(c)
    col = np.concatenate([np.asarray(c, np.int64) for c in rows]).astype(np.int32)
    return rp, col


@pytest.mark.parametrize("r", RS)
@pytest.mark.parametrize("big", [False, True])
def test_gpu_converter_column_precondition(r, big):
    """Columns must be non-negative and strictly increasing inside each row:
    a decrease, a repeat or a negative column anywhere is refused by the
    count pass; a drop across a row boundary (empty rows between included)
    is not.  `big` puts the rows in a 64-row group too large for the LDS
    stage (the global-memory walk) beside small groups."""
    from spc5_amd import Spc5Error
    from spc5_amd.device import block_count, convert_csr
    rng = np.random.default_rng(r)
    base = []
    for i in range(64 * 6):
        if i % 7 == 3:
            base.append([])
        elif big and 64 <= i < 128:
            base.append(list(range(i, i + 400)))
        else:
            base.append(sorted(rng.choice(np.arange(50, 2000), 12, replace=False).tolist()))
    row = 70  # inside the big group when big
    good_rp, good_col = _rows_csr(base)
    v = rng.uniform(-1, 1, len(good_col))
    assert block_count(dev(good_rp), dev(good_col), r) > 0
    convert_csr(dev(good_rp), dev(good_col), dev(v), r)
    for fault in ("swap", "repeat", "negative_first", "negative_later"):
        rows = [list(c) for c in base]
        c = rows[row]
        if fault == "swap":
            c[4], c[5] = c[5], c[4]
        elif fault == "repeat":
            c[6] = c[5]
        elif fault == "negative_first":
            c[0] = -3
        else:
            c[7] = -1
        rp, col = _rows_csr(rows)
        with pytest.raises(Spc5Error, match="increasing"):
            block_count(dev(rp), dev(col), r)
        with pytest.raises(Spc5Error, match="increasing"):
            convert_csr(dev(rp), dev(col), dev(v[: len(col)]), r)
    # the last column of a row above the first of the next: accepted
    rows = [list(c) for c in base]
    rows[row] = [1990, 1995]
    rows[row + 1] = [0, 1]
    rp, col = _rows_csr(rows)
    assert block_count(dev(rp), dev(col), r) > 0


@pytest.mark.parametrize("case", golden_cases())
@pytest.mark.parametrize("tag", ["f64", "f32"])
@pytest.mark.parametrize("r", RS)
def test_gpu_spmv_matches_reference_kernel(case, tag, r):
    g = load_golden(case)
    rp = g["rowptr"]
    n = len(rp) - 1
    ncols = int(g["ncols"])
    x = g["x64"] if tag == "f64" else g["x32"]
    y0 = g["y064"] if tag == "f64" else g["y032"]
    y = gpu_spmv_from_arrays(r, n, ncols, g[f"rs_{tag}_r{r}"], g[f"hdr_{tag}_r{r}"], g[f"val_{tag}_r{r}"], x, y0)
    err = normwise_error(y, g[f"y_{tag}_r{r}"][:n], rp, g["val64"], x[:ncols], y0[:n])
    assert err <= TOL[tag], err


@pytest.mark.parametrize("case", golden_cases())
@pytest.mark.parametrize("tag", ["f64", "f32"])
def test_reference_c_abi_on_gpu(case, tag):
    """The reference's own entry points (core_SPC5_<r>rVc_Spmv_<T>) through
    the Python mirror: convert in place on the GPU, then y += A x."""
    import spc5_amd.spc5 as S
    g = load_golden(case)
    rp, col = g["rowptr"], g["col"]
    n = len(rp) - 1
    ncols = int(g["ncols"])
    x = (g["x64"] if tag == "f64" else g["x32"]).copy()
    y0 = g["y064"] if tag == "f64" else g["y032"]
    for r, fmt in zip(RS, (S.SPC5_MATRIX_TYPE.FORMAT_1rVc_WT, S.SPC5_MATRIX_TYPE.FORMAT_2rVc,
                           S.SPC5_MATRIX_TYPE.FORMAT_4rVc, S.SPC5_MATRIX_TYPE.FORMAT_8rVc)):
        m = S.SPC5Mat(DT[tag])
        m.format = S.SPC5_MATRIX_TYPE.FORMAT_CSR
        m.numberOfRows, m.numberOfColumns, m.numberOfNNZ = n, ncols, len(col)
        m.rowsSize, m.valuesColumnIndexes = rp.copy(), col.copy()
        m.values = g["val64"].astype(DT[tag])
        S.CSR_to_SPC5(m, fmt)
        assert m.format == fmt
        assert np.array_equal(m.rowsSize, g[f"rs_{tag}_r{r}"])  # N+1 entries, stale tail kept
        assert np.array_equal(m.blocksColumnIndexesWithMasks, g[f"hdr_{tag}_r{r}"])
        assert m.numberOfBlocks == int(g[f"nb_{tag}_r{r}"])
        y = y0.copy()
        S.SPC5_Spmv(m, x, y)
        err = normwise_error(y[:n], g[f"y_{tag}_r{r}"][:n], rp, g["val64"], x[:ncols], y0[:n])
        assert err <= TOL[tag], (r, err)


def test_unknown_format_raises():
    import spc5_amd.spc5 as S
    m = S.SPC5Mat(np.float64)
    m.format = S.SPC5_MATRIX_TYPE.FORMAT_CSR
    with pytest.raises(ValueError):
        S.CSR_to_SPC5(m, 42)
    with pytest.raises(ValueError):
        S.SPC5_Spmv(m, np.zeros(1), np.zeros(1))


def _oracle_case(rp, col, v, r, seed, ncols):
    dt = v.dtype
    vs = O.VS_OF[np.dtype(dt)]
    n = len(rp) - 1
    rng = np.random.default_rng(seed)
    x = np.zeros(ncols + vs - 1, dt)
    x[:ncols] = rng.uniform(-1, 1, ncols)
    y0 = np.zeros(n + 7, dt)
    y0[:n] = rng.uniform(-1, 1, n)
    bptr, hdr, vout, nb = O.convert(r, rp, col, v)
    yref = y0.copy()
    if nb:
        O.beta_spmv(r, n, bptr, hdr, vout, x, yref)
    return x, y0, bptr, hdr, vout, yref


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("tag", ["f64", "f32"])
@pytest.mark.parametrize("r", RS)
@pytest.mark.parametrize("shape", ["rand", "clustered", "tall", "wide"])
@pytest.mark.parametrize("row_sort", [False, True])
def test_gpu_spmv_vs_oracle(seed, tag, r, shape, row_sort):
    dt = DT[tag]
    if shape == "rand":
        rp, col, v = random_csr(333 + seed, 500, 0.02, seed, dt)
        ncols = 500
    elif shape == "clustered":
        rp, col, v = clustered_csr(2003 + seed, 3000, seed, dtype=dt, empty_frac=0.2)
        ncols = 3000
    elif shape == "tall":
        rp, col, v = clustered_csr(1501, 97, seed, mean_len=20, dtype=dt)
        ncols = 97
    else:
        rp, col, v = clustered_csr(7, 20000, seed, mean_len=300, spread=9000, dtype=dt)
        ncols = 20000
    n = len(rp) - 1
    x, y0, bptr, hdr, vout, yref = _oracle_case(rp, col, v, r, seed, ncols)
    y = gpu_spmv_from_arrays(r, n, ncols, bptr, hdr, vout, x, y0, row_sort=row_sort)
    assert normwise_error(y, yref[:n], rp, v, x[:ncols], y0[:n]) <= TOL[tag]


@pytest.mark.parametrize("tag", ["f64", "f32"])
@pytest.mark.parametrize("r", RS)
@pytest.mark.parametrize("row_sort", [False, True])
def test_gpu_long_and_dense_rows(tag, r, row_sort):
    """Block-rows far longer than a tile's value target and than one LDS
    batch; full masks (64 or 128 values per block)."""
    dt = DT[tag]
    rng = np.random.default_rng(r)
    sets = []
    n, ncols = 37, 30000
    for i in range(n):
        if i % 5 == 0:
            sets.append(np.arange(0, 12000, dtype=np.int64))  # dense run
        elif i % 5 == 1:
            sets.append(np.sort(rng.choice(ncols, 9000, replace=False)))
        elif i % 5 == 2:
            sets.append(np.zeros(0, np.int64))
        else:
            sets.append(np.sort(rng.choice(ncols, 40, replace=False)))
    rowptr = np.zeros(n + 1, np.int32)
    for i, c in enumerate(sets):
        rowptr[i + 1] = rowptr[i] + len(c)
    col = np.concatenate(sets).astype(np.int32)
    v = rng.uniform(-1, 1, len(col)).astype(dt)
    x, y0, bptr, hdr, vout, yref = _oracle_case(rowptr, col, v, r, r, ncols)
    y = gpu_spmv_from_arrays(r, n, ncols, bptr, hdr, vout, x, y0, row_sort=row_sort)
    assert normwise_error(y, yref[:n], rowptr, v, x[:ncols], y0[:n]) <= TOL[tag]


@pytest.mark.parametrize("r", RS)
def test_gpu_empty_and_tiny(r):
    for n in (1, 3, 9):
        rp = np.zeros(n + 1, np.int32)
        col = np.zeros(0, np.int32)
        v = np.zeros(0)
        x, y0, bptr, hdr, vout, yref = _oracle_case(rp, col, v, r, 0, 4)
        y = gpu_spmv_from_arrays(r, n, 4, bptr, hdr, vout, x, y0)
        assert np.array_equal(y, y0[:n])
    rp = np.array([0, 1], np.int32)
    col = np.array([0], np.int32)
    v = np.array([2.5])
    x, y0, bptr, hdr, vout, yref = _oracle_case(rp, col, v, r, 0, 1)
    y = gpu_spmv_from_arrays(r, 1, 1, bptr, hdr, vout, x, y0)
    assert y[0] == y0[0] + 2.5 * x[0]


@pytest.mark.parametrize("r", RS)
def test_gpu_closed_forms(r):
    """Laplacian x ones (0 inside, 1 on edges, 2 at corners) and the
    reference driver's dense i+j matrix with x[j] = j % 98: exact."""
    from spc5_amd.device import DeviceMatrix
    n = 64
    rp, col, v = laplacian2d(n)
    N = n * n
    m = DeviceMatrix.from_csr(dev(rp), dev(col), dev(v), r)
    y = torch.zeros(N, dtype=torch.float64, device="cuda")
    m.spmv(torch.ones(N, dtype=torch.float64, device="cuda"), y)
    expect = np.zeros((n, n))
    expect[0, :] += 1
    expect[-1, :] += 1
    expect[:, 0] += 1
    expect[:, -1] += 1
    assert np.array_equal(host(y), expect.ravel())
    n = 96
    rp, col, v = dense_csr(n)
    m = DeviceMatrix.from_csr(dev(rp), dev(col), dev(v), r)
    x = torch.from_numpy((np.arange(n) % 98).astype(np.float64)).cuda()
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    m.spmv(x, y)
    ii = np.arange(n)[:, None]
    jj = np.arange(n)[None, :]
    assert np.array_equal(host(y), ((ii + jj) * (jj % 98)).sum(axis=1).astype(np.float64))


@pytest.mark.parametrize("r", (1, 4))
def test_laplacian512_bytes_match_reference_digests(r):
    """512x512 5-point Laplacian (N=262,144): GPU-converted bytes have the
    reference converter's SHA-256 digests; y within 1e-12 of the oracle."""
    from spc5_amd import synth
    from spc5_amd.device import DeviceMatrix
    ref = json.load(open(os.path.join(GOLDEN, "lap512.json")))
    rp, col, v = synth.laplacian2d(512)
    n = len(rp) - 1
    m = DeviceMatrix.from_csr(dev(rp), dev(col), dev(v), r)
    rr = ref[f"r{r}"]
    assert m.nblocks == rr["nblocks"]
    rs_full = rp.copy()
    rs_full[: (n + r - 1) // r + 1] = host(m.rows_size)
    assert sha(rs_full) == rr["rowsSize_sha256"]
    assert sha(host(m.headers)) == rr["headers_sha256"]
    assert sha(host(m.values)) == rr["values_sha256"]
    x = synth.uniform(n, 0)
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    m.spmv(dev(x), y)
    xp = np.zeros(n + 15)
    xp[:n] = x
    yref = np.zeros(n + 7)
    O.beta_spmv(r, n, host(m.rows_size), host(m.headers), host(m.values), xp, yref)
    assert normwise_error(host(y), yref[:n], rp, v, x, np.zeros(1)) <= 1e-12


@pytest.mark.parametrize("r", RS)
def test_clustered_runs_full_size_properties(r):
    """4,194,304-row clustered-runs matrix (about 134M values), double:
    mask popcount = nnz, linearity A(x1+x2) = A x1 + A x2, and exact CSR
    sums on 2,000 sampled rows (numpy, double) within 1e-12."""
    from spc5_amd import synth
    from spc5_amd.device import DeviceMatrix
    N = 4_194_304
    rp, col, v = synth.clustered_runs(N)
    m = DeviceMatrix.from_csr(dev(rp), dev(col), dev(v), r)
    assert m.info()["nnz"] == len(v)
    x1 = dev(synth.uniform(N, 3))
    x2 = dev(synth.uniform(N, 4))
    y1 = torch.zeros(N, dtype=torch.float64, device="cuda")
    y2 = torch.zeros_like(y1)
    y12 = torch.zeros_like(y1)
    m.spmv(x1, y1)
    m.spmv(x2, y2)
    m.spmv(x1 + x2, y12)
    scale = 512.0 * 2.0
    assert float((y12 - y1 - y2).abs().max()) <= 1e-12 * scale
    rows = np.random.default_rng(0).choice(N, 2000, replace=False)
    hx = host(x1)
    yh = host(y1)
    for i in rows:
        s = rp[i]
        e = rp[i + 1]
        exact = float(np.dot(v[s:e], hx[col[s:e]]))
        assert abs(yh[i] - exact) <= 1e-12 * max(1.0, float(np.abs(v[s:e]).sum()))
    del m
    torch.cuda.empty_cache()


BIG_CASES = [("clustered", "f64", r) for r in RS] + [("clustered", "f32", r) for r in RS] + \
    [("stencil27", "f64", 2), ("stencil27", "f64", 4)]


@pytest.mark.parametrize("case,tag,r", BIG_CASES)
def test_full_size_converter_digests(case, tag, r):
    """Full-size GPU conversion, byte for byte: the clustered-runs matrix
    (4,194,304 rows, double beta(r,8) and float beta(r,16)) and the 27-point
    stencil 160^3 (double beta(2,8), beta(4,8)) give the block counts and
    the SHA-256 digests of rowsSize (block-row pointers), headers and values
    that the reference converter gives on the same input
    (tests/golden/big_digests.json, written by make_golden.py)."""
    from spc5_amd import synth
    from spc5_amd.device import DeviceMatrix
    ref = json.load(open(os.path.join(GOLDEN, "big_digests.json")))[case]
    rp, col, v = synth.clustered_runs(4_194_304, seed=1, value_seed=2) if case == "clustered" else synth.stencil27(160)
    assert len(rp) - 1 == ref["n"] and int(rp[-1]) == ref["nnz"]
    m = DeviceMatrix.from_csr(dev(rp), dev(col), dev(v.astype(DT[tag])), r)
    want = ref[f"{tag}_r{r}"]
    assert m.nblocks == want["nblocks"]
    assert sha(host(m.rows_size)) == want["rowsSize_sha256"]
    assert sha(host(m.headers)) == want["headers_sha256"]
    assert sha(host(m.values)) == want["values_sha256"]
    del m
    torch.cuda.empty_cache()


@pytest.mark.parametrize("r", RS)
def test_clustered_runs_full_size_float(r):
    """The clustered-runs matrix in float, the reference's float format
    beta(r,16), at full size: mask popcount = nnz, linearity within float
    rounding, and 2,000 sampled rows against numpy's double sums of the
    float values, normwise 1e-5 (sum |a_ij| |x_j| per row as the scale)."""
    from spc5_amd import synth
    from spc5_amd.device import DeviceMatrix
    N = 4_194_304
    rp, col, v64 = synth.clustered_runs(N)
    v = v64.astype(np.float32)
    m = DeviceMatrix.from_csr(dev(rp), dev(col), dev(v), r)
    assert m.vs == 16 and m.info()["nnz"] == len(v)
    x1 = dev(synth.uniform(N, 3).astype(np.float32))
    x2 = dev(synth.uniform(N, 4).astype(np.float32))
    y1 = torch.zeros(N, dtype=torch.float32, device="cuda")
    y2 = torch.zeros_like(y1)
    y12 = torch.zeros_like(y1)
    m.spmv(x1, y1)
    m.spmv(x2, y2)
    m.spmv(x1 + x2, y12)
    assert float((y12 - y1 - y2).abs().max()) <= 1e-5 * 512.0 * 2.0
    rows = np.random.default_rng(1).choice(N, 2000, replace=False)
    hx = host(x1).astype(np.float64)
    yh = host(y1)
    for i in rows:
        s, e = rp[i], rp[i + 1]
        exact = float(np.dot(v[s:e].astype(np.float64), hx[col[s:e]]))
        assert abs(float(yh[i]) - exact) <= 1e-5 * max(1.0, float(np.abs(v[s:e]).sum()))
    del m
    torch.cuda.empty_cache()


@pytest.mark.parametrize("tag", ["f64", "f32"])
@pytest.mark.parametrize("r", RS)
@pytest.mark.parametrize("row_sort", [False, True])
def test_gpu_stencil27(tag, r, row_sort):
    """27-point stencil: every block of a block-row covers the same r rows,
    the structure the row-sorted batches are for."""
    from spc5_amd import synth
    rp, col, v = synth.stencil27(20)
    v = v.astype(DT[tag])
    n = len(rp) - 1
    x, y0, bptr, hdr, vout, yref = _oracle_case(rp, col, v, r, 7, n)
    y = gpu_spmv_from_arrays(r, n, n, bptr, hdr, vout, x, y0, row_sort=row_sort)
    assert normwise_error(y, yref[:n], rp, v, x[:n], y0[:n]) <= TOL[tag]


@pytest.mark.parametrize("r", (1, 4))
@pytest.mark.parametrize("row_sort", [False, True])
def test_gpu_x_beyond_4gib(r, row_sort):
    """x spanning more than 4 GiB (ncols * 8 bytes): the kernel body with
    64-bit x addressing.  A few rows with columns spread up to the end of x;
    reference from the CSR and the x entries it touches."""
    from spc5_amd.device import DeviceMatrix
    ncols = (1 << 29) + 4099  # 4 GiB + 32 KiB of doubles
    rng = np.random.default_rng(11)
    n = 300
    sets = [np.unique(np.concatenate([rng.integers(0, ncols, 6), rng.integers(ncols - 40, ncols, 3),
                                      [i % 50]])) for i in range(n)]
    rowptr = np.zeros(n + 1, np.int32)
    for i, c in enumerate(sets):
        rowptr[i + 1] = rowptr[i] + len(c)
    col = np.concatenate(sets).astype(np.int32)
    v = rng.uniform(-1, 1, len(col))
    x = torch.zeros(ncols, dtype=torch.float64, device="cuda")
    used = np.unique(col)
    xu = rng.uniform(-1, 1, len(used))
    x[dev(used.astype(np.int64))] = dev(xu)
    xmap = dict(zip(used.tolist(), xu.tolist()))
    m = DeviceMatrix.from_csr(dev(rowptr), dev(col), dev(v), r, ncols=ncols)
    m.row_sort = row_sort
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    m.spmv(x, y)
    yref = np.array([sum(v[k] * xmap[int(col[k])] for k in range(rowptr[i], rowptr[i + 1])) for i in range(n)])
    scale = np.abs(v).max() * np.diff(rowptr).max()
    assert np.abs(host(y) - yref).max() <= 1e-12 * scale
    del x
    torch.cuda.empty_cache()


@pytest.mark.parametrize("r", RS)
@pytest.mark.parametrize("kind", ["f64", "f32vs16", "bf16view"])
def test_save_load_round_trip(tmp_path, r, kind):
    """A matrix written by spc5_hip_mat_save reads back (Python reader and
    the C-ABI loader) with identical arrays and an identical product."""
    import ctypes as C
    from spc5_amd._lib import check, lib
    from spc5_amd.device import DeviceMatrix
    from spc5_amd.solver import lowp
    rp, col, v = clustered_csr(1500, 2000, r, empty_frac=0.1)
    m = DeviceMatrix.from_csr(dev(rp), dev(col), dev(v if kind != "f32vs16" else v.astype(np.float32)), r,
                              ncols=2000)
    if kind == "bf16view":
        m = lowp(m, "bf16")
    p = str(tmp_path / "m.spc5")
    m.save(p)
    m2 = DeviceMatrix.load(p)
    assert (m2.r, m2.vs, m2.nrows, m2.ncols, m2.nnz, m2.nblocks) == (m.r, m.vs, m.nrows, m.ncols, m.nnz, m.nblocks)
    nbr = (m.nrows + r - 1) // r
    hb = m.nblocks * m.header_bytes()
    assert torch.equal(m.rows_size[: nbr + 1].cpu() - m.rows_size[0].cpu(), m2.rows_size[: nbr + 1].cpu())
    assert torch.equal(m.headers[:hb].cpu(), m2.headers[:hb].cpu())
    bits = torch.int16 if m.values.dtype == torch.bfloat16 else m.values.dtype
    assert torch.equal(m.values[: m.nnz].cpu().view(bits), m2.values[: m.nnz].cpu().view(bits))
    x = torch.from_numpy(np.random.default_rng(r).uniform(-1, 1, 2000)).cuda().to(m.vec_dtype)
    y1 = torch.zeros(1500, dtype=m.vec_dtype, device="cuda")
    y2 = torch.zeros_like(y1)
    m.spmv(x, y1)
    m2.spmv(x, y2)
    assert torch.equal(y1, y2)
    h = C.c_void_p(0)  # the C-ABI loader
    check(lib().spc5_hip_mat_load(p.encode(), C.byref(h)), "spc5_hip_mat_load")
    y3 = torch.zeros_like(y1)
    check(lib().spc5_hip_spmv(h, x.data_ptr(), y3.data_ptr(), None), "spc5_hip_spmv")
    torch.cuda.synchronize()
    assert torch.equal(y1, y3)
    check(lib().spc5_hip_mat_destroy(h), "destroy")
