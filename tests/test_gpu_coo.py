"""COO -> CSR on the GPU (the reference's COO_unsorted_to_CSR /
COO_sorted_to_CSR): arrays bit-identical to a stable numpy sort by
(row, column) with duplicates kept in input order; the sorted path checks
the order; out-of-range indices are errors; a MatrixMarket file goes
through the reader, the GPU conversion and the beta(r,VS) product."""
import numpy as np
import pytest
import torch

import oracle as O
from conftest import normwise_error

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def ref_csr(i, j, v, nrows):
    order = np.lexsort((j, i))  # stable: equal (i, j) keep input order
    rowptr = np.zeros(nrows + 1, np.int64)
    np.add.at(rowptr, i + 1, 1)
    return np.cumsum(rowptr).astype(np.int32), j[order], v[order]


@pytest.mark.parametrize("seed,nrows,ncols,nnz", [(0, 1, 1, 1), (1, 50, 70, 400), (2, 3000, 100, 50000),
                                                  (3, 100000, 200000, 700000), (4, 7, 7, 0)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_unsorted_matches_stable_sort(seed, nrows, ncols, nnz, dtype):
    from spc5_amd.mm import coo_to_csr
    rng = np.random.default_rng(seed)
    i = rng.integers(0, nrows, nnz).astype(np.int32)
    j = rng.integers(0, ncols, nnz).astype(np.int32)
    if nnz > 10:  # duplicates
        i[-5:], j[-5:] = i[:5], j[:5]
    v = rng.uniform(-1, 1, nnz).astype(dtype)
    rp, col, val = coo_to_csr(dev(i), dev(j), dev(v), nrows, ncols)
    wrp, wcol, wval = ref_csr(i.astype(np.int64), j, v, nrows)
    assert np.array_equal(rp.cpu().numpy(), wrp)
    assert np.array_equal(col.cpu().numpy(), wcol)
    assert np.array_equal(val.cpu().numpy(), wval)


def test_sorted_path_and_errors():
    from spc5_amd._lib import Spc5Error
    from spc5_amd.mm import coo_to_csr
    i = np.array([0, 0, 1, 3], np.int32)
    j = np.array([1, 4, 0, 2], np.int32)
    v = np.array([1.0, 2.0, 3.0, 4.0])
    rp, col, val = coo_to_csr(dev(i), dev(j), dev(v), 4, 5, sorted=True)
    assert rp.cpu().tolist() == [0, 2, 3, 3, 4] and col.cpu().tolist() == j.tolist()
    with pytest.raises(Spc5Error):
        coo_to_csr(dev(i[::-1].copy()), dev(j[::-1].copy()), dev(v), 4, 5, sorted=True)
    with pytest.raises(Spc5Error):
        coo_to_csr(dev(i), dev(j), dev(v), 3, 5)  # row 3 out of range
    with pytest.raises(Spc5Error):
        coo_to_csr(dev(i), dev(j), dev(v), 4, 4)  # column 4 out of range


@pytest.mark.parametrize("r", (1, 4))
def test_matrixmarket_file_to_product(tmp_path, r):
    from spc5_amd.device import DeviceMatrix
    from spc5_amd.mm import load_csr
    rng = np.random.default_rng(r)
    n = 400
    ent = {}
    for _ in range(3000):
        a, b = sorted(rng.integers(1, n + 1, 2))
        ent[(b, a)] = rng.uniform(-1, 1)
    lines = ["%%MatrixMarket matrix coordinate real symmetric", f"{n} {n} {len(ent)}"]
    lines += [f"{a} {b} {x!r}" for (a, b), x in ent.items()]
    p = tmp_path / "s.mtx"
    p.write_text("\n".join(lines) + "\n")
    rowptr, col, val, ncols = load_csr(str(p))
    # dense reference from the file's own entries
    D = np.zeros((n, n))
    for (a, b), x in ent.items():
        D[a - 1, b - 1] = x
        D[b - 1, a - 1] = x
    rp, cl, vv = rowptr.cpu().numpy(), col.cpu().numpy(), val.cpu().numpy()
    dense = np.zeros((n, n))
    for i in range(n):
        dense[i, cl[rp[i]:rp[i + 1]]] = vv[rp[i]:rp[i + 1]]
    assert np.array_equal(dense, D)
    m = DeviceMatrix.from_csr(rowptr, col, val, r, ncols=ncols)
    x = rng.uniform(-1, 1, n)
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    m.spmv(dev(x), y)
    assert normwise_error(y.cpu().numpy(), D @ x, rp, vv, x, np.zeros(1)) <= 1e-12
