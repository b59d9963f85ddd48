"""Multi-right-hand-side product Y = alpha A X + beta Y (batching form).
No reference counterpart; pinned by numpy float64 on the decoded matrix.
Bar: normwise 1e-12 (double), 1e-5 (float), with k in {1,3,4,16,64,70} and
several column chunks kt (the matrix re-read per chunk)."""
import numpy as np
import pytest
import torch

import oracle as O
from matgen import clustered_csr

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def dense_of(rp, col, val, ncols):
    n = len(rp) - 1
    A = np.zeros((n, ncols))
    for i in range(n):
        A[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    return A


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
@pytest.mark.parametrize("r", (1, 2, 4, 8))
@pytest.mark.parametrize("k,kt", [(1, 0), (3, 0), (4, 0), (16, 0), (16, 4), (64, 0), (64, 16), (70, 0), (70, 32)])
def test_spmm(dtype, tol, r, k, kt):
    from spc5_amd.device import DeviceMatrix
    rp, col, val = clustered_csr(601, 450, seed=r + k, dtype=dtype, empty_frac=0.1)
    A = DeviceMatrix.from_csr(dev(rp), dev(col), dev(val), r, ncols=450)
    rng = np.random.default_rng(k)
    X = rng.uniform(-1, 1, (450, k)).astype(dtype)
    Y0 = rng.uniform(-1, 1, (601, k)).astype(dtype)
    Y = dev(Y0.copy())
    A.spmm(dev(X), Y, alpha=1.0, beta=1.0, kt=kt)
    Ad = dense_of(rp, col, val.astype(np.float64), 450)
    ref = Ad @ X.astype(np.float64) + Y0
    scale = max(np.abs(Ad).sum(axis=1).max() * np.abs(X).max(), np.abs(Y0).max())
    err = np.abs(Y.cpu().numpy().astype(np.float64) - ref).max() / scale
    assert err <= tol, err


def test_spmm_strided_and_beta0():
    from spc5_amd.device import DeviceMatrix
    rp, col, val = clustered_csr(300, 300, seed=3)
    A = DeviceMatrix.from_csr(dev(rp), dev(col), dev(val), 4)
    rng = np.random.default_rng(2)
    Xbig = dev(rng.uniform(-1, 1, (300, 24)))
    X = Xbig[:, 4:12]                      # ldx = 24, k = 8
    Y = torch.full((300, 8), float("nan"), dtype=torch.float64, device="cuda")
    A.spmm(X, Y, alpha=2.0, beta=0.0)      # beta = 0: Y is not read
    ref = 2.0 * dense_of(rp, col, val, 300) @ X.cpu().numpy()
    assert np.abs(Y.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max() * 10


def test_spmm_wide_x_rows():
    """X with a leading dimension that puts it beyond 4 GiB: the SpMM body
    with 64-bit X addressing; reference from the CSR."""
    import numpy as np
    import torch
    from spc5_amd.device import DeviceMatrix
    rng = np.random.default_rng(5)
    n, ncols, k, ldx = 200, 70_000, 8, 8200  # ncols * ldx * 8 > 4 GiB
    sets = [np.unique(np.concatenate([rng.integers(0, ncols, 5), [ncols - 1 - i % 7]])) for i in range(n)]
    rowptr = np.zeros(n + 1, np.int32)
    for i, c in enumerate(sets):
        rowptr[i + 1] = rowptr[i] + len(c)
    col = np.concatenate(sets).astype(np.int32)
    v = rng.uniform(-1, 1, len(col))
    X = torch.zeros((ncols, ldx), dtype=torch.float64, device="cuda")
    used = np.unique(col)
    xu = rng.uniform(-1, 1, (len(used), k))
    X[torch.from_numpy(used.astype(np.int64)).cuda(), :k] = torch.from_numpy(xu).cuda()
    xmap = {int(c): xu[j] for j, c in enumerate(used)}
    A = DeviceMatrix.from_csr(torch.from_numpy(rowptr).cuda(), torch.from_numpy(col).cuda(),
                              torch.from_numpy(v).cuda(), 2, ncols=ncols)
    Y = torch.zeros((n, k), dtype=torch.float64, device="cuda")
    A.spmm(X[:, :k], Y, beta=0.0)  # a row-major view with leading dimension ldx
    torch.cuda.synchronize()
    Yref = np.array([sum(v[j] * xmap[int(col[j])] for j in range(rowptr[i], rowptr[i + 1])) for i in range(n)])
    scale = np.abs(v).max() * np.diff(rowptr).max()
    assert np.abs(Y.cpu().numpy() - Yref).max() <= 1e-12 * scale
    del X
    torch.cuda.empty_cache()
