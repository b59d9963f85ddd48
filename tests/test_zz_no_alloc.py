"""The converter's count pass without allocation
(spc5_hip_convert_count_noalloc): same block-row pointers and block count
as spc5_hip_convert_count, called through the C ABI, twice in a row with a
smaller matrix the second time (the fixed scratch keeps the first call's
contents), for every (r, VS); its error paths leave the scratch usable."""
import ctypes as C

import numpy as np
import pytest
import torch

from matgen import clustered_csr

pytestmark = pytest.mark.gpu

_P = C.c_void_p
_I = C.c_int
_I64 = C.c_int64


def _lib():
    from spc5_amd._lib import lib
    L = lib()
    for name in ("spc5_hip_convert_count", "spc5_hip_convert_count_noalloc"):
        f = getattr(L, name)
        f.restype = _I
        f.argtypes = [_I, _I, _I64, _P, _P, _P, C.POINTER(_I64), _P]
    return L


def _count(fn, vs, r, rp, col):
    n = rp.numel() - 1
    nbr = (n + r - 1) // r
    out = torch.full((nbr + 1,), -7, dtype=torch.int32, device="cuda")
    nb = _I64(-1)
    stream = _P(torch.cuda.current_stream().cuda_stream)
    rc = fn(vs, r, n, _P(rp.data_ptr()), _P(col.data_ptr()) if col.numel() else None, _P(out.data_ptr()),
            C.byref(nb), stream)
    torch.cuda.synchronize()
    return rc, out.cpu().numpy(), nb.value


@pytest.mark.parametrize("vs", (8, 16))
@pytest.mark.parametrize("r", (1, 2, 4, 8))
def test_noalloc_count_equals_count(vs, r):
    L = _lib()
    for rows, cols, seed in ((2999, 1500, 3), (411, 700, 4)):  # second call smaller
        rp, col, _ = clustered_csr(rows, cols, seed=seed + r, empty_frac=0.1)
        d_rp = torch.from_numpy(rp).cuda()
        d_col = torch.from_numpy(col).cuda()
        rc0, want, nb0 = _count(L.spc5_hip_convert_count, vs, r, d_rp, d_col)
        rc1, got, nb1 = _count(L.spc5_hip_convert_count_noalloc, vs, r, d_rp, d_col)
        assert rc0 == 0 and rc1 == 0, L.spc5_hip_last_error()
        assert nb1 == nb0 > 0
        assert np.array_equal(got, want)


def test_noalloc_count_errors_and_limit():
    L = _lib()
    rp, col, _ = clustered_csr(300, 300, seed=9)
    good = (torch.from_numpy(rp).cuda(), torch.from_numpy(col).cuda())
    # unsorted columns in one row: rejected, and the next call starts clean
    bad_col = col.copy()
    i0, i1 = rp[5], rp[6]
    assert i1 - i0 >= 2
    bad_col[i0], bad_col[i0 + 1] = bad_col[i0 + 1], bad_col[i0]
    rc, _, _ = _count(L.spc5_hip_convert_count_noalloc, 8, 2, good[0], torch.from_numpy(bad_col).cuda())
    assert rc != 0 and b"increasing" in L.spc5_hip_last_error()
    rc, got, nb = _count(L.spc5_hip_convert_count_noalloc, 8, 2, *good)
    rc0, want, nb0 = _count(L.spc5_hip_convert_count, 8, 2, *good)
    assert rc == 0 and rc0 == 0 and nb == nb0 and np.array_equal(got, want)
    # more block-rows than the fixed scratch holds: refused before any launch
    out = torch.zeros(16, dtype=torch.int32, device="cuda")
    nbv = _I64(-1)
    stream = _P(torch.cuda.current_stream().cuda_stream)
    rc = L.spc5_hip_convert_count_noalloc(8, 1, 40_000_000, _P(good[0].data_ptr()), _P(good[1].data_ptr()),
                                          _P(out.data_ptr()), C.byref(nbv), stream)
    assert rc != 0 and b"scratch" in L.spc5_hip_last_error()
    assert nbv.value == 0
