"""Batching form: uneven matrices with different r in one persistent
hipGraph launch.  Bar: every y_i within 1e-12 (double) / 1e-5 (float)
normwise of the oracle's beta(r,VS) product; several graph replays with
eager work and synchronisation between them, each checked."""
import numpy as np
import pytest
import torch

import oracle as O
from conftest import normwise_error
from matgen import clustered_csr, random_csr

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("nmats,nrows,waves", [(1, 50, 0), (3, 200, 1), (5, 3000, 3), (12, 2500, 8), (12, 2500, 17),
                                                (12, 2500, 1000), (30, 4000, 0)])
def test_batch_persistent_wave_counts(nmats, nrows, waves):
    """The persistent waves walk the tile list with a stride of the wave
    count: a single tile, one wave, fewer waves than tiles, more waves than
    tiles and the default count all cover every tile exactly once (each y
    against the oracle after one product)."""
    from spc5_amd.batch import Batch
    from spc5_amd.device import DeviceMatrix
    rng = np.random.default_rng(nmats * 7 + waves)
    mats, xs, ys, refs = [], [], [], []
    for i in range(nmats):
        n = int(rng.integers(1, nrows + 1))
        rp, col, v = clustered_csr(n, n, seed=100 + i, empty_frac=0.1)
        m = DeviceMatrix.from_csr(dev(rp), dev(col), dev(v), 2, ncols=n)
        m.tile_values = 256  # many small tiles
        x = rng.uniform(-1, 1, n)
        y0 = rng.uniform(-1, 1, n)
        mats.append(m)
        xs.append(dev(x))
        ys.append(dev(y0))
        bptr, hdr, vout, nb = O.convert(2, rp, col, v)
        xp = np.zeros(n + 8)
        xp[:n] = x
        yr = np.zeros(n + 8)
        if nb:
            O.beta_spmv(2, n, bptr, hdr, vout, xp, yr)
        refs.append((rp, v, x, y0, yr[:n]))
    b = Batch(mats, xs, ys, waves=waves)
    b.spmv(graph=False)
    torch.cuda.synchronize()
    for (rp, v, x, y0, yr), y in zip(refs, ys):
        assert normwise_error(y.cpu().numpy(), y0 + yr, rp, v, x, y0) <= 1e-12, (b.info(), len(rp) - 1)
    b.close()


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
@pytest.mark.parametrize("same_r", [0, 1, 2, 4, 8])
def test_batch_uneven_matrices(dtype, tol, same_r):
    """same_r 0: r cycles through 1, 2, 4, 8 (the kernel with every r);
    otherwise every matrix has that r (the kernel with only its code)."""
    from spc5_amd.batch import Batch
    from spc5_amd.device import DeviceMatrix
    rng = np.random.default_rng(0)
    mats, xs, ys, refs = [], [], [], []
    for i in range(40):
        n = int(rng.integers(1, 900))
        nc = int(rng.integers(1, 700))
        if i % 3 == 0:
            rp, col, v = random_csr(n, nc, 0.03, seed=i, dtype=dtype)
        else:
            rp, col, v = clustered_csr(n, nc, seed=i, dtype=dtype, empty_frac=0.1)
        r = same_r if same_r else (1, 2, 4, 8)[i % 4]
        m = DeviceMatrix.from_csr(dev(rp), dev(col), dev(v), r, ncols=nc)
        x = rng.uniform(-1, 1, nc).astype(dtype)
        y0 = rng.uniform(-1, 1, n).astype(dtype)
        mats.append(m)
        xs.append(dev(x))
        ys.append(dev(y0))
        vs = O.VS_OF[np.dtype(dtype)]
        bptr, hdr, vout, nb = O.convert(r, rp, col, v)
        xp = np.zeros(nc + vs, dtype)
        xp[:nc] = x
        yr = np.zeros(n + 8, dtype)
        if nb:
            O.beta_spmv(r, n, bptr, hdr, vout, xp, yr)
        refs.append((rp, v, x, y0, yr[:n].astype(np.float64)))
    b = Batch(mats, xs, ys)
    assert b.info()["ntiles"] >= 40
    for rep in range(3):
        b.spmv(graph=(rep != 1))
        torch.cuda.synchronize()
        for (rp, v, x, y0, yr), y in zip(refs, ys):
            want = y0.astype(np.float64) + (rep + 1) * yr
            assert normwise_error(y.cpu().numpy(), want, rp, v * (rep + 1), x, y0) <= tol
        # eager work between replays
        torch.ones(1000, device="cuda").sum().item()
    b.close()


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
@pytest.mark.parametrize("r", [2, 4])
def test_batch_row_sorted(dtype, tol, r):
    """Every matrix with the same r in {2, 4} and row sorting on: the batch
    runs the row-sorted tile, so each y is bit-identical to the matrix's own
    row-sorted product (same tiles, same routine) and matches the oracle."""
    from spc5_amd.batch import Batch
    from spc5_amd.device import DeviceMatrix
    rng = np.random.default_rng(11)
    mats, xs, ys, singles, refs = [], [], [], [], []
    for i in range(30):
        n = int(rng.integers(1, 1500))
        nc = int(rng.integers(1, 1200))
        if i % 3 == 0:
            rp, col, v = random_csr(n, nc, 0.03, seed=100 + i, dtype=dtype)
        else:
            rp, col, v = clustered_csr(n, nc, seed=100 + i, dtype=dtype, empty_frac=0.1)
        m = DeviceMatrix.from_csr(dev(rp), dev(col), dev(v), r, ncols=nc)
        m.row_sort = True
        x = rng.uniform(-1, 1, nc).astype(dtype)
        y0 = rng.uniform(-1, 1, n).astype(dtype)
        ys1 = dev(y0)
        m.spmv(dev(x), ys1)
        singles.append(ys1)
        mats.append(m)
        xs.append(dev(x))
        ys.append(dev(y0))
        vs = O.VS_OF[np.dtype(dtype)]
        bptr, hdr, vout, nb = O.convert(r, rp, col, v)
        xp = np.zeros(nc + vs, dtype)
        xp[:nc] = x
        yr = np.zeros(n + 8, dtype)
        if nb:
            O.beta_spmv(r, n, bptr, hdr, vout, xp, yr)
        refs.append((rp, v, x, y0, yr[:n].astype(np.float64)))
    b = Batch(mats, xs, ys)
    b.spmv(graph=True)
    torch.cuda.synchronize()
    for (rp, v, x, y0, yr), y, y1 in zip(refs, ys, singles):
        got = y.cpu().numpy()
        assert np.array_equal(got, y1.cpu().numpy())
        assert normwise_error(got, y0.astype(np.float64) + yr, rp, v, x, y0) <= tol
    b.close()


@pytest.mark.parametrize("n,bw,fill,seed", [(1000, 4, 0.3, 7), (5003, 256, 0.3, 8), (777, 31, 0.9, 9)])
def test_banded_generator_gpu_equals_host(n, bw, fill, seed):
    """The device generator of the uneven batch is byte-identical to the
    host generator (same hashing and value formula)."""
    from spc5_amd import synth
    rp, col, val = synth.banded(n, bw, fill, seed)
    drp, dcol, dval = synth.banded_gpu(n, bw, fill, seed)
    assert np.array_equal(drp.cpu().numpy(), rp)
    assert np.array_equal(dcol.cpu().numpy(), col)
    assert np.array_equal(dval.cpu().numpy().view(np.uint64), val.view(np.uint64))


def _oracle_y(r, rp, col, v, x, y0, vs, msz):
    """y0 + A x by the oracle's beta(r,VS) product (reference arithmetic)."""
    bptr, hdr, vout, nb = O.convert(r, rp, col, v, vs=vs, msz=msz)
    n, nc = len(rp) - 1, len(x)
    xp = np.zeros(nc + vs, v.dtype)
    xp[:nc] = x
    yr = np.zeros(n + 8, v.dtype)
    if nb:
        O.beta_spmv(r, n, bptr, hdr, vout, xp, yr, vs=vs, msz=msz)
    return y0.astype(np.float64) + yr[:n].astype(np.float64)


@pytest.mark.parametrize("kind", ("f64", "lowp_f32"))
@pytest.mark.parametrize("r_rule", ("autotuned", "mixed"))
def test_batch_banded_recipe(kind, r_rule):
    """The uneven batch's own recipe at reduced size: 60 banded matrices
    with the batch's (rows, bandwidth, seed) table, rows scaled down 40x
    (25 .. 5,000), bandwidth capped below the row count, 30 % fill, each with
    its own x.  r per matrix from the autotuner's model ("autotuned") or
    forced to cycle 1, 2, 4, 8 ("mixed": the kernel with every r's body);
    double values, or float copies of the double matrices (beta(r,8),
    float sums) as the benchmark times them.  One hipGraph launch, replayed
    three times with eager work between, every y against the oracle
    (normwise 1e-12 double, 1e-5 float)."""
    from spc5_amd import autotune, synth
    from spc5_amd.batch import Batch
    from spc5_amd.device import DeviceMatrix
    from spc5_amd.solver import lowp
    rng = np.random.default_rng(21)
    mats, xs, ys, cases = [], [], [], []
    rs_used = set()
    for i, (n, bw, seed) in enumerate(synth.uneven_batch_params(60)):
        n = max(25, n // 40)
        bw = min(bw, n - 1)
        drp, dcol, dval = synth.banded_gpu(n, bw, 0.3, seed)
        if r_rule == "autotuned":
            r, _ = autotune.choose_r(drp, dcol, dval, mode="model")
        else:
            r = (1, 2, 4, 8)[i % 4]
        rs_used.add(r)
        m = DeviceMatrix.from_csr(drp, dcol, dval, r)
        rp, col, v = drp.cpu().numpy(), dcol.cpu().numpy(), dval.cpu().numpy()
        if kind == "lowp_f32":
            m = lowp(m, "f32")
            v = v.astype(np.float32)
        x = rng.uniform(-1, 1, n).astype(v.dtype)
        y0 = rng.uniform(-1, 1, n).astype(v.dtype)
        mats.append(m)
        xs.append(dev(x))
        ys.append(dev(y0))
        cases.append((r, rp, col, v, x, y0))
    if r_rule == "mixed":
        assert rs_used == {1, 2, 4, 8}
    b = Batch(mats, xs, ys)
    tol = 1e-12 if kind == "f64" else 1e-5
    for rep in range(3):
        b.spmv(graph=True)
        torch.cuda.synchronize()
        for (r, rp, col, v, x, y0), y in zip(cases, ys):
            # y0 + (rep+1) A x: the product of the matrix scaled by rep+1
            want = _oracle_y(r, rp, col, v * (rep + 1), x, y0, 8, 1)
            assert normwise_error(y.cpu().numpy(), want, rp, v * (rep + 1), x, y0) <= tol, (rep, r, len(rp) - 1)
        torch.ones(1000, device="cuda").sum().item()
    b.close()


@pytest.mark.parametrize("row_sort", (False, True))
def test_batch_full_size_equals_each_product(row_sort):
    """The uneven batch at full size (2,000 banded matrices, rows
    log-uniform in [1e3, 2e5], bandwidth 4..256, fill 0.3: 74 M rows, 6.0 G
    values, r per matrix from the autotuner's model), one graph launch: every
    sampled matrix's y is bit-identical to that matrix's own product (same
    tiles, same arithmetic), which is pinned against the oracle separately."""
    from spc5_amd import autotune, synth
    from spc5_amd.batch import Batch
    from spc5_amd.device import DeviceMatrix
    params = synth.uneven_batch_params(2000)
    mats, rows, nnz = [], 0, 0
    for (n, bw, seed) in params:
        rp, col, val = synth.banded_gpu(n, bw, 0.3, seed)
        r, _ = autotune.choose_r(rp, col, val, mode="model")
        mats.append(DeviceMatrix.from_csr(rp, col, val, r))
        rows += n
        nnz += val.numel()
        del rp, col, val
    assert len(mats) == 2000 and rows > 70_000_000 and nnz > 5_000_000_000
    for m in mats:
        m.row_sort = row_sort
    g = torch.Generator(device="cuda").manual_seed(9)
    xs = [torch.empty(m.ncols, dtype=torch.float64, device="cuda").uniform_(-1, 1, generator=g) for m in mats]
    ys = [torch.empty(m.nrows, dtype=torch.float64, device="cuda").uniform_(-1, 1, generator=g) for m in mats]
    y0 = {i: ys[i].clone() for i in range(0, 2000, 83)}
    b = Batch(mats, xs, ys)
    b.spmv(graph=True)
    torch.cuda.synchronize()
    for i, yi in y0.items():
        mats[i].spmv(xs[i], yi)
        assert torch.equal(yi, ys[i]), (i, mats[i].r, mats[i].nrows)
    b.close()
    del b, xs, ys, mats
    torch.cuda.empty_cache()
