"""Autotuner on the GPU: the model and the measurement agree on the
candidate set, and measure mode returns the fastest measured r."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_choose_r_modes():
    from spc5_amd import autotune, synth
    rp, col, val = synth.banded_gpu(200_000, 64, 0.3, seed=5)
    r_model, det = autotune.choose_r(rp, col, val, mode="model")
    assert r_model in autotune.RS and set(det["predicted_s"]) == set(autotune.RS)
    r_meas, mdet = autotune.choose_r(rp, col, val, mode="measure")
    t = mdet["measured_s"]
    assert r_meas == min(t, key=t.get)
    assert t[r_model] <= 1.5 * t[r_meas]  # the model's pick is close to the measured best


def test_tune_spmm_kt():
    from spc5_amd import autotune, synth
    from spc5_amd.device import DeviceMatrix
    rp, col, val = synth.banded_gpu(50_000, 16, 0.5, seed=2)
    A = DeviceMatrix.from_csr(rp, col, val, 2)
    kt, cand = autotune.tune_spmm_kt(A, 24)
    assert kt == min(cand, key=cand.get) and kt in (24, 16, 8)
    torch.cuda.synchronize()


def _product_error(m, rp, col, val):
    import numpy as np
    n = len(rp) - 1
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, m.ncols)
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    m.spmv(torch.from_numpy(x).cuda(), y)
    ref = np.zeros(n)
    np.add.at(ref, np.repeat(np.arange(n), np.diff(rp)), val * x[col])
    scale = max(float(np.abs(np.add.reduceat(np.abs(val), rp[:-1])).max()), 1.0)
    return float(np.abs(y.cpu().numpy() - ref).max()) / scale


def test_tile_values_option_and_tuning():
    """SPC5_HIP_OPT_TILE_VALUES rebuilds the tile plan (about nnz / target
    tiles, at most 128 rows each), keeps the product exact to rounding, and
    tune_tiles leaves the matrix on its fastest candidate."""
    from spc5_amd import autotune, synth
    from spc5_amd._lib import Spc5Error
    from spc5_amd.device import DeviceMatrix
    rp, col, val = synth.banded_gpu(120_000, 24, 0.5, seed=3)
    hrp, hcol, hval = (t.cpu().numpy() for t in (rp, col, val))
    m = DeviceMatrix.from_csr(rp, col, val, 2)
    nnz = len(hval)
    default_tiles = m.info()["ntiles"]
    for target in (256, 1024, 4096, 0):
        m.tile_values = target
        nt = m.info()["ntiles"]
        if target:
            assert nt >= nnz // target - 1 and nt >= 120_000 // 128
        else:
            assert nt == default_tiles
        assert _product_error(m, hrp, hcol, hval) <= 1e-12
    with pytest.raises(Spc5Error):
        m.tile_values = 10
    best, times = autotune.tune_tiles(m, candidates=(0, 512, 2048, 8192), reps=3)
    assert best == min(times, key=times.get) and set(times) == {0, 512, 2048, 8192}
    assert m.tile_values == best
    assert _product_error(m, hrp, hcol, hval) <= 1e-12


def test_batch_keeps_its_tiles_and_wave_tuning():
    """A batch built before a matrix's tile plan is rebuilt keeps using its
    own tiles (correct products after the rebuild); tune_batch_waves returns
    the fastest of its candidates."""
    import numpy as np
    from spc5_amd import autotune, synth
    from spc5_amd.batch import Batch
    from spc5_amd.device import DeviceMatrix
    mats, xs, ys, host = [], [], [], []
    for i, (n, bw) in enumerate(((3000, 8), (5000, 40), (800, 3))):
        rp, col, val = synth.banded_gpu(n, bw, 0.4, seed=10 + i)
        mats.append(DeviceMatrix.from_csr(rp, col, val, 4))
        x = torch.linspace(-1, 1, n, dtype=torch.float64, device="cuda")
        xs.append(x)
        ys.append(torch.zeros(n, dtype=torch.float64, device="cuda"))
        host.append(tuple(t.cpu().numpy() for t in (rp, col, val)) + (x.cpu().numpy(),))
    b = Batch(mats, xs, ys)
    for m in mats:
        m.tile_values = 64  # new plans; the batch keeps the old ones
    torch.cuda.synchronize()
    b.spmv(graph=True)
    torch.cuda.synchronize()
    for (rp, col, val, x), y in zip(host, ys):
        ref = np.zeros(len(rp) - 1)
        np.add.at(ref, np.repeat(np.arange(len(rp) - 1), np.diff(rp)), val * x[col])
        assert np.abs(y.cpu().numpy() - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()) * 50
    b.close()
    waves, times = autotune.tune_batch_waves(lambda w: Batch(mats, xs, ys, waves=w), fractions=(0.5, 1.0), reps=2)
    assert waves == min(times, key=times.get)


def test_tune_batch_tiles():
    """tune_batch_tiles times a batch built at every tile size, leaves every
    matrix set to the fastest, and a batch built then multiplies correctly."""
    import numpy as np
    from spc5_amd import autotune, synth
    from spc5_amd.batch import Batch
    from spc5_amd.device import DeviceMatrix
    mats, xs, ys, host = [], [], [], []
    for i, (n, bw) in enumerate(((30_000, 40), (12_000, 8), (50_000, 100))):
        rp, col, val = synth.banded_gpu(n, bw, 0.3, seed=20 + i)
        mats.append(DeviceMatrix.from_csr(rp, col, val, 4))
        x = torch.linspace(-1, 1, n, dtype=torch.float64, device="cuda")
        xs.append(x)
        ys.append(torch.zeros(n, dtype=torch.float64, device="cuda"))
        host.append(tuple(t.cpu().numpy() for t in (rp, col, val)) + (x.cpu().numpy(),))
    best, times = autotune.tune_batch_tiles(mats, lambda: Batch(mats, xs, ys), candidates=(0, 1024, 4096), reps=2)
    assert set(times) == {0, 1024, 4096} and best == min(times, key=times.get)
    assert all(m.tile_values == best for m in mats)
    for y in ys:
        y.zero_()
    b = Batch(mats, xs, ys)
    assert b.info()["ntiles"] == sum(m.info()["ntiles"] for m in mats)
    b.spmv(graph=False)
    torch.cuda.synchronize()
    for (rp, col, val, x), y in zip(host, ys):
        ref = np.zeros(len(rp) - 1)
        np.add.at(ref, np.repeat(np.arange(len(rp) - 1), np.diff(rp)), val * x[col])
        assert np.abs(y.cpu().numpy() - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()) * 50
    b.close()
