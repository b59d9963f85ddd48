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
