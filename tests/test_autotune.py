"""Autotuner model (CPU only): the least-squares fit recovers known
coefficients, and the committed MI355X calibration picks the measured
fastest r on its own timings as often as recorded."""
import json
import os

import numpy as np

from spc5_amd import autotune as A


def _synthetic_measurements(coeffs, rng):
    out = []
    for i in range(12):
        n = int(rng.integers(100_000, 4_000_000))
        nnz = int(n * rng.uniform(3, 80))
        for r in A.RS:
            nb = int(nnz / rng.uniform(1.5, 4.0 * r))
            m = {"r": r, "nrows": n, "nnz": nnz, "nblocks": nb}
            m["seconds"] = float(np.dot(A._model_vector(coeffs), A.features(n, n, nnz, nb, r, 8, 8, 8)))
            out.append(m)
    return out


def test_fit_recovers_known_coefficients():
    want = {"bytes": 2.0e-13, "nblocks": 3.0e-12, "batches": 5.0e-10,
            "nnz": {"1": 1.0e-12, "2": 1.5e-12, "4": 2.0e-12, "8": 2.5e-12}}
    got = A.fit_coefficients(_synthetic_measurements(want, np.random.default_rng(3)))
    assert np.allclose(A._model_vector(got), A._model_vector(want), rtol=1e-6, atol=0)


def test_features_layout():
    f = A.features(1000, 1000, 5000, 900, 4, 8, 8, 8)
    assert f[0] == A.spmv_bytes(1000, 1000, 5000, 900, 4, 8, 8, 8)
    assert f[1] == 900 and f[2] == max(900 / 64.0, 5000 / 384.0)
    assert list(f[3:]) == [0.0, 0.0, 5000.0, 0.0]
    assert A.features(1000, 1000, 5000, 900, 4, 4, 4, 16)[2] == max(900 / 64.0, 5000 / 512.0)


def test_committed_calibration_choices():
    with open(A.DEFAULT_FILE) as f:
        d = json.load(f)
    c = A.load_coefficients()
    assert c == d["coefficients"]
    raw = d["measurements"]
    mats = [raw[i:i + 4] for i in range(0, len(raw), 4)]
    agree, regret = 0, []
    for m in mats:
        pred = [A.predict({s["r"]: s["nblocks"]}, s["nrows"], s["nrows"], s["nnz"])[s["r"]] for s in m]
        t = [s["seconds"] for s in m]
        agree += int(np.argmin(pred)) == int(np.argmin(t))
        regret.append(t[int(np.argmin(pred))] / min(t) - 1.0)
    assert agree >= 18 and np.mean(regret) < 0.03


def test_default_model_without_file(tmp_path):
    assert A.load_coefficients(os.path.join(tmp_path, "none.json")) == A._DEFAULT
