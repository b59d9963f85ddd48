"""Autotuner model (CPU only): the non-negative least-squares fit recovers
known per-r costs, the committed MI355X calibration has non-negative costs,
and left out of the fit one matrix at a time it picks the measured fastest r
for at least 20 of its 24 matrices."""
import json
import os

import numpy as np

from spc5_amd import autotune as A


def _coeffs(seed):
    rng = np.random.default_rng(seed)
    return {str(r): {"value": float(rng.uniform(0.5, 3.0) * 1e-12), "block": float(rng.uniform(0.0, 4.0) * 1e-12),
                     "batch": float(rng.uniform(0.5, 8.0) * 1e-10)} for r in A.RS}


def _synthetic_measurements(coeffs, rng):
    out = []
    for _ in range(16):
        n = int(rng.integers(100_000, 4_000_000))
        nnz = int(n * rng.uniform(3, 80))
        for r in A.RS:
            # values per block on both sides of the batch cap's 6 (384 / 64),
            # so that batches are not proportional to blocks
            nb = int(nnz / rng.uniform(1.5, 12.0))
            m = {"r": r, "nrows": n, "nnz": nnz, "nblocks": nb}
            m["seconds"] = float(np.dot(A._model_vector(coeffs), A.features(n, n, nnz, nb, r, 8, 8, 8)))
            out.append(m)
    return out


def test_fit_recovers_known_coefficients():
    want = _coeffs(3)
    got = A.fit_coefficients(_synthetic_measurements(want, np.random.default_rng(3)))
    assert np.allclose(A._model_vector(got), A._model_vector(want), rtol=1e-4, atol=1e-20)


def test_fit_costs_are_never_negative():
    # timings a signed fit would explain with negative costs (the per-block
    # cost falls as blocks grow): the fit clips such costs at zero
    rng = np.random.default_rng(5)
    ms = _synthetic_measurements(_coeffs(5), rng)
    for m in ms:
        m["seconds"] *= 1.0 - 0.3 * m["nblocks"] / (m["nnz"] + 1.0)
    got = A._model_vector(A.fit_coefficients(ms))
    assert (got >= 0).all()


def test_features_layout():
    f = A.features(1000, 1000, 5000, 900, 4, 8, 8, 8)
    assert f.shape == (12,)
    assert list(f[6:9]) == [5000.0, 900.0, max(900 / 64.0, 5000 / 384.0)]
    assert not f[:6].any() and not f[9:].any()
    assert A.features(1000, 1000, 5000, 900, 4, 4, 4, 16)[8] == max(900 / 64.0, 5000 / 512.0)


def test_committed_calibration():
    with open(A.DEFAULT_FILE) as f:
        d = json.load(f)
    c = A.load_coefficients()
    assert c == d["coefficients"]
    assert (A._model_vector(c) >= 0).all()
    raw = d["measurements"]
    loo = A.leave_one_out(raw)
    assert loo["matrices"] == 24 and loo["agree"] >= 20, loo
    assert loo["mean_loss"] < 0.02
    assert loo == d["leave_one_out"]
    # on its own timings (in sample) as well
    groups = {}
    for m in raw:
        groups.setdefault((m["nrows"], m["nnz"]), []).append(m)
    agree = 0
    for g in groups.values():
        pred = A.predict({s["r"]: s["nblocks"] for s in g}, g[0]["nrows"], g[0]["nrows"], g[0]["nnz"])
        agree += min(pred, key=pred.get) == min(g, key=lambda s: s["seconds"])["r"]
    assert agree >= 20


def test_default_model_without_file(tmp_path):
    assert A.load_coefficients(os.path.join(tmp_path, "none.json")) == A._default_coefficients()
    assert (A._model_vector(A._default_coefficients()) >= 0).all()
