"""The C++ drop-in header (include/spc5_hip.hpp) in a reference-style
driver (tests/drivers/spc5_compare.cpp): COO -> CSR, CSR product x16, every
beta(r,VS) format converted on the GPU and applied x16 through the
reference API (SPC5_Spmv and the prepared SPC5_split_omp/SPC5_Spmv_omp
pair).  Bar: normwise error vs the CSR product <= 1e-12 (double) / 1e-5
(float), block counts equal SPC5_block_count, SPC5_iterate visits nnz
entries."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "drivers", "spc5_compare")


@pytest.mark.parametrize("real,tol", [("double", 1e-12), ("float", 1e-5)])
@pytest.mark.parametrize("kind,arg", [("dense", "64"), ("lap", "40"), ("dense", "33")])
def test_cpp_dropin_driver(kind, arg, real, tol):
    assert os.path.exists(DRIVER), "tests/drivers/spc5_compare not built (run __graft_entry__.build())"
    p = subprocess.run([DRIVER, kind, arg, real], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "RESULT DONE" in p.stdout
    errs = [float(x) for x in re.findall(r"-> normwise ([0-9.eE+-]+)", p.stdout)]
    assert len(errs) == 8
    assert max(errs) <= tol, p.stdout
