"""bench.py's host-side paths (CPU only): the CPU baseline in its own
process, on the CPUs it is given, with the reference's thread binding; and
the refusal of more than one GPU."""
import json
import os
import subprocess
import sys

import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref (the compiled reference) is not built")
def test_cpu_baseline_child_process():
    cpus = sorted(os.sched_getaffinity(0))[:2]
    env = dict(os.environ, OMP_PROC_BIND="true", OMP_PLACES="cores", OMP_NUM_THREADS=str(len(cpus)))
    res = subprocess.run([sys.executable, BENCH, "--cpu-baseline-child", "--workload", "clustered", "--n", "20000",
                          "--cpu-r", "2", "--cpu-list", ",".join(str(c) for c in cpus)],
                         env=env, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr[-2000:]
    out = json.loads(res.stdout.strip().splitlines()[-1])
    assert out["kind"] == "reference" and out["unit"] == "GFLOP/s"
    assert out["cores"] == len(cpus)
    assert out["omp_proc_bind"] == "true" and out["omp_places"] == "cores"
    assert len(out["samples_gflops"]) == 5 and all(v > 0 for v in out["samples_gflops"])
    assert out["min_gflops"] <= out["value"] <= out["max_gflops"]
    assert "beta(2,8)" in out["sample"]


def test_more_than_one_gpu_is_refused():
    res = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 2
    assert "one GPU" in res.stderr
