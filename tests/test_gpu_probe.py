"""The read-dominated streaming probe behind the measured ceiling: runs
over several tile sizes and rejects malformed sizes before any launch."""
import ctypes as C

import pytest
import torch

pytestmark = pytest.mark.gpu


def test_stream_probe_runs_and_checks_sizes():
    from spc5_amd._lib import lib
    src = torch.ones(1 << 20, dtype=torch.float64, device="cuda")
    dst = torch.zeros(4096, dtype=torch.float64, device="cuda")
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for kib in (1, 8, 64):
        rc = lib().spc5_hip_stream_probe(C.c_void_p(src.data_ptr()), src.numel() * 8, C.c_void_p(dst.data_ptr()),
                                         dst.numel() * 8, kib * 1024, s)
        assert rc == 0, lib().spc5_hip_last_error()
    torch.cuda.synchronize()
    assert not bool(torch.isnan(dst).any())
    for rb, wb, tile in ((100, 16, 1024), (1024, 16, 1000), (-16, 16, 1024)):
        assert lib().spc5_hip_stream_probe(C.c_void_p(src.data_ptr()), rb, C.c_void_p(dst.data_ptr()), wb, tile,
                                           s) != 0
