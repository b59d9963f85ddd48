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


def test_access_probe_modes_and_arguments():
    """The access probe runs its three modes over a double beta(1,8) matrix's
    tiles, leaves its output alone, and refuses other formats, unknown modes
    and a window wider than the matrix."""
    import numpy as np

    from spc5_amd import synth
    from spc5_amd._lib import lib
    from spc5_amd.device import DeviceMatrix
    rp, col, val = synth.clustered_runs(20000, seed=5, value_seed=6)
    d = [torch.from_numpy(a).cuda() for a in (rp, col, val)]
    x = torch.from_numpy(np.linspace(-1.0, 1.0, 20000)).cuda()
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    m = DeviceMatrix.from_csr(*d, 1)
    for mode in (0, 1, 2):
        rc = lib().spc5_hip_access_probe(m._h, mode, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), s)
        assert rc == 0, lib().spc5_hip_last_error()
    torch.cuda.synchronize()
    assert float(out[0]) == 0.0
    assert lib().spc5_hip_access_probe(m._h, 3, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), s) != 0
    m2 = DeviceMatrix.from_csr(*d, 2)
    assert lib().spc5_hip_access_probe(m2._h, 0, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), s) != 0
    m32 = DeviceMatrix.from_csr(d[0], d[1], d[2].to(torch.float32), 1)
    assert lib().spc5_hip_access_probe(m32._h, 0, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), s) != 0
    small = [torch.from_numpy(a).cuda() for a in synth.clustered_runs(500, seed=5, value_seed=6)]
    ms = DeviceMatrix.from_csr(*small, 1)
    assert lib().spc5_hip_access_probe(ms._h, 2, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), s) != 0
    assert lib().spc5_hip_access_probe(ms._h, 1, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), s) == 0
    torch.cuda.synchronize()
