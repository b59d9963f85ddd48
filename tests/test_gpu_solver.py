"""Precision form on the GPU: reduced-precision value copies, the general
SpMV epilogue, CG and mixed-precision IR-CG.  No reference counterpart;
pinned by numpy float64:

* value rounding: bit-identical to torch's double->float and
  double->bfloat16 conversions (both round to nearest even through float);
* float / bfloat16 SpMV: within 1e-5 (normwise, float accumulation) of the
  float64 product with the rounded values, and within 2^-8 of the exact
  double matrix for bfloat16 (value rounding bound);
* y = alpha A x + beta y (double): within 1e-12 normwise of numpy;
* CG / IR-CG: true relative residual, recomputed in numpy float64, below
  1e-10, and the solution within 1e-7 of numpy.linalg.solve.
"""
import numpy as np
import pytest
import torch

from conftest import normwise_error
from matgen import clustered_csr, laplacian2d

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def csr_dense(rp, col, val, ncols):
    n = len(rp) - 1
    A = np.zeros((n, ncols))
    for i in range(n):
        A[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    return A


def stencil27_small(n=12, diag=26.001):
    from spc5_amd import synth
    return synth.stencil27(n, diag=diag)


@pytest.mark.parametrize("r", (1, 2, 4, 8))
@pytest.mark.parametrize("row_sort", [False, True])
def test_lowp_values_and_spmv(r, row_sort):
    from spc5_amd.device import DeviceMatrix
    from spc5_amd.solver import lowp
    rp, col, val = clustered_csr(700, 900, seed=r, empty_frac=0.05)
    A64 = DeviceMatrix.from_csr(dev(rp), dev(col), dev(val), r, ncols=900)
    v64 = A64.values.cpu()
    x = np.random.default_rng(r).uniform(-1, 1, 900)
    Ad = csr_dense(rp, col, val, 900)
    for kind, dt, tol_exact in (("f32", torch.float32, 1e-6), ("bf16", torch.bfloat16, 2.0 ** -8)):
        Alo = lowp(A64, kind)
        Alo.row_sort = row_sort
        assert Alo.values.dtype == dt
        ref_bits = v64.to(dt)
        assert torch.equal(Alo.values.cpu().view(torch.int16 if dt == torch.bfloat16 else torch.int32),
                           ref_bits.view(torch.int16 if dt == torch.bfloat16 else torch.int32))
        # the rounded matrix in float64, same permutation as the stream
        rounded = np.zeros_like(val)
        order_vals = ref_bits.to(torch.float64).numpy()
        # decode permuted order back to CSR order through the oracle
        import oracle as O
        bptr = A64.rows_size.cpu().numpy()
        rows, cols, _ = O.decode(r, 700, bptr, A64.headers.cpu().numpy(), v64.numpy())
        Alo_d = np.zeros((700, 900))
        Alo_d[rows, cols] = order_vals
        xf = x.astype(np.float32)
        y = torch.zeros(700, dtype=torch.float32, device="cuda")
        Alo.spmv(dev(xf), y)
        yref = Alo_d @ xf.astype(np.float64)
        assert normwise_error(y.cpu().numpy(), yref, rp, val, xf, np.zeros(1)) <= 1e-5
        assert normwise_error(y.cpu().numpy(), Ad @ xf.astype(np.float64), rp, val, xf, np.zeros(1)) <= tol_exact
        del rounded


@pytest.mark.parametrize("r", (2, 4))
@pytest.mark.parametrize("kind", ["f32", "bf16"])
@pytest.mark.parametrize("row_sort", [False, True])
def test_lowp_stencil_row_sort(r, kind, row_sort):
    """Float and bfloat16 copies of the 27-point stencil (the structure the
    row-sorted batches are for): the product equals the rounded matrix's
    to float accumulation error."""
    from spc5_amd.device import DeviceMatrix
    from spc5_amd.solver import lowp
    rp, col, val = stencil27_small(16)
    n = len(rp) - 1
    A64 = DeviceMatrix.from_csr(dev(rp), dev(col), dev(val), r)
    Alo = lowp(A64, kind)
    Alo.row_sort = row_sort
    dt = torch.float32 if kind == "f32" else torch.bfloat16
    rounded = torch.from_numpy(val).to(dt).to(torch.float64).numpy()  # CSR order, same rounding
    xf = np.random.default_rng(r).uniform(-1, 1, n).astype(np.float32)
    y = torch.zeros(n, dtype=torch.float32, device="cuda")
    Alo.spmv(dev(xf), y)
    yref = csr_dense(rp, col, rounded, n) @ xf.astype(np.float64)
    assert normwise_error(y.cpu().numpy(), yref, rp, val, xf, np.zeros(1)) <= 1e-5


@pytest.mark.parametrize("alpha,beta", [(1.0, 1.0), (1.0, 0.0), (-1.0, 1.0), (2.5, -0.5)])
def test_spmv_ex_double(alpha, beta):
    from spc5_amd.device import DeviceMatrix
    from spc5_amd.solver import spmv_ex
    rp, col, val = clustered_csr(500, 500, seed=7)
    A = DeviceMatrix.from_csr(dev(rp), dev(col), dev(val), 4)
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, 500)
    y0 = rng.uniform(-1, 1, 500)
    y = dev(y0.copy())
    spmv_ex(A, dev(x), y, alpha, beta)
    yref = alpha * (csr_dense(rp, col, val, 500) @ x) + beta * y0
    assert normwise_error(y.cpu().numpy(), yref, rp, val * max(1, abs(alpha)), x, y0) <= 1e-12


@pytest.mark.parametrize("r", (1, 4, 8))
def test_cg_double(r):
    from spc5_amd.device import DeviceMatrix
    from spc5_amd.solver import cg
    rp, col, val = laplacian2d(24)
    n = len(rp) - 1
    A = DeviceMatrix.from_csr(dev(rp), dev(col), dev(val), r)
    b = np.random.default_rng(r).uniform(-1, 1, n)
    x, st = cg(A, dev(b), rtol=1e-10, maxit=5000)
    xh = x.cpu().numpy()
    Ad = csr_dense(rp, col, val, n)
    relres = np.linalg.norm(b - Ad @ xh) / np.linalg.norm(b)
    assert relres <= 1e-10, (relres, st)
    assert st["relres"] == pytest.approx(relres, rel=1e-3, abs=1e-14)
    xs = np.linalg.solve(Ad, b)
    assert np.linalg.norm(xh - xs) / np.linalg.norm(xs) <= 1e-7


def test_cg_many_tiles_matches_recurrence():
    """A 5.29M-row Laplacian has more tiles (one p.q partial each) than one
    pass of the partial-sum kernel covers (32,768): three CG iterations
    equal the textbook recurrence run in torch on the same (separately
    verified) product, to summation-order rounding."""
    from spc5_amd import synth
    from spc5_amd.device import DeviceMatrix
    from spc5_amd.solver import cg
    rp, col, val = synth.laplacian2d(2300)
    n = len(rp) - 1
    A = DeviceMatrix.from_csr(dev(rp), dev(col), dev(val), 1)
    assert n // 128 > 32768
    b = dev(synth.uniform(n, 21))
    x, st = cg(A, b, rtol=1e-14, maxit=3)
    assert st["iterations"] == 3

    def mul(v):
        out = torch.zeros_like(v)
        A.spmv(v, out)
        return out

    xr = torch.zeros_like(b)
    r = b.clone()
    p = r.clone()
    rr = torch.dot(r, r)
    for _ in range(3):
        q = mul(p)
        alpha = rr / torch.dot(p, q)
        xr += alpha * p
        r -= alpha * q
        rr_new = torch.dot(r, r)
        p = r + (rr_new / rr) * p
        rr = rr_new
    torch.cuda.synchronize()
    assert float(torch.linalg.vector_norm(x - xr) / torch.linalg.vector_norm(xr)) <= 1e-12
    assert st["relres"] == pytest.approx(float(torch.sqrt(rr) / torch.linalg.vector_norm(b)), rel=1e-9)


@pytest.mark.parametrize("kind", ["f32", "bf16"])
@pytest.mark.parametrize("r", (2, 4))
def test_ir_cg_mixed(kind, r):
    from spc5_amd.device import DeviceMatrix
    from spc5_amd.solver import ir_cg, lowp
    rp, col, val = stencil27_small(12)
    n = len(rp) - 1
    A64 = DeviceMatrix.from_csr(dev(rp), dev(col), dev(val), r)
    Alo = lowp(A64, kind)
    b = np.random.default_rng(5).uniform(-1, 1, n)
    x, st = ir_cg(A64, Alo, dev(b), rtol=1e-10)
    xh = x.cpu().numpy()
    Ad = csr_dense(rp, col, val, n)
    relres = np.linalg.norm(b - Ad @ xh) / np.linalg.norm(b)
    assert relres <= 1e-10, (relres, st)
    assert st["outer"] >= 1
    xs = np.linalg.solve(Ad, b)
    assert np.linalg.norm(xh - xs) / np.linalg.norm(xs) <= 1e-7


@pytest.mark.parametrize("tail", [None, 0.0, 1e-6])
@pytest.mark.parametrize("vectors", ["f64", "f32"])
@pytest.mark.parametrize("kind", ["f32", "bf16"])
@pytest.mark.parametrize("r", (1, 4))
def test_mixed_cg_reliable_updates(kind, r, vectors, tail):
    """Reliable-update CG: float iteration, double residual/solution, the
    true residual below 1e-10 and the solution equal to a dense solve.
    tail None is the default: a double tail for bfloat16 values only."""
    from spc5_amd.device import DeviceMatrix
    from spc5_amd.solver import lowp, mixed_cg
    rp, col, val = stencil27_small(12)
    n = len(rp) - 1
    A64 = DeviceMatrix.from_csr(dev(rp), dev(col), dev(val), r)
    Alo = lowp(A64, kind)
    b = np.random.default_rng(6).uniform(-1, 1, n)
    x, st = mixed_cg(A64, Alo, dev(b), rtol=1e-10, vectors=vectors, tail_rtol=tail)
    xh = x.cpu().numpy()
    if tail == 0.0 or (tail is None and kind == "f32"):
        assert st["tail_iterations"] == 0
    else:
        assert 0 < st["tail_iterations"] < st["iterations"]
    Ad = csr_dense(rp, col, val, n)
    relres = np.linalg.norm(b - Ad @ xh) / np.linalg.norm(b)
    assert relres <= 1e-10, (relres, st)
    assert st["relres"] == pytest.approx(relres, rel=1e-2, abs=1e-13)
    assert st["outer"] >= 2  # at least one intermediate update besides the last
    xs = np.linalg.solve(Ad, b)
    assert np.linalg.norm(xh - xs) / np.linalg.norm(xs) <= 1e-7


def test_mixed_cg_tail_with_another_r_and_many_tiles():
    """The double tail runs A64's product, whose tile partials go to the
    buffer the float product uses: with A64 at another r than the float
    matrix, more tiles than the float matrix and more than 2,048 tiles (the
    vector kernels' partial count), the buffer must hold A64's longer list."""
    import scipy.sparse as sp

    from spc5_amd.device import DeviceMatrix
    from spc5_amd.solver import lowp, mixed_cg
    rp, col, val = stencil27_small(64)
    n = len(rp) - 1
    A64 = DeviceMatrix.from_csr(dev(rp), dev(col), dev(val), 1)
    A64.tile_values = 1024
    base = DeviceMatrix.from_csr(dev(rp), dev(col), dev(val), 4)
    base.tile_values = 2048
    Alo = lowp(base, "f32")
    nt64, ntlo = A64.info()["ntiles"], Alo.info()["ntiles"]
    assert nt64 > ntlo > 2048, (nt64, ntlo)
    b = np.random.default_rng(8).uniform(-1, 1, n)
    x, st = mixed_cg(A64, Alo, dev(b), rtol=1e-10, tail_rtol=1e-6)
    assert st["tail_iterations"] > 0
    A = sp.csr_matrix((val, col, rp), shape=(n, n))
    xh = x.cpu().numpy()
    relres = np.linalg.norm(b - A @ xh) / np.linalg.norm(b)
    assert relres <= 1e-10, (relres, st)


def test_cg_graph_replay_matches_direct_launches(monkeypatch):
    """The CG chunk recorded into a hipGraph and replayed gives the same
    iterates, bit for bit, as launching every kernel directly (double CG and
    the reliable-update solver), over many replays."""
    from spc5_amd.device import DeviceMatrix
    from spc5_amd.solver import cg, lowp, mixed_cg
    rp, col, val = stencil27_small(12)
    n = len(rp) - 1
    A64 = DeviceMatrix.from_csr(dev(rp), dev(col), dev(val), 4)
    Alo = lowp(A64, "f32")
    b = dev(np.random.default_rng(11).uniform(-1, 1, n))
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("SPC5_CG_GRAPH", flag)
        x1, s1 = cg(A64, b, rtol=1e-12, maxit=500)
        x2, s2 = mixed_cg(A64, Alo, b, rtol=1e-11, vectors="f32")
        x3, s3 = mixed_cg(A64, Alo, b, rtol=1e-11, vectors="f64")
        torch.cuda.synchronize()
        out[flag] = (x1.cpu().numpy(), s1, x2.cpu().numpy(), s2, x3.cpu().numpy(), s3)
    g, d = out["1"], out["0"]
    assert g[1]["iterations"] == d[1]["iterations"] >= 16
    assert np.array_equal(g[0], d[0])
    for i in (2, 4):
        assert g[i + 1]["iterations"] == d[i + 1]["iterations"] and g[i + 1]["outer"] == d[i + 1]["outer"]
        assert np.array_equal(g[i], d[i])


def test_mixed_cg_nonzero_guess_and_limits():
    """A nonzero initial guess is refined, maxit=0 leaves x alone, bad
    arguments are rejected."""
    from spc5_amd._lib import Spc5Error
    from spc5_amd.device import DeviceMatrix
    from spc5_amd.solver import lowp, mixed_cg
    rp, col, val = stencil27_small(8)
    n = len(rp) - 1
    A64 = DeviceMatrix.from_csr(dev(rp), dev(col), dev(val), 2)
    Alo = lowp(A64, "f32")
    b = np.random.default_rng(8).uniform(-1, 1, n)
    x0 = np.random.default_rng(9).uniform(-1, 1, n)
    x = dev(x0.copy())
    x, st = mixed_cg(A64, Alo, dev(b), x=x, rtol=1e-10, maxit=0)
    assert np.array_equal(x.cpu().numpy(), x0) and st["iterations"] == 0
    x, st = mixed_cg(A64, Alo, dev(b), x=x, rtol=1e-10)
    Ad = csr_dense(rp, col, val, n)
    assert np.linalg.norm(b - Ad @ x.cpu().numpy()) / np.linalg.norm(b) <= 1e-10
    with pytest.raises(Spc5Error):
        mixed_cg(A64, Alo, dev(b), rtol=1e-10, delta=1.5)
    with pytest.raises(Spc5Error):
        mixed_cg(A64, Alo, dev(b), rtol=1e-10, vectors="f16")
    with pytest.raises(Spc5Error):
        mixed_cg(A64, A64, dev(b), rtol=1e-10)


def test_cg_float_reference_format():
    """Float CG on the reference float format (VS=16): converges to 1e-5."""
    from spc5_amd.device import DeviceMatrix
    from spc5_amd.solver import cg
    rp, col, val = laplacian2d(16)
    n = len(rp) - 1
    A = DeviceMatrix.from_csr(dev(rp), dev(col), dev(val.astype(np.float32)), 2)
    assert A.vs == 16
    b = np.random.default_rng(2).uniform(-1, 1, n).astype(np.float32)
    x, st = cg(A, dev(b), rtol=1e-5, maxit=2000)
    Ad = csr_dense(rp, col, val, n)
    xh = x.cpu().numpy().astype(np.float64)
    assert np.linalg.norm(b - Ad @ xh) / np.linalg.norm(b) <= 2e-5


def test_solvers_full_size_true_residual():
    """The precision-form workload at full size (27-point stencil 160^3,
    4,096,000 rows, b uniform in [-1, 1] with seed 7): double CG and the
    reliable-update CG with float values, both on row-sorted beta(2,8), reach
    a true relative residual ||b - A x|| / ||b|| (scipy, float64, on the
    host) below 1e-10 with a little rounding headroom, and the residual each
    solver reports agrees with it."""
    import scipy.sparse as sp

    from spc5_amd import synth
    from spc5_amd.device import DeviceMatrix
    from spc5_amd.solver import cg, lowp, mixed_cg
    rp, col, val = synth.stencil27(160)
    n = len(rp) - 1
    b = synth.uniform(n, 7)
    A64 = DeviceMatrix.from_csr(dev(rp), dev(col), dev(val), 2)
    A64.row_sort = True
    Alo = lowp(A64, "f32")
    Alo.row_sort = True
    A = sp.csr_matrix((val, col, rp), shape=(n, n))
    bn = np.linalg.norm(b)
    x, st = cg(A64, dev(b), rtol=1e-10)
    rel = np.linalg.norm(b - A @ x.cpu().numpy()) / bn
    assert rel <= 1.2e-10 and st["relres"] == pytest.approx(rel, rel=0.05), (rel, st)
    x, st = mixed_cg(A64, Alo, dev(b), rtol=1e-10)
    rel = np.linalg.norm(b - A @ x.cpu().numpy()) / bn
    assert rel <= 1.2e-10 and st["relres"] == pytest.approx(rel, rel=0.05), (rel, st)
    assert st["outer"] >= 2 and st["tail_iterations"] == 0
