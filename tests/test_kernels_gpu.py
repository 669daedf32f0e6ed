"""GPU parity of the Gaussian kernel sums against the FP64 oracle.

FP64 path: relative norm-wise error <= 1e-12 (both sides float64; the GPU
uses exp2 on scaled coordinates).  FP32 compensated path: <= 1e-5 norm-wise
relative error against the FP64 oracle, the bound the workload specification
states for reduced precision.  Every call goes through the C ABI.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {'float64': 1e-12, 'float32': 1e-5}


@pytest.fixture(scope="module")
def ku():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lddmm_amd import kernelFunctions_util as ku
    return ku


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def pts(rng, n, dim=3, scale=3.0, shift=0.0):
    return rng.normal(size=(n, dim)) * scale + shift


SHAPES = [  # n, m, dim, k, sigma
    (1, 1, 3, 3, [1.0]),
    (37, 53, 3, 3, [2.0]),
    (300, 700, 3, 1, [1.5]),
    (513, 257, 2, 2, [1.0]),
    (1000, 1000, 3, 6, [1.0, 2.5]),
    (2000, 3000, 3, 3, [2.5]),
    (700, 900, 3, 4, [0.7, 1.4, 3.0]),
]


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("n,m,dim,k,sig", SHAPES)
def test_applyK(ku, dtype, n, m, dim, k, sig):
    rng = np.random.default_rng(n * 7 + m)
    y, x, a = pts(rng, n, dim, shift=100.0), pts(rng, m, dim, shift=100.0), rng.normal(size=(m, k))
    got = ku.applyK(y, x, a, 'gauss', sig, 10, dtype=dtype)
    ref = O.applyK(y, x, a, sig)
    assert got.shape == ref.shape
    assert rel(got, ref) < TOL[dtype]


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("mode", ["plain", "sym", "lddmm"])
@pytest.mark.parametrize("n,m,dim,k,sig", SHAPES)
def test_applyDiffKT(ku, dtype, mode, n, m, dim, k, sig):
    if mode != "plain":
        m = n
    rng = np.random.default_rng(n + 11 * m)
    y = pts(rng, n, dim)
    x = y if mode != "plain" else pts(rng, m, dim)
    p, a = rng.normal(size=(n, k)), rng.normal(size=(m, k))
    kw = dict(sym=(mode == "sym"), lddmm=(mode == "lddmm"), regweight=0.6)
    got = ku.applyDiffKT(y, x, p, a, 'gauss', sig, 10, dtype=dtype, **kw)
    ref = O.applyDiffKT(y, x, p, a, sig, **kw)
    assert got.shape == ref.shape
    assert rel(got, ref) < TOL[dtype]


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("n,m,kb,sig", [(1, 1, 3, [1.]), (400, 650, 3, [1.2]), (900, 300, 2, [0.8, 2.]),
                                         (1500, 1500, 4, [1.0])])
def test_applyktensor(ku, dtype, n, m, kb, sig):
    rng = np.random.default_rng(n + m)
    y, x = pts(rng, n), pts(rng, m)
    ay, ax = rng.uniform(0.1, 1, n), rng.uniform(0.1, 1, m)
    by, bx = rng.normal(size=(n, kb)), rng.normal(size=(m, kb))
    got = ku.applyktensor(y, x, ay, ax, by, bx, 'gauss', sig, 10, dtype=dtype)
    ref = O.applyktensor(y, x, ay, ax, by, bx, sig)
    assert rel(got, ref) < TOL[dtype]


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("option", ["1", "2", "1and2"])
@pytest.mark.parametrize("n,k,sig", [(1, 3, [1.]), (333, 3, [1.3]), (1200, 2, [0.9, 2.0]), (600, 5, [1.0])])
def test_applyDiffK(ku, dtype, option, n, k, sig):
    rng = np.random.default_rng(n * 3 + k)
    x, a1, a2 = pts(rng, n), rng.normal(size=(n, 3)), rng.normal(size=(n, k))
    got = ku.applyDiffK(x, a1, a2, 'gauss', sig, 10, dtype=dtype, option=option)
    ref = O.applyDiffK(x, a1, a2, sig, option=option)
    assert rel(got, ref) < TOL[dtype] or np.abs(ref).max() == 0


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("option", ["11", "12", "11and12"])
@pytest.mark.parametrize("n,k,sig", [(2, 3, [1.]), (333, 3, [1.3]), (1200, 2, [0.9, 2.0]), (600, 5, [1.0])])
def test_applyDDiffK(ku, dtype, option, n, k, sig):
    rng = np.random.default_rng(n * 5 + k)
    # two points must stay within a few sigma: far apart, float32 underflows to 0
    x = pts(rng, n, scale=3.0 if n > 2 else 0.5)
    a1, a2, p = rng.normal(size=(n, k)), rng.normal(size=(n, k)), rng.normal(size=(n, 3))
    got = ku.applyDDiffK(x, a1, a2, p, 'gauss', sig, 10, dtype=dtype, option=option)
    ref = O.applyDDiffK(x, a1, a2, p, sig, option=option)
    assert rel(got, ref) < TOL[dtype]


@pytest.mark.parametrize("n,m,kb,sig", [(1, 1, 1, [1.]), (90, 130, 3, [1.2]), (200, 77, 1, [0.8, 2.])])
def test_dense_weight_sums(ku, n, m, kb, sig):
    rng = np.random.default_rng(n + m + kb)
    y, x = pts(rng, n), pts(rng, m)
    beta = rng.normal(size=(n, m, kb))
    assert rel(ku.applykmat(y, x, beta, 'gauss', sig, 10), O.applykmat(y, x, beta, sig)) < 1e-12
    b2 = rng.normal(size=(n, m))
    assert rel(ku.applykdiffmat(y, x, b2, 'gauss', sig, 10), O.applykdiffmat(y, x, b2, sig)) < 1e-12


def test_empty_sources_and_targets(ku):
    y = np.ones((5, 3))
    out = ku.applyK(y, np.zeros((0, 3)), np.zeros((0, 3)), 'gauss', [1.], 10)
    np.testing.assert_array_equal(out, np.zeros((5, 3)))
    out = ku.applyK(np.zeros((0, 3)), y, np.ones((5, 2)), 'gauss', [1.], 10)
    assert out.shape == (0, 2)
    out = ku.applyDiffKT(y, np.zeros((0, 3)), np.ones((5, 3)), np.zeros((0, 3)), 'gauss', [1.], 10)
    np.testing.assert_array_equal(out, np.zeros((5, 3)))


def test_far_apart_clouds_underflow_to_zero(ku):
    y = np.zeros((10, 3))
    x = np.full((10, 3), 1e4)
    for dt in ("float64", "float32"):
        out = ku.applyK(y, x, np.ones((10, 3)), 'gauss', [1.], 10, dtype=dt)
        assert np.all(np.isfinite(out)) and np.abs(out).max() == 0.0


def test_torch_inputs_stay_on_device(ku):
    import torch
    rng = np.random.default_rng(5)
    y, x, a = pts(rng, 100), pts(rng, 120), rng.normal(size=(120, 3))
    t = [torch.tensor(v, device="cuda") for v in (y, x, a)]
    out = ku.applyK(*t, 'gauss', [1.5], 10)
    assert isinstance(out, torch.Tensor) and out.is_cuda
    assert rel(out.cpu().numpy(), O.applyK(y, x, a, [1.5])) < 1e-12


def test_full_size_rows_match_oracle(ku):
    """Kernel-sum workload at full size (N = M = 50,000 uniform in [0,50]^3):
    400 random rows of the GPU result against the oracle over all sources."""
    from lddmm_amd import synth
    n = 50_000
    y = synth.uniform_cloud(n, seed=1)
    x = synth.uniform_cloud(n, seed=2)
    rng = np.random.default_rng(3)
    a = rng.normal(size=(n, 3))
    p = rng.normal(size=(n, 3))
    rows = np.sort(rng.choice(n, 400, replace=False))
    for sig in (2.5, 10.0):
        refK = O.applyK(y[rows], x, a, [sig])
        refD = O.applyDiffKT(y[rows], x, p[rows], a, [sig])
        for dt in ("float64", "float32"):
            gK = ku.applyK(y, x, a, 'gauss', [sig], 10, dtype=dt)[rows]
            gD = ku.applyDiffKT(y, x, p, a, 'gauss', [sig], 10, dtype=dt)[rows]
            assert rel(gK, refK) < TOL[dt], (sig, dt)
            assert rel(gD, refD) < TOL[dt], (sig, dt)


def test_kernel_class_dropin(ku):
    from lddmm_amd.kernelFunctions import Kernel
    rng = np.random.default_rng(9)
    x, y, a, p = pts(rng, 200), pts(rng, 150), rng.normal(size=(200, 3)), rng.normal(size=(150, 3))
    K = Kernel(name='gauss', sigma=1.7)
    assert rel(K.applyK(x, a), O.applyK(x, x, a, [1.7])) < 1e-12
    assert rel(K.applyK(x, a, firstVar=y), O.applyK(y, x, a, [1.7])) < 1e-12
    assert rel(K.applyDiffKT(x, p, a, firstVar=y), O.applyDiffKT(y, x, p, a, [1.7])) < 1e-12
    pa = rng.normal(size=(200, 3))
    assert rel(K.applyDiffKT(x, pa, a, regweight=0.3, lddmm=True),
               O.applyDiffKT(x, x, pa, a, [1.7], regweight=0.3, lddmm=True)) < 1e-12
    assert rel(K.applyDiffK1and2(x, pa, a), O.applyDiffK(x, pa, a, [1.7], option='1and2')) < 1e-12
    assert rel(K.applyDDiffK11and12(x, a, a, pa), O.applyDDiffK(x, a, a, pa, [1.7])) < 1e-12
    K.pk_dtype = 'float32'
    assert rel(K.applyK(x, a), O.applyK(x, x, a, [1.7])) < 1e-5
    # affine part follows the reference formula (base/kernelFunctions.py:476-487)
    Ka = Kernel(name='gauss', sigma=1.7, affine='affine', w1=0.3, w2=0.2)
    xx = x - 0.0
    expect = O.applyK(x, x, a, [1.7]) + 0.3 * (xx @ (xx.T @ a)) + 0.2 * a.sum(0)
    assert rel(Ka.applyK(x, a), expect) < 1e-12


@pytest.mark.parametrize("n,m,dim,k,sig", [(1, 1, 3, 3, [1.0]), (300, 700, 3, 1, [1.5]), (513, 257, 2, 2, [1.0]),
                                           (1000, 1000, 3, 6, [1.0, 2.5]), (2000, 3000, 3, 3, [2.5]),
                                           (700, 900, 3, 18, [1.2])])
def test_mfma_engine_parity(ku, n, m, dim, k, sig):
    """FP32 applyK / plain applyDiffKT with MFMA accumulation against the oracle."""
    from lddmm_amd import _backend as B
    rng = np.random.default_rng(n + 5 * m + k)
    y, x = pts(rng, n, dim, shift=20.0), pts(rng, m, dim, shift=20.0)
    a, p = rng.normal(size=(m, k)), rng.normal(size=(n, k))
    B.set_pair_engine('mfma')
    try:
        gK = ku.applyK(y, x, a, 'gauss', sig, 10, dtype='float32')
        gD = ku.applyDiffKT(y, x, p, a, 'gauss', sig, 10, dtype='float32')
    finally:
        B.set_pair_engine('valu')
    assert rel(gK, O.applyK(y, x, a, sig)) < 1e-5
    assert rel(gD, O.applyDiffKT(y, x, p, a, sig)) < 1e-5


def test_kernel_solve_matches_dense_solve(ku):
    """Kernel.solve with the textbook sequence (conjugate gradients on the
    device) against a dense LU solve of the Gram matrix; K(x, solve(xi)) must
    reproduce xi."""
    from lddmm_amd.kernelFunctions import Kernel
    rng = np.random.default_rng(21)
    x = rng.uniform(0, 12, size=(250, 3))
    xi = rng.normal(size=(250, 3))
    K = Kernel(name='gauss', sigma=1.0)
    a = K.solve(x, xi, None, iterMax=400, sequence='textbook')
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    G = np.exp(-d2 / 2.0)
    expect = np.linalg.solve(G, xi)
    assert isinstance(a, np.ndarray)
    assert rel(G @ a, xi) < 1e-6
    assert rel(a, expect) < 1e-4
    assert np.all(K.solve(x, np.zeros_like(xi), None) == 0)


def test_stream_handle_is_torch_current_stream(ku):
    """The C calls are issued on torch's current stream, also inside a side stream."""
    import torch
    from lddmm_amd import _backend as B
    assert (B.stream_handle().value or 0) == torch.cuda.current_stream().cuda_stream
    s = torch.cuda.Stream()
    rng = np.random.default_rng(8)
    y, a = rng.normal(size=(500, 3)), rng.normal(size=(500, 3))
    yd, ad = torch.tensor(y, device="cuda"), torch.tensor(a, device="cuda")
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        assert (B.stream_handle().value or 0) == s.cuda_stream
        out = ku.applyK(yd, yd, ad, 'gauss', [1.5], 10)
    torch.cuda.current_stream().wait_stream(s)
    assert rel(out.cpu().numpy(), O.applyK(y, y, a, [1.5])) < 1e-12
    t = torch.tensor(y, device="cuda")
    assert B.to_dev(t) is t                                   # already device float64 contiguous
    assert B.to_dev(t.T).is_contiguous() and B.to_dev(t.float()).dtype == torch.float64


@pytest.mark.parametrize("sigma", [1.0, 2.5])
def test_fp64_gaussian_factor_over_its_range(ku, sigma):
    """One source and targets at distances whose Gaussian goes from 1 down to
    the subnormal range (exponent e = r^2 / 2 sigma^2 up to 750): each output
    is a single factor exp(-e), checked element by element against numpy.
    The float64 path rounds the centred, scaled coordinates, which perturbs
    e by a few ulp of e itself, hence the e-proportional tolerance."""
    rng = np.random.default_rng(17)
    e = np.concatenate([[0.0, 1e-12, 0.5, 1.0], rng.uniform(0, 750, 4000)])
    r = np.sqrt(2 * e) * sigma
    dirs = rng.normal(size=(e.size, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    x = np.array([[3.0, -2.0, 1.0]])
    y = x + r[:, None] * dirs
    got = ku.applyK(y, x, np.ones((1, 1)), 'gauss', [sigma], 10, dtype='float64')[:, 0]
    ee = ((y - x) ** 2).sum(1) / (2 * sigma ** 2)
    ref = np.exp(-ee)
    tol = (2e-15 + 4e-15 * ee) * ref + 1e-322
    assert np.all(np.abs(got - ref) <= tol), np.max(np.abs(got - ref) / np.maximum(ref, 1e-300))
    assert got[0] == 1.0
