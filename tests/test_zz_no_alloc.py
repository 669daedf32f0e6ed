"""lddmm_applyK_noalloc (applyK with its partial sums in a static device
array) against lddmm_applyK, called through the C ABI.

When both plan the same source chunks the two run the same kernels on the
same data, so the results are compared bit for bit; a plan that the static
array cannot hold runs with fewer chunks (another summation order) and is
compared to the usual tolerances (FP64 1e-12, FP32 1e-5 against the FP64
oracle).  Two calls in a row, the second smaller, show that nothing carries
over between calls; sizes beyond the array are rejected before any launch.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {'float64': 1e-12, 'float32': 1e-5}


@pytest.fixture(scope="module")
def B():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lddmm_amd import _backend
    return _backend


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _dev(B, arr):
    return B.to_dev(np.ascontiguousarray(arr, dtype=np.float64))


def _call(B, fn, kd, y, x, a, k, dim):
    import torch
    n, m = y.shape[0], x.shape[0]
    out = torch.full((n, k), float('nan'), dtype=torch.float64, device=y.device)
    rc = fn(ctypes.byref(kd), B.ptr(y), n, B.ptr(x), m, dim, B.ptr(a), k, B.ptr(out), B.stream_handle())
    B.check(rc, fn.__name__)
    torch.cuda.synchronize()
    return out


KERNELS = [  # name, sigma, kind, order
    ("gauss", [2.5], 0, 0),
    ("gauss2", [1.0, 3.0], 0, 0),
    ("laplacian3", [2.0], 1, 3),
]


@pytest.mark.parametrize("prec", ["float64", "float32"])
@pytest.mark.parametrize("kname,sigma,kind,order", KERNELS)
def test_noalloc_matches_applyK_two_calls(B, prec, kname, sigma, kind, order):
    rng = np.random.default_rng(11)
    kd = B.kernel_desc(sigma, -1.0, prec, kind, order)
    # first call large (two weight groups for k = 5), second smaller
    for (n, m, dim, k) in ((9000, 7000, 3, 5), (700, 900, 2, 3)):
        print(f"[noalloc] {kname} {prec} n={n} m={m} dim={dim} k={k}", flush=True)
        y = rng.normal(size=(n, dim)) * 4 + 50
        x = rng.normal(size=(m, dim)) * 4 + 50
        a = rng.normal(size=(m, k))
        yd, xd, ad = _dev(B, y), _dev(B, x), _dev(B, a)
        got = _call(B, B.LIB.lddmm_applyK_noalloc, kd, yd, xd, ad, k, dim).cpu().numpy()
        want = _call(B, B.LIB.lddmm_applyK, kd, yd, xd, ad, k, dim).cpu().numpy()
        assert np.array_equal(got, want)
        rows = rng.choice(n, size=min(n, 300), replace=False)
        ref = O.applyK(y[rows], x, a, sigma, kernel=('laplacian', order) if kind else O.GAUSS)
        assert rel(got[rows], ref) < TOL[prec]


def test_noalloc_allocates_nothing(B):
    import torch
    rng = np.random.default_rng(3)
    kd = B.kernel_desc([2.0], -1.0, 'float32')
    small = [_dev(B, rng.normal(size=s)) for s in ((500, 3), (400, 3), (400, 3))]
    _call(B, B.LIB.lddmm_applyK_noalloc, kd, small[0], small[1], small[2], 3, 3)   # first call per device
    B.check(B.LIB.lddmm_release_workspace(), "release_workspace")
    n, m = 200_000, 20_000                # the workspace path would need ~10 MB of partials
    y, x, a = (_dev(B, rng.normal(size=s) * 10) for s in ((n, 3), (m, 3), (m, 3)))
    out = torch.empty((n, 3), dtype=torch.float64, device=y.device)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    rc = B.LIB.lddmm_applyK_noalloc(ctypes.byref(kd), B.ptr(y), n, B.ptr(x), m, 3, B.ptr(a), 3, B.ptr(out),
                                    B.stream_handle())
    B.check(rc, "applyK_noalloc")
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    want = _call(B, B.LIB.lddmm_applyK, kd, y, x, a, 3, 3)
    assert torch.equal(out, want)


@pytest.mark.parametrize("prec", ["float64", "float32"])
def test_noalloc_large_targets(B, prec):
    """800k targets x 4 weight columns x 40k sources: applyK's plan takes 3 to
    5 source chunks (77-128 MB of partials) on most occupancies, more than
    the static array holds, so the no-allocation plan uses fewer chunks;
    compared to applyK and, on sampled rows, to the oracle."""
    rng = np.random.default_rng(5)
    n, m, k = 800_000, 40_000, 4
    print(f"[noalloc] large {prec} n={n} m={m} k={k}", flush=True)
    kd = B.kernel_desc([3.0], -1.0, prec)
    y = rng.uniform(0, 60, size=(n, 3))
    x = rng.uniform(0, 60, size=(m, 3))
    a = rng.normal(size=(m, k))
    yd, xd, ad = _dev(B, y), _dev(B, x), _dev(B, a)
    got = _call(B, B.LIB.lddmm_applyK_noalloc, kd, yd, xd, ad, k, 3).cpu().numpy()
    want = _call(B, B.LIB.lddmm_applyK, kd, yd, xd, ad, k, 3).cpu().numpy()
    assert rel(got, want) < (1e-13 if prec == 'float64' else 1e-6)
    rows = rng.choice(n, size=400, replace=False)
    assert rel(got[rows], O.applyK(y[rows], x, a, [3.0])) < TOL[prec]


def test_noalloc_rejects_sizes_beyond_scratch(B):
    import torch
    rng = np.random.default_rng(9)
    n, m, k = 3_000_000, 64, 3           # 3M x 3 x 8 B = 72 MB of partials for one chunk
    kd = B.kernel_desc([1.0], -1.0, 'float64')
    y = torch.zeros((n, 3), dtype=torch.float64, device='cuda')
    x, a = _dev(B, rng.normal(size=(m, 3))), _dev(B, rng.normal(size=(m, k)))
    out = torch.full((n, k), 7.0, dtype=torch.float64, device='cuda')
    rc = B.LIB.lddmm_applyK_noalloc(ctypes.byref(kd), B.ptr(y), n, B.ptr(x), m, 3, B.ptr(a), k, B.ptr(out),
                                    B.stream_handle())
    assert rc == B.EINVAL
    assert b"static scratch" in B.LIB.lddmm_last_error()
    torch.cuda.synchronize()
    assert bool((out == 7.0).all())
