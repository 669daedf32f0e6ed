"""Property tests (hypothesis) over random small shapes.

CPU: identities of the oracle that hold for every input -- adjoint symmetry
<p, K(y, x) a> = <a, K(x, y) p>, invariance of the kernel sums under a common
rigid motion of both point sets (rotating the weights with applyDiffKT),
linearity in the weights.
GPU: the HIP kernel sums against the oracle for random target / source
counts (1-300, not multiples of any tile), dimensions 2 and 3, weight columns
1-5 (more than one 4-column group), one or two scales, every applyDiffKT mode,
FP64 (1e-11) and compensated FP32 (1e-5).  Examples are derandomized so that
every run draws the same shapes.
"""
import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

from oracle import oracle as O

SETTINGS = dict(max_examples=25, deadline=None, derandomize=True)


def shapes():
    return st.tuples(st.integers(1, 300), st.integers(1, 300), st.sampled_from([2, 3]), st.integers(1, 5),
                     st.sampled_from([(1.3,), (0.7,), (0.9, 2.2)]), st.integers(0, 2 ** 31 - 1))


def data(n, m, dim, k, seed):
    r = np.random.default_rng(seed)
    return (r.normal(size=(n, dim)) * 2, r.normal(size=(m, dim)) * 2, r.normal(size=(m, k)),
            r.normal(size=(n, k)))


def rotation(seed, dim):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(dim, dim)))
    return q if np.linalg.det(q) > 0 else -q


@settings(**SETTINGS)
@given(shapes())
def test_oracle_adjoint_symmetry(sh):
    n, m, dim, k, sig, seed = sh
    y, x, a, p = data(n, m, dim, k, seed)
    lhs = (p * O.applyK(y, x, a, list(sig))).sum()
    rhs = (a * O.applyK(x, y, p, list(sig))).sum()
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


@settings(**SETTINGS)
@given(shapes())
def test_oracle_rigid_invariance_and_linearity(sh):
    n, m, dim, k, sig, seed = sh
    y, x, a, p = data(n, m, dim, k, seed)
    R, t = rotation(seed + 1, dim), np.arange(1, dim + 1) * 0.7
    s = list(sig)
    np.testing.assert_allclose(O.applyK(y @ R.T + t, x @ R.T + t, a, s), O.applyK(y, x, a, s), rtol=1e-10,
                               atol=1e-12)
    if k == dim:
        # applyDiffKT is a vector in the point space: it turns with the points
        np.testing.assert_allclose(O.applyDiffKT(y @ R.T + t, x @ R.T + t, p, a, s), O.applyDiffKT(y, x, p, a, s) @ R.T,
                                   rtol=1e-9, atol=1e-11)
    a2 = np.random.default_rng(seed + 2).normal(size=a.shape)
    np.testing.assert_allclose(O.applyK(y, x, 2 * a - 3 * a2, s),
                               2 * O.applyK(y, x, a, s) - 3 * O.applyK(y, x, a2, s), rtol=1e-10, atol=1e-11)


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def ku():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lddmm_amd import kernelFunctions_util
    return kernelFunctions_util


TOL = {"float64": 1e-11, "float32": 1e-5}


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_hip_kernel_sums_random_shapes(ku, dtype):
    @settings(**SETTINGS)
    @given(shapes(), st.sampled_from(["plain", "sym", "lddmm"]))
    def check(sh, mode):
        n, m, dim, k, sig, seed = sh
        y, x, a, p = data(n, m, dim, k, seed)
        s = list(sig)
        got = ku.applyK(y, x, a, 'gauss', s, 10, dtype=dtype)
        assert rel(got, O.applyK(y, x, a, s)) < TOL[dtype], (sh, "applyK")
        if mode == "plain":
            got = ku.applyDiffKT(y, x, p, a, 'gauss', s, 10, dtype=dtype)
            ref = O.applyDiffKT(y, x, p, a, s)
        else:
            kw = {"sym": True} if mode == "sym" else {"lddmm": True, "regweight": 0.4}
            got = ku.applyDiffKT(x, x, a, a[::-1].copy(), 'gauss', s, 10, dtype=dtype, **kw)
            ref = O.applyDiffKT(x, x, a, a[::-1].copy(), s, **kw)
        assert rel(got, ref) < TOL[dtype], (sh, mode)
    check()
