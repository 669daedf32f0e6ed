"""GPU parity of Kernel.solve (lddmm_kernel_solve: conjugate gradients with the
CG scalars on the device) against the oracle's restatement of linearcg
(base/conjugateGradient.py:351-408, reached from base/kernelFunctions.py:676-680).

Both sequences ('reference': linearcg as it executes, with its first direction
bound to the residual array; 'textbook') are compared iterate for iterate at
truncations 1-100 (the host polls the device stop flag every 16 products, so
truncations above 16 also cover an early stop), for one and several columns,
point counts above one workgroup's stride, a Laplacian and a multi-scale
kernel, the float32 product form, and the host loop used for kernels with an
affine part.  Beyond a few products, the comparisons use well conditioned
Gram matrices (point spacing about twice sigma): on badly conditioned ones CG
amplifies rounding differences of the products by many orders of magnitude
(checked on the CPU by perturbing the oracle's products by 1e-15).  FP64 tolerance 1e-9 relative up to 20 products and 1e-7 for
the run of about 50-60 products that stops early (CG amplifies the
products' rounding over the iterations), FP32 products 1e-4.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lddmm_amd.kernelFunctions import Kernel
    return dict(torch=torch, Kernel=Kernel)


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def problem(n, k, seed, box=8.0):
    r = np.random.default_rng(seed)
    return r.uniform(0, box, size=(n, 3)), r.normal(size=(n, k))


@pytest.mark.parametrize("sequence", ["reference", "textbook"])
@pytest.mark.parametrize("iters,sigma,box", [(1, 1.5, 8.0), (2, 1.5, 8.0), (10, 1.5, 8.0), (17, 1.5, 8.0),
                                             (100, 0.5, 16.0)])
def test_solve_matches_oracle(env, sequence, iters, sigma, box):
    """Up to 17 products on a badly conditioned Gram matrix (sigma 1.5 at a
    point spacing of about 0.9); 100 on a well conditioned one (sigma 0.5 at
    a spacing of about 1.8), where the iteration stops early (after about 50
    products) on the oracle and on the device alike; over a badly
    conditioned matrix 100 products amplify the products' rounding
    differences to 1e-3."""
    K = env["Kernel"](name='gauss', sigma=sigma)
    x, b = problem(700, 3, 3, box=box)
    got = K.solve(x, b, None, iterMax=iters, sequence=sequence)
    ref, products = O.kernel_solve(x, b, [sigma], iter_max=iters, sequence=sequence)
    print(f"[solve] {sequence} iters {iters}: oracle products {products}, rel {rel(got, ref):.2e}", flush=True)
    assert isinstance(got, np.ndarray) and got.shape == b.shape
    assert rel(got, ref) < (1e-9 if iters < 100 else 1e-7)
    if iters == 100:
        assert products < 100


def test_solve_default_is_reference_sequence(env):
    K = env["Kernel"](name='gauss', sigma=1.5)
    x, b = problem(300, 3, 4)
    ref, _ = O.kernel_solve(x, b, [1.5], iter_max=12, sequence='reference')
    assert rel(K.solve(x, b, None, iterMax=12), ref) < 1e-9


@pytest.mark.parametrize("case", ["k1_large", "laplacian", "multiscale", "vector_rhs"])
def test_solve_shapes_and_kernels(env, case):
    torch, Kernel = env["torch"], env["Kernel"]
    if case == "k1_large":
        K, sig, kern = Kernel(name='gauss', sigma=2.0), [2.0], O.GAUSS
        x, b = problem(5000, 1, 5, box=30.0)
    elif case == "laplacian":
        K, sig, kern = Kernel(name='laplacian', sigma=0.7, order=3), [0.7], ('laplacian', 3)
        x, b = problem(600, 3, 6, box=16.0)
    elif case == "multiscale":
        K, sig, kern = Kernel(name='gauss', sigma=[0.5, 1.0]), [0.5, 1.0], O.GAUSS
        x, b = problem(600, 2, 7, box=16.0)
    else:
        K, sig, kern = Kernel(name='gauss', sigma=0.8), [0.8], O.GAUSS
        x, b = problem(500, 1, 8, box=16.0)
        b = b[:, 0]
    for seq in ("reference", "textbook"):
        got = K.solve(torch.tensor(x, device='cuda'), torch.tensor(b, device='cuda'), None, iterMax=20,
                      sequence=seq)
        assert isinstance(got, torch.Tensor) and tuple(got.shape) == b.shape
        bb = b if b.ndim == 2 else b[:, None]
        ref, _ = O.linearcg(lambda v: O.applyK(x, x, v, sig, kernel=kern), bb, 20, seq)
        assert rel(got.cpu().numpy().reshape(bb.shape), ref) < 1e-9, (case, seq)


def test_solve_float32_products(env):
    K = env["Kernel"](name='gauss', sigma=0.8)
    K.pk_dtype = 'float32'
    x, b = problem(700, 3, 9, box=16.0)
    for seq in ("reference", "textbook"):
        got = K.solve(x, b, None, iterMax=10, sequence=seq)
        ref, _ = O.kernel_solve(x, b, [0.8], iter_max=10, sequence=seq)
        assert rel(got, ref) < 1e-4, seq


def test_solve_affine_kernel_host_loop(env):
    """With an affine part the products are Kernel.applyK calls (kernel sum +
    affine terms); the sequence is the same."""
    Kernel = env["Kernel"]
    K = Kernel(name='gauss', sigma=0.8)
    K.setAffine('euclidean', w1=0.3, w2=0.2, center=[8.0, 8.0, 8.0])
    x, b = problem(400, 3, 10, box=16.0)
    kaff = ('euclidean', 0.3, 0.2, np.array([8.0, 8.0, 8.0]))
    op = lambda v: O.applyK(x, x, v, [0.8]) + O.affineApplyK(x, x, v, kaff)
    for seq in ("reference", "textbook"):
        got = K.solve(x, b, None, iterMax=6, sequence=seq)
        ref, _ = O.linearcg(op, b, 6, seq)
        assert rel(got, ref) < 1e-9, seq


def test_solve_zero_and_tiny_rhs(env):
    K = env["Kernel"](name='gauss', sigma=1.5)
    x, b = problem(200, 3, 11)
    for seq in ("reference", "textbook"):
        assert np.all(K.solve(x, np.zeros_like(b), None, sequence=seq) == 0)
        assert np.all(K.solve(x, 1e-7 * b, None, sequence=seq) == 0)
