"""SecondOrderPointSetMatching (geodesic shooting from the initial momentum
a0) on the GPU against the float64 oracle: objective and gradient at a
non-zero a0 (point-set measure and L2 data terms, float64 and compensated
float32), the initialMomentum option, and a registration that lowers the
objective with every accepted step and ends where the oracle agrees."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lddmm_amd import synth
    from lddmm_amd.pointSets import PointSet
    from lddmm_amd.kernelFunctions import Kernel
    from lddmm_amd.secondOrderPointSetMatching import SecondOrderPointSetMatching
    return synth, PointSet, Kernel, SecondOrderPointSetMatching, torch


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def problem(seed=8, n=300):
    from lddmm_amd import synth
    rng = np.random.default_rng(seed)
    x0 = synth.fibonacci_sphere(n, 5.0)
    w0 = rng.uniform(0.5, 1.5, size=n)
    y = x0 * 1.1 + 0.2 * rng.normal(size=x0.shape)
    wy = rng.uniform(0.5, 1.5, size=n)
    return rng, x0, w0, y, wy


def make(mods, kind, dtype, x0, w0, y, wy, **extra):
    synth, PointSet, Kernel, SOPM, torch = mods
    opts = {'KparDiff': Kernel(name='gauss', sigma=2.0), 'KparDist': Kernel(name='gauss', sigma=1.0),
            'errorType': kind, 'timeStep': 0.1, 'pk_dtype': dtype, 'mode': 'quiet', 'maxIter': 15,
            'sigmaError': 0.8}
    opts.update(extra)
    return SOPM(Template=PointSet(x0, weights=w0), Target=PointSet(y, weights=wy), options=opts)


def oracle_pair(kind, x0, w0, a0, y, wy):
    tgt = (y, wy) if kind == 'measure' else y
    k = 'point_measure' if kind == 'measure' else 'L2'
    obj, _ = O.second_order_matching_objective(x0, w0, a0, tgt, [2.0], [1.0], 0.1, kind=k, sigma_error=0.8)
    g = O.second_order_matching_gradient(x0, w0, a0, tgt, [2.0], [1.0], 0.1, kind=k, sigma_error=0.8)
    return obj, g


@pytest.mark.parametrize("dtype,tol", [("float64", 1e-10), ("float32", 2e-5)])
@pytest.mark.parametrize("kind", ["measure", "L2"])
def test_objective_and_gradient_match_oracle(mods, kind, dtype, tol):
    torch = mods[-1]
    rng, x0, w0, y, wy = problem()
    a0 = 0.05 * rng.normal(size=x0.shape)
    m = make(mods, kind, dtype, x0, w0, y, wy)
    m.control['a0'] = torch.tensor(a0, device='cuda')
    m.obj = None
    obj = m.objectiveFun()
    g = m.getGradient()['a0'].cpu().numpy()
    ref, gr = oracle_pair(kind, x0, w0, a0, y, wy)
    print(f"[second order] {kind} {dtype}: objective rel err {abs(obj - ref) / abs(ref):.2e}, "
          f"gradient rel err {rel(g, gr):.2e}", flush=True)
    assert abs(obj - ref) / abs(ref) < tol
    assert rel(g, gr) < tol


def test_initial_momentum_option(mods):
    rng, x0, w0, y, wy = problem(seed=3)
    a0 = 0.05 * rng.normal(size=x0.shape)
    m = make(mods, 'measure', 'float64', x0, w0, y, wy, initialMomentum=a0)
    ref, _ = oracle_pair('measure', x0, w0, a0, y, wy)
    assert abs(m.objectiveFun() - ref) / abs(ref) < 1e-10
    xt = m.state['xt'].cpu().numpy()
    xr, _ = O.secondOrderEvolution(x0, a0, [2.0], 0.1)
    assert rel(xt, xr) < 1e-10


@pytest.mark.parametrize("kind", ["measure", "L2"])
def test_registration_decreases_and_ends_on_oracle(mods, kind):
    rng, x0, w0, y, wy = problem(seed=5)
    m = make(mods, kind, 'float64', x0, w0, y, wy)
    m.optimizeMatching()
    h = np.asarray(m.bfgs_history)
    print(f"[second order] {kind}: objective {h[0]:.4f} -> {h[-1]:.4f} in {len(h) - 1} iterations", flush=True)
    assert np.all(np.diff(h) <= 1e-12 * abs(h[0]))
    assert h[-1] < 0.5 * h[0]
    a0 = m.control['a0'].cpu().numpy()
    ref, _ = oracle_pair(kind, x0, w0, a0, y, wy)
    m.obj = None
    assert abs(m.objectiveFun() - ref) / abs(ref) < 1e-10
