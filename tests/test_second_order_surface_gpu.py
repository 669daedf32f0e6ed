"""SecondOrderSurfaceMatching (geodesic shooting from the initial momentum a0
with the surface varifold / current / measure data terms) on the GPU against
the float64 oracle (second_order_matching_objective / _gradient with surface
data terms, pinned by finite differences in a0): objective and gradient at a
non-zero a0 in float64 (<= 1e-10) and compensated float32 (<= 2e-5), the
initialMomentum option, the internalCost warning, and registrations that
lower the objective at every accepted step and end where the oracle agrees.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SIG_DIFF, SIG_DIST, DT, SIG_ERR = 2.0, 1.0, 0.1, 0.8


@pytest.fixture(scope="module")
def mods():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lddmm_amd import synth
    from lddmm_amd.surfaces import Surface
    from lddmm_amd.kernelFunctions import Kernel
    from lddmm_amd.secondOrderSurfaceMatching import SecondOrderSurfaceMatching
    return synth, Surface, Kernel, SecondOrderSurfaceMatching, torch


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def problem(seed=8, subdiv=3):
    from lddmm_amd import synth
    v, f = synth.icosphere(subdiv, radius=6.0)
    v2 = synth.bump_deform(v, f, 10, 1.8, 0.6, seed=seed)
    return np.random.default_rng(seed), v, f, v2


def make(mods, kind, dtype, v, f, v2, **extra):
    synth, Surface, Kernel, SOSM, torch = mods
    opts = {'KparDiff': Kernel(name='gauss', sigma=SIG_DIFF), 'KparDist': Kernel(name='gauss', sigma=SIG_DIST),
            'errorType': kind, 'timeStep': DT, 'pk_dtype': dtype, 'mode': 'quiet', 'maxIter': 20,
            'sigmaError': SIG_ERR}
    opts.update(extra)
    return SOSM(Template=Surface((f, v)), Target=Surface((f, v2)), options=opts)


def oracle_pair(kind, v, f, a0, v2):
    kw = dict(kind=kind, sigma_error=SIG_ERR)
    obj, _ = O.second_order_matching_objective(v, f, a0, (v2, f), [SIG_DIFF], [SIG_DIST], DT, **kw)
    g = O.second_order_matching_gradient(v, f, a0, (v2, f), [SIG_DIFF], [SIG_DIST], DT, **kw)
    return obj, g


@pytest.mark.parametrize("dtype,tol", [("float64", 1e-10), ("float32", 2e-5)])
@pytest.mark.parametrize("kind", ["varifold", "current", "measure"])
def test_objective_and_gradient_match_oracle(mods, kind, dtype, tol):
    torch = mods[-1]
    rng, v, f, v2 = problem()
    a0 = 0.05 * rng.normal(size=v.shape)
    m = make(mods, kind, dtype, v, f, v2)
    m.control['a0'] = torch.tensor(a0, device='cuda')
    m.obj = None
    obj = m.objectiveFun()
    g = m.getGradient()['a0'].cpu().numpy()
    ref, gr = oracle_pair(kind, v, f, a0, v2)
    print(f"[second order surface] {kind} {dtype}: objective rel err {abs(obj - ref) / abs(ref):.2e}, "
          f"gradient rel err {rel(g, gr):.2e}", flush=True)
    assert abs(obj - ref) / abs(ref) < tol
    assert rel(g, gr) < tol


def test_initial_momentum_option_and_internal_cost_warning(mods, caplog):
    import logging
    rng, v, f, v2 = problem(seed=3)
    a0 = 0.05 * rng.normal(size=v.shape)
    with caplog.at_level(logging.INFO):
        m = make(mods, 'varifold', 'float64', v, f, v2, initialMomentum=a0, internalCost=[['elastic', 50.]])
    assert m.options['internalCost'] is None and m.icost is None
    assert any('Hybrid models not implemented' in r.message for r in caplog.records)
    ref, _ = oracle_pair('varifold', v, f, a0, v2)
    assert abs(m.objectiveFun() - ref) / abs(ref) < 1e-10
    xr, _ = O.secondOrderEvolution(v, a0, [SIG_DIFF], DT)
    assert rel(m.state['xt'].cpu().numpy(), xr) < 1e-10


@pytest.mark.parametrize("kind,dtype", [("varifold", "float64"), ("current", "float64"),
                                        ("varifold", "float32")])
def test_registration_decreases_and_ends_on_oracle(mods, kind, dtype):
    rng, v, f, v2 = problem(seed=5)
    m = make(mods, kind, dtype, v, f, v2)
    m.optimizeMatching()
    h = np.asarray(m.bfgs_history)
    print(f"[second order surface] {kind} {dtype}: objective {h[0]:.4f} -> {h[-1]:.4f} in "
          f"{m.bfgs_iterations} iterations", flush=True)
    assert np.all(np.diff(h) <= 1e-12 * abs(h[0]))
    assert h[-1] < 0.5 * h[0]
    a0 = m.control['a0'].cpu().numpy()
    ref, _ = oracle_pair(kind, v, f, a0, v2)
    m.resetPK('float64')
    m.obj = None
    assert abs(m.objectiveFun() - ref) / abs(ref) < 1e-10
    # the deformed surface is the geodesic's endpoint
    xr, _ = O.secondOrderEvolution(v, a0, [SIG_DIFF], DT)
    assert rel(m.fvDef.vertices, xr[-1]) < (1e-9 if dtype == 'float64' else 1e-5)
