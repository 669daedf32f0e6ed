"""GPU parity of the measure data terms (surface, surface vs point set, point
sets) and of SurfaceMatching with errorType 'measure', against the oracle.
FP64 <= 1e-11 relative, FP32 compensated: energies <= 1e-5, gradients <= 2e-5."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lddmm_amd import surfaceDistances as sd, pointset_distances as psd, synth
    from lddmm_amd.surfaces import Surface
    from lddmm_amd.pointSets import PointSet
    from lddmm_amd.kernelFunctions import Kernel
    return sd, psd, synth, Surface, PointSet, Kernel, torch


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


TOL_E = {'float64': 1e-11, 'float32': 1e-5}
TOL_G = {'float64': 1e-11, 'float32': 2e-5}


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("subdiv,sig,weighted", [(2, [1.0], False), (3, [0.7], True), (4, [1.0, 2.5], False)])
def test_surface_measure(mods, dtype, subdiv, sig, weighted):
    sd, psd, synth, Surface, PointSet, Kernel, torch = mods
    (v, f), (v2, _) = synth.plain_matching_pair(seed=4, subdiv=subdiv, nbumps=10)
    A, Bt = Surface((f, v)), Surface((f, v2))
    a, b = (v, f), (v2, f)
    if weighted:
        w = np.random.default_rng(1).uniform(0.5, 1.5, size=v.shape[0])
        A.updateWeights(w)
        a = (v, f, w)
    K = Kernel(name='gauss', sigma=sig)
    K.pk_dtype = dtype
    r0, rD = O.measureNorm0(b, sig), O.measureNormDef(a, b, sig)
    assert abs(sd.measureNorm0(Bt, K) - r0) / abs(r0) < TOL_E[dtype]
    assert abs(sd.measureNormDef(A, Bt, K) - rD) / abs(rD) < TOL_E[dtype]
    assert rel(sd.measureNormGradient(A, Bt, K), O.measureNormGradient(a, b, sig)) < TOL_G[dtype]


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_point_set_measures(mods, dtype):
    sd, psd, synth, Surface, PointSet, Kernel, torch = mods
    (v, f), (v2, _) = synth.plain_matching_pair(seed=6, subdiv=3, nbumps=10)
    rng = np.random.default_rng(2)
    pts = v2 + 0.1 * rng.normal(size=v2.shape)
    wts = rng.uniform(0.2, 1.0, size=v2.shape[0])
    PS = PointSet(pts, weights=wts)
    A = Surface((f, v))
    K = Kernel(name='gauss', sigma=1.2)
    K.pk_dtype = dtype
    assert abs(sd.measureNormPS0(PS, K) - O.measureNormPS0(pts, wts, [1.2])) \
        / O.measureNormPS0(pts, wts, [1.2]) < TOL_E[dtype]
    r = O.measureNormPSDef((v, f), pts, wts, [1.2])
    assert abs(sd.measureNormPSDef(A, PS, K) - r) / abs(r) < TOL_E[dtype]
    assert rel(sd.measureNormPSGradient(A, PS, K), O.measureNormPSGradient((v, f), pts, wts, [1.2])) < TOL_G[dtype]
    X = PointSet(v, weights=rng.uniform(0.5, 1.5, size=v.shape[0]))
    r0 = O.pointMeasureNorm0(pts, wts, [1.2])
    assert abs(psd.measureNorm0(PS, K) - r0) / r0 < TOL_E[dtype]
    rD = O.pointMeasureNormDef(X.vertices, X.weights, pts, wts, [1.2])
    assert abs(psd.measureNormDef(X, PS, K) - rD) / abs(rD) < TOL_E[dtype]
    assert rel(psd.measureNormGradient(X, PS, K),
               O.pointMeasureNormGradient(X.vertices, X.weights, pts, wts, [1.2])) < TOL_G[dtype]


def test_measure_matching_objective_and_gradient(mods):
    sd, psd, synth, Surface, PointSet, Kernel, torch = mods
    from lddmm_amd.surfaceMatching import SurfaceMatching
    (v, f), (v2, _) = synth.plain_matching_pair(seed=7, subdiv=3, radius=10.0, nbumps=20, width=3.0)
    opts = {'KparDiff': Kernel(name='gauss', sigma=2.5), 'KparDist': Kernel(name='gauss', sigma=1.0),
            'errorType': 'measure', 'timeStep': 0.1, 'pk_dtype': 'float64', 'mode': 'quiet'}
    sm = SurfaceMatching(Template=Surface((f, v)), Target=Surface((f, v2)), options=opts)
    at = np.random.default_rng(11).normal(size=(10, v.shape[0], 3)) * 0.05
    sm.control['at'] = torch.tensor(at, device='cuda')
    sm.obj = None
    ref, _ = O.matching_objective(v, f, at, (v2, f), [2.5], [1.0], kind='measure')
    assert abs(sm.objectiveFun() - ref) / abs(ref) < 1e-10
    g = sm.getGradient()['at'].cpu().numpy()
    assert rel(g, O.matching_gradient(v, f, at, (v2, f), [2.5], [1.0], kind='measure')) < 1e-10
    sm2 = SurfaceMatching(Template=Surface((f, v)), Target=Surface((f, v2)),
                          options=dict(opts, maxIter=10, KparDiff=Kernel(name='gauss', sigma=2.5),
                                       KparDist=Kernel(name='gauss', sigma=1.0)))
    sm2.optimizeMatching()
    assert sm2.bfgs_history[-1] < 0.5 * sm2.bfgs_history[0]


def test_surface_device_view_after_vertex_override(mods):
    """A device view built with replaced vertices, then one with the surface's
    own vertices, uses the right coordinates each time."""
    sd, psd, synth, Surface, PointSet, Kernel, torch = mods
    (v, f), (v2, _) = synth.plain_matching_pair(seed=2, subdiv=2, nbumps=5)
    A, Bt = Surface((f, v)), Surface((f, v2))
    K = Kernel(name='gauss', sigma=1.0)
    e1 = float(sd.measureNormDef_device(A, Bt, K, vertices=torch.tensor(v2, device='cuda')).item())
    e2 = sd.measureNormDef(A, Bt, K)
    assert abs(e1 - O.measureNormDef((v2, f), (v2, f), [1.0])) < 1e-9 * abs(e1)
    assert abs(e2 - O.measureNormDef((v, f), (v2, f), [1.0])) < 1e-11 * abs(e2)


@pytest.mark.parametrize("kind", ["measure", "L2"])
def test_point_set_matching(mods, kind):
    """PointSetMatching objective / gradient at a non-zero control against the
    oracle (measure) or the closed-form L2 term; a short run decreases it."""
    sd, psd, synth, Surface, PointSet, Kernel, torch = mods
    from lddmm_amd.pointSetMatching import PointSetMatching
    rng = np.random.default_rng(8)
    x0 = synth.fibonacci_sphere(400, 5.0)
    w0 = rng.uniform(0.5, 1.5, size=400)
    y = x0 * 1.1 + 0.2 * rng.normal(size=x0.shape)
    wy = rng.uniform(0.5, 1.5, size=400)
    opts = {'KparDiff': Kernel(name='gauss', sigma=2.0), 'KparDist': Kernel(name='gauss', sigma=1.0),
            'errorType': kind, 'timeStep': 0.1, 'pk_dtype': 'float64', 'mode': 'quiet', 'maxIter': 10}
    pm = PointSetMatching(Template=PointSet(x0, weights=w0), Target=PointSet(y, weights=wy), options=opts)
    at = rng.normal(size=(10, 400, 3)) * 0.05
    pm.control['at'] = torch.tensor(at, device='cuda')
    pm.obj = None
    obj = pm.objectiveFun()
    if kind == 'measure':
        ref, _ = O.matching_objective(x0, w0, at, (y, wy), [2.0], [1.0], kind='point_measure')
        gr = O.matching_gradient(x0, w0, at, (y, wy), [2.0], [1.0], kind='point_measure')
    else:
        xt = O.landmarkDirectEvolutionEuler(x0, at, [2.0])
        ref = sum(0.1 * (at[t] * O.applyK(xt[t], xt[t], at[t], [2.0])).sum() for t in range(10)) \
            + ((xt[-1] - y) ** 2).sum()
        px1 = -2 * (xt[-1] - y)
        dat, xt2, _ = O.landmarkHamiltonianGradient(x0, at, px1, [2.0], 1.0)
        gr = np.stack([O.applyK(xt2[t], xt2[t], dat[t], [2.0]) for t in range(10)]) / 10
    assert abs(obj - ref) / abs(ref) < 1e-10
    assert rel(pm.getGradient()['at'].cpu().numpy(), gr) < 1e-10
    pm2 = PointSetMatching(Template=PointSet(x0, weights=w0), Target=PointSet(y, weights=wy),
                           options=dict(opts, KparDiff=Kernel(name='gauss', sigma=2.0),
                                        KparDist=Kernel(name='gauss', sigma=1.0)))
    pm2.optimizeMatching()
    assert pm2.bfgs_history[-1] < 0.7 * pm2.bfgs_history[0]
