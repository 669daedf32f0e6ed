"""GPU parity of the kernel's affine part in first-order and geodesic shooting.

Kernel(affine='affine'|'euclidean') adds w1 (y - c)(x - c)^T a + w2 sum(a)
(or its rotation part) to every K(x_t) a_t of the flow, its energy, the
lddmm covector and the Euclidean gradient.  The device loops
(lddmm_shoot_forward / lddmm_shoot_adjoint) are compared with the oracle's
restatement of Kernel.applyK / applyDiffKT (base/kernelFunctions.py:476-505,
:601-625): FP64 <= 1e-11 relative, FP32 compensated <= 2e-5 (the affine part
itself is computed in float64).  SurfaceMatching with affineKernel=True must
reproduce the oracle objective and gradient.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {'float64': 1e-11, 'float32': 2e-5}


@pytest.fixture(scope="module")
def env():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lddmm_amd import pointEvolution as pe, synth
    from lddmm_amd.kernelFunctions import Kernel
    from lddmm_amd.surfaces import Surface
    from lddmm_amd.surfaceMatching import SurfaceMatching
    return dict(torch=torch, pe=pe, synth=synth, Kernel=Kernel, Surface=Surface, SurfaceMatching=SurfaceMatching)


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("kind", ["affine", "euclidean"])
@pytest.mark.parametrize("n,dim,T", [(1, 3, 2), (700, 3, 5), (3000, 3, 4), (500, 2, 4)])
def test_shooting_with_affine_kernel(env, dtype, kind, n, dim, T):
    pe, Kernel = env["pe"], env["Kernel"]
    rng = np.random.default_rng(n + 7 * T + (kind == 'affine'))
    x0 = rng.normal(size=(n, dim)) * 3
    at = rng.normal(size=(T, n, dim)) * 0.2 / np.sqrt(n)
    px1 = rng.normal(size=(n, dim))
    c = rng.normal(size=dim)
    K = Kernel(name='gauss', sigma=[1.5], dim=dim)
    K.setAffine(kind, w1=0.3, w2=0.8, center=c)
    K.pk_dtype = dtype
    kaff = (kind, 0.3, 0.8, c)
    xr = O.landmarkDirectEvolutionEuler(x0, at, [1.5], kaff=kaff)
    xt, e = pe.shoot(x0, at, K, regweight=0.7, energy=True)
    assert rel(xt.cpu().numpy(), xr) < TOL[dtype]
    er = sum(0.7 / T * (at[t] * (O.applyK(xr[t], xr[t], at[t], [1.5]) + O.affineApplyK(xr[t], xr[t], at[t], kaff))
                        ).sum() for t in range(T))
    assert abs(e.item() - er) <= TOL[dtype] * abs(er) + 1e-14
    dat, _ = pe.landmarkHamiltonianGradient(x0, at, px1, K, 0.7)
    dr, _, pr = O.landmarkHamiltonianGradient(x0, at, px1, [1.5], 0.7, kaff=kaff)
    assert rel(dat, dr) < TOL[dtype]
    g, pxt = pe.shoot_adjoint(xt, at, px1, K, 0.7, euclidean=True, covector=True)
    assert rel(pxt.cpu().numpy(), pr) < TOL[dtype]
    gr = np.stack([O.applyK(xr[t], xr[t], dr[t], [1.5]) + O.affineApplyK(xr[t], xr[t], dr[t], kaff)
                   for t in range(T)]) / T
    assert rel(g.cpu().numpy(), gr) < 5 * TOL[dtype]


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("kind", ["affine", "euclidean"])
@pytest.mark.parametrize("n,dim,T,sig", [(1, 3, 2, [1.5]), (700, 3, 5, [1.5]), (3000, 3, 4, [1.5]),
                                         (500, 2, 4, [1.5]), (400, 3, 3, [0.8, 2.0])])
def test_geodesic_shooting_with_affine_kernel(env, dtype, kind, n, dim, T, sig):
    """Geodesic (EPDiff) forward, covectors and da0 with the kernel's affine
    part (lddmm_epdiff_forward / _adjoint) against the oracle's
    secondOrderEvolution / secondOrderCovector / secondOrderGradient, whose
    affine terms (Kernel.applyK, applyDiffKT, applyDDiffK11and12,
    applyDiffK1and2) are pinned by finite differences (tests/test_oracle.py)."""
    pe, Kernel = env["pe"], env["Kernel"]
    rng = np.random.default_rng(3 * n + T + (kind == 'affine'))
    x0 = rng.normal(size=(n, dim)) * 3
    a0 = rng.normal(size=(n, dim)) * 0.3 / np.sqrt(n)
    px1, pa1 = rng.normal(size=(n, dim)), rng.normal(size=(n, dim)) * 0.1
    c = rng.normal(size=dim)
    K = Kernel(name='gauss', sigma=sig, dim=dim)
    K.setAffine(kind, w1=0.3, w2=0.8, center=c)
    K.pk_dtype = dtype
    kaff = (kind, 0.3, 0.8, c)
    fw = pe.secondOrderEvolution(x0, a0, K, 1 / T)
    xr, ar = O.secondOrderEvolution(x0, a0, sig, 1 / T, kaff=kaff)
    assert rel(fw['xt'], xr) < TOL[dtype]
    assert rel(fw['at'], ar) < TOL[dtype]
    res = pe.secondOrderGradient(x0, a0, px1, pa1, K, 1 / T)
    dr, _, _, pxr, par = O.secondOrderGradient(x0, a0, px1, pa1, sig, 1 / T, kaff=kaff)
    assert rel(res['pxt'], pxr) < TOL[dtype]
    assert rel(res['pat'], par) < TOL[dtype]
    assert rel(res['da0'], dr) < TOL[dtype]
    # the affine part matters at this size
    xr0, _ = O.secondOrderEvolution(x0, a0, sig, 1 / T)
    assert rel(xr, xr0) > 100 * TOL[dtype]


@pytest.mark.parametrize("affine", ["euclidean", "affine"])
def test_second_order_surface_matching_affine_kernel(env, affine):
    """SecondOrderSurfaceMatching with affineKernel=True: objective and
    gradient against the oracle's second-order objective with the kernel's
    affine part; a short registration decreases the objective."""
    torch, S, Kernel = env["torch"], env["Surface"], env["Kernel"]
    from lddmm_amd.secondOrderSurfaceMatching import SecondOrderSurfaceMatching
    (v, f), (v2, _) = env["synth"].plain_matching_pair(seed=9, subdiv=3, radius=10.0, nbumps=20, width=3.0)
    opts = {'KparDiff': Kernel(name='gauss', sigma=2.5), 'KparDist': Kernel(name='gauss', sigma=1.0),
            'errorType': 'varifold', 'sigmaError': 1.0, 'timeStep': 0.1, 'pk_dtype': 'float64',
            'mode': 'quiet', 'affineKernel': True, 'affine': affine, 'maxIter': 10}
    sm = SecondOrderSurfaceMatching(Template=S((f, v)), Target=S((f, v2)), options=opts)
    assert sm.options['KparDiff'].affine == affine
    w1 = 1 / (0.01 if affine == 'euclidean' else 10.0)
    kaff = (affine, w1, 1 / 0.1, v.mean(axis=0))
    # small start: with the default rotWeight 0.01 (w1 = 100) a momentum of
    # 1e-3 per vertex already overflows the geodesic within 10 steps (in the
    # oracle as on the device)
    a0 = np.random.default_rng(12).normal(size=v.shape) * 1e-6
    sm.setInitialControl(a0)
    ref, _ = O.second_order_matching_objective(v, f, a0, (v2, f), [2.5], [1.0], 0.1, kind='varifold', kaff=kaff)
    assert abs(sm.objectiveFun() - ref) / abs(ref) < 1e-10
    g = sm.getGradient()['a0'].cpu().numpy()
    gr = O.second_order_matching_gradient(v, f, a0, (v2, f), [2.5], [1.0], 0.1, kind='varifold', kaff=kaff)
    assert rel(g, gr) < 1e-10
    sm.optimizeMatching()
    h = sm.bfgs_history
    assert len(h) >= 3 and h[-1] < h[0]


@pytest.mark.parametrize("affine", ["euclidean", "affine"])
def test_surface_matching_affine_kernel(env, affine):
    torch, S, Kernel = env["torch"], env["Surface"], env["Kernel"]
    (v, f), (v2, _) = env["synth"].plain_matching_pair(seed=7, subdiv=3, radius=10.0, nbumps=20, width=3.0)
    opts = {'KparDiff': Kernel(name='gauss', sigma=2.5), 'KparDist': Kernel(name='gauss', sigma=1.0),
            'errorType': 'varifold', 'sigmaError': 1.0, 'timeStep': 0.1, 'pk_dtype': 'float64',
            'regWeight': 1.0, 'mode': 'quiet', 'affineKernel': True, 'affine': affine}
    sm = env["SurfaceMatching"](Template=S((f, v)), Target=S((f, v2)), options=opts)
    assert sm.options['affine'] == 'none' and sm.options['KparDiff'].affine == affine
    # base/pointSetMatching.py:193-205 weights: w1 = 1/rotWeight (euclidean) or
    # 1/affineWeight, w2 = 1/transWeight, centred on the template
    w1 = 1 / (0.01 if affine == 'euclidean' else 10.0)
    kaff = (affine, w1, 1 / 0.1, v.mean(axis=0))
    at = np.random.default_rng(11).normal(size=(10, v.shape[0], 3)) * 1e-6
    sm.control['at'] = torch.tensor(at, device='cuda')
    sm.obj = None
    ref, _ = O.matching_objective(v, f, at, (v2, f), [2.5], [1.0], kaff=kaff)
    assert abs(sm.objectiveFun() - ref) / abs(ref) < 1e-10
    g = sm.getGradient()['at'].cpu().numpy()
    assert rel(g, O.matching_gradient(v, f, at, (v2, f), [2.5], [1.0], kaff=kaff)) < 1e-10
