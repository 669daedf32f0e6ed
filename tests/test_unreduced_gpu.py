"""GPU parity of the unreduced (control-point) surface matching.

SurfaceMatching(unreduced=True) carries control points 'ct' beside 'at'; the
semi-reduced flow, its objective and both gradients run on device kernel
sums and must match the oracle's restatement
(base/pointEvolutionSemiReduced.py:7-90, base/surfaceMatching.py unreduced
branches) to 1e-10 relative in FP64; a short BFGS run over both controls
(device two-loop over the concatenated controls) must decrease the objective.
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
    from lddmm_amd import synth
    from lddmm_amd.kernelFunctions import Kernel
    from lddmm_amd.surfaces import Surface
    from lddmm_amd.surfaceMatching import SurfaceMatching
    return dict(torch=torch, synth=synth, Kernel=Kernel, Surface=Surface, SurfaceMatching=SurfaceMatching)


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("variant", ["plain", "fidelity+costs+laplacian+affine"])
def test_unreduced_objective_and_gradients(env, variant):
    torch, S, Kernel = env["torch"], env["Surface"], env["Kernel"]
    (v, f), (v2, _) = env["synth"].plain_matching_pair(seed=7, subdiv=3, radius=10.0, nbumps=20, width=3.0)
    N = v.shape[0]
    opts = {'KparDiff': Kernel(name='gauss', sigma=2.5), 'KparDist': Kernel(name='gauss', sigma=1.0),
            'errorType': 'varifold', 'sigmaError': 1.0, 'timeStep': 0.1, 'pk_dtype': 'float64',
            'regWeight': 1.0, 'mode': 'quiet', 'unreduced': True, 'unreducedWeight': 2.0}
    kw = dict(regweight=1.0, unreduced_weight=2.0 * 1000.0 / N)
    if variant != "plain":
        opts.update({'KparDiff': Kernel(name='laplacian', sigma=2.5, order=3), 'fidelityWeight': 0.4,
                     'internalCost': [['elastic', 5.0], ['h1', 0.5]], 'affineKernel': True,
                     'affine': 'euclidean'})
        kw.update(fidelity_weight=0.4, icost=[('elastic', 5.0), ('h1', 0.5)], kernel_diff=('laplacian', 3),
                  kaff=('euclidean', 100.0, 10.0, v.mean(axis=0)))
    sm = env["SurfaceMatching"](Template=S((f, v)), Target=S((f, v2)), options=opts)
    rng = np.random.default_rng(11)
    T = 10
    scale = 0.02 if variant == "plain" else 1e-5
    at = rng.normal(size=(T, N, 3)) * scale
    ct = np.broadcast_to(v, (T, N, 3)) + rng.normal(size=(T, N, 3)) * 0.2
    sm.control['at'] = torch.tensor(at, device='cuda')
    sm.control['ct'] = torch.tensor(ct, device='cuda')
    sm.obj = None
    ref, _ = O.matching_objective_unreduced(v, f, ct, at, (v2, f), [2.5], [1.0], **kw)
    assert abs(sm.objectiveFun() - ref) / abs(ref) < 1e-10
    g = sm.getGradient()
    gc, ga = O.matching_gradient_unreduced(v, f, ct, at, (v2, f), [2.5], [1.0], **kw)
    assert rel(g['at'].cpu(), ga) < 1e-10
    assert rel(g['ct'].cpu(), gc) < 1e-10
    # Jacobian of the semi-reduced flow: the divergence at the control points
    # (base/pointEvolutionSemiReduced.py:78-81)
    _, st = sm.objectiveFunDef(sm.control, withJacobian=True)
    Jr = O.jacobian(ct, at, [2.5], kernel=kw.get('kernel_diff', O.GAUSS), kaff=kw.get('kaff'))
    assert st['Jt'].shape == (T + 1, N, 1)
    assert rel(st['Jt'].cpu(), Jr) < 1e-10


def test_unreduced_registration_decreases(env, tmp_path):
    S, Kernel = env["Surface"], env["Kernel"]
    (v, f), (v2, _) = env["synth"].plain_matching_pair(seed=1, subdiv=3, radius=10.0, nbumps=20, width=3.0)
    opts = {'KparDiff': Kernel(name='gauss', sigma=2.5), 'KparDist': Kernel(name='gauss', sigma=1.0),
            'errorType': 'varifold', 'pk_dtype': 'float64', 'maxIter': 8, 'mode': 'quiet', 'unreduced': True,
            'writeOutput': True, 'saveRate': 4, 'outputDir': str(tmp_path)}
    sm = env["SurfaceMatching"](Template=S((f, v)), Target=S((f, v2)), options=opts)
    sm.optimizeMatching()
    h = sm.bfgs_history
    assert len(h) >= 3
    assert all(h[i + 1] <= h[i] for i in range(len(h) - 1))
    assert h[-1] < 0.8 * h[0]
    assert float((sm.control['ct'] - sm.x0_dev).abs().max()) > 0      # control points moved
    # the evolution files of the final control: exp(J_t) from the divergence at
    # c_t, velocity K(x_t, c_t) a_t
    ct, at = sm.control['ct'].cpu().numpy(), sm.control['at'].cpu().numpy()
    Jr = O.jacobian(ct, at, [2.5])
    xr = O.semiReducedEvolution(v, ct, at, [2.5], 0.0, O.GAUSS, None)
    nv = v.shape[0]
    for k in (0, 7):
        lines = open(tmp_path / ("evolution%03d.vtk" % k)).read().splitlines()
        i = lines.index("SCALARS Jacobian float 1") + 2
        assert rel([float(t) for t in lines[i:i + nv]], np.exp(Jr[k, :, 0])) < 1e-12
        i = lines.index("VECTORS velocity float") + 1
        vel = [[float(u) for u in t.split()] for t in lines[i:i + nv]]
        assert rel(vel, O.applyK(xr[k], ct[k], at[k], [2.5])) < 1e-11


def test_control_point_reset(env):
    """unreducedResetRate (base/surfaceMatching.py:1462-1489): at a reset check
    with a new lowest objective and a control point further than 5% of the
    largest template displacement from the trajectory, c_t <- x_t (t < T)
    exactly, the optimiser is told to restart, and the objective recomputed
    at the moved control points is the oracle's; without a new lowest
    objective, off the reset period, or with the control points on the
    trajectory, nothing moves."""
    torch, S, Kernel = env["torch"], env["Surface"], env["Kernel"]
    (v, f), (v2, _) = env["synth"].plain_matching_pair(seed=7, subdiv=2, radius=10.0, nbumps=20, width=3.0)
    N, T = v.shape[0], 10
    opts = {'KparDiff': Kernel(name='gauss', sigma=2.5), 'KparDist': Kernel(name='gauss', sigma=1.0),
            'errorType': 'varifold', 'pk_dtype': 'float64', 'mode': 'quiet', 'unreduced': True,
            'unreducedWeight': 2.0, 'unreducedResetRate': 2}
    sm = env["SurfaceMatching"](Template=S((f, v)), Target=S((f, v2)), options=opts)
    rng = np.random.default_rng(5)
    at = rng.normal(size=(T, N, 3)) * 0.05
    ct = np.broadcast_to(v, (T, N, 3)) + rng.normal(size=(T, N, 3)) * 0.3
    sm.control['at'] = torch.tensor(at, device='cuda')
    sm.control['ct'] = torch.tensor(ct, device='cuda')
    sm.obj = None
    obj = sm.objectiveFun()
    xt = sm.state['xt'].clone()
    sm.iter = 0
    sm.endOfIteration()                      # iteration 1: off the reset period
    assert not sm.reset and sm.ct_resets == []
    assert torch.equal(sm.control['ct'], torch.tensor(ct, device='cuda'))
    sm.endOfIteration()                      # iteration 2: reset
    assert sm.reset and sm.ct_resets == [2] and sm.obj_unreduced_save == obj
    assert torch.equal(sm.control['ct'], xt[:-1])
    assert torch.equal(sm.controlTry['ct'], xt[:-1])
    assert torch.equal(sm.control['at'], torch.tensor(at, device='cuda'))
    sm.obj = None
    kw = dict(regweight=1.0, unreduced_weight=2.0 * 1000.0 / N)
    ref, _ = O.matching_objective_unreduced(v, f, xt[:-1].cpu().numpy(), at, (v2, f), [2.5], [1.0], **kw)
    obj_new = sm.objectiveFun()
    assert abs(obj_new - ref) / abs(ref) < 1e-10
    # not a new lowest objective at the next check: no reset
    sm.reset = False
    sm.obj = obj + 1.0
    sm.iter = 3
    sm.endOfIteration()
    assert not sm.reset and sm.ct_resets == [2]
    # a new lowest objective with the control points on the trajectory: the
    # distance (0) is below 5% of the displacement
    sm.control['ct'] = sm.state['xt'][:-1].clone()
    sm.obj = 0.0
    sm.iter = 5
    sm.endOfIteration()
    d = float(((sm.control['ct'] - sm.state['xt'][:-1]) ** 2).sum(-1).max().sqrt())
    d0 = float(((sm.state['xt'][0] - sm.state['xt'][-1]) ** 2).sum(-1).max().sqrt())
    assert d <= 0.05 * d0 and not sm.reset and sm.ct_resets == [2]


def test_registration_with_control_point_resets(env):
    """A registration with resets every 3 iterations: at least one reset
    happens, the objective ends well below its start (a reset moves the
    control points, so the objective may rise at a restart), and the final
    objective is the oracle's at the final controls."""
    S, Kernel = env["Surface"], env["Kernel"]
    (v, f), (v2, _) = env["synth"].plain_matching_pair(seed=1, subdiv=3, radius=10.0, nbumps=20, width=3.0)
    N = v.shape[0]
    opts = {'KparDiff': Kernel(name='gauss', sigma=2.5), 'KparDist': Kernel(name='gauss', sigma=1.0),
            'errorType': 'varifold', 'pk_dtype': 'float64', 'maxIter': 12, 'mode': 'quiet', 'unreduced': True,
            'unreducedResetRate': 3}
    sm = env["SurfaceMatching"](Template=S((f, v)), Target=S((f, v2)), options=opts)
    sm.optimizeMatching()
    assert len(sm.ct_resets) >= 1
    h = sm.bfgs_history
    assert h[-1] < 0.8 * h[0]
    ct, at = sm.control['ct'].cpu().numpy(), sm.control['at'].cpu().numpy()
    ref, _ = O.matching_objective_unreduced(v, f, ct, at, (v2, f), [2.5], [1.0], regweight=1.0,
                                            unreduced_weight=1000.0 / N)
    sm.obj = None
    assert abs(sm.objectiveFun() - ref) / abs(ref) < 1e-10
