"""GPU parity of the kernel divergence and the log-Jacobian it drives.

applyDiv (lddmm_applyDiv) against the FP64 oracle, which is pinned as the trace
of the finite-difference Jacobian of applyK: FP64 <= 1e-12 relative, FP32
compensated <= 2e-5 (the divergence is a signed sum, like applyDiffKT's).
Kernel.applyDivergence adds the affine part; withJacobian in the shooting
functions (J_{t+1} = J_t + dt div v_t(x_t), base/pointEvolution.py:149-152,
:366-367, :459-463) is compared with the oracle's trajectory and Jacobian.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {'float64': 1e-12, 'float32': 2e-5}


@pytest.fixture(scope="module")
def env():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lddmm_amd import kernelFunctions_util as ku, pointEvolution as pe
    from lddmm_amd.kernelFunctions import Kernel
    return dict(torch=torch, ku=ku, pe=pe, Kernel=Kernel)


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


SHAPES = [  # n, m, dim, sigma
    (1, 1, 3, [1.0]),
    (37, 53, 3, [2.0]),
    (513, 257, 2, [1.0]),
    (1000, 1000, 3, [1.0, 2.5]),
    (2000, 3000, 3, [2.5]),
    (700, 900, 3, [0.7, 1.4, 3.0]),
]


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("n,m,dim,sig", SHAPES)
def test_applyDiv_gauss(env, dtype, n, m, dim, sig):
    rng = np.random.default_rng(3 * n + m)
    y, x, a = rng.normal(size=(n, dim)) * 3, rng.normal(size=(m, dim)) * 3, rng.normal(size=(m, dim))
    got = env["ku"].applyDiv(y, x, a, 'gauss', sig, 10, dtype=dtype)
    ref = O.applyDiv(y, x, a, sig)
    assert got.shape == ref.shape == (n, 1)
    if n == 1 and m == 1:
        return      # one point: the divergence at y != x is a single term
    assert rel(got, ref) < TOL[dtype]


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("order", [0, 1, 2, 3, 4])
def test_applyDiv_laplacian(env, dtype, order):
    rng = np.random.default_rng(40 + order)
    y, x, a = rng.normal(size=(600, 3)) * 2, rng.normal(size=(800, 3)) * 2, rng.normal(size=(800, 3))
    sig = [1.3, 2.2]
    got = env["ku"].applyDiv(y, x, a, 'laplacian', sig, order, dtype=dtype)
    ref = O.applyDiv(y, x, a, sig, kernel=('laplacian', order))
    if order == 0:
        assert np.all(got == 0)
    else:
        assert rel(got, ref) < TOL[dtype]


def test_applyDiv_edge_cases(env):
    ku = env["ku"]
    y = np.ones((5, 3))
    out = ku.applyDiv(y, np.zeros((0, 3)), np.zeros((0, 3)), 'gauss', [1.], 10)
    np.testing.assert_array_equal(out, np.zeros((5, 1)))
    assert ku.applyDiv(np.zeros((0, 3)), y, y, 'gauss', [1.], 10).shape == (0, 1)
    # a single point: y_i - x_i = 0, so its own term vanishes
    x = np.array([[0.3, -1.0, 2.0]])
    assert ku.applyDiv(x, x, np.ones((1, 3)), 'gauss', [1.], 10)[0, 0] == 0.0
    with pytest.raises(ValueError):
        ku.applyDiv(y, y, np.ones((5, 2)), 'gauss', [1.], 10)


def test_applyDiv_torch_inputs_stay_on_device(env):
    torch = env["torch"]
    rng = np.random.default_rng(9)
    y, x, a = rng.normal(size=(300, 3)), rng.normal(size=(400, 3)), rng.normal(size=(400, 3))
    out = env["ku"].applyDiv(*[torch.tensor(v, device="cuda") for v in (y, x, a)], 'gauss', [1.5], 10)
    assert isinstance(out, torch.Tensor) and out.is_cuda and tuple(out.shape) == (300, 1)
    assert rel(out.cpu().numpy(), O.applyDiv(y, x, a, [1.5])) < TOL['float64']


def test_applyDiv_full_size_rows(env):
    """50k x 50k in both precisions, 256 target rows checked against the oracle."""
    torch = env["torch"]
    rng = np.random.default_rng(21)
    n = 50_000
    x = rng.normal(size=(n, 3)) * 10
    a = rng.normal(size=(n, 3))
    rows = rng.choice(n, 256, replace=False)
    ref = O.applyDiv(x[rows], x, a, [2.5])
    xd, ad = torch.tensor(x, device="cuda"), torch.tensor(a, device="cuda")
    for dtype in ("float64", "float32"):
        got = env["ku"].applyDiv(xd, xd, ad, 'gauss', [2.5], 10, dtype=dtype).cpu().numpy()[rows]
        assert rel(got, ref) < TOL[dtype], dtype


@pytest.mark.parametrize("kind", [None, "affine", "euclidean"])
def test_kernel_applyDivergence(env, kind):
    Kernel = env["Kernel"]
    rng = np.random.default_rng(17)
    x, a, y = rng.normal(size=(400, 3)) * 2, rng.normal(size=(400, 3)), rng.normal(size=(300, 3)) * 2
    K = Kernel(name='gauss', sigma=[1.7])
    c = np.array([0.2, -0.4, 0.1])
    if kind:
        K.setAffine(kind, w1=0.6, w2=0.9, center=c)
    kaff = (kind, 0.6, 0.9, c) if kind else None
    ref = O.applyDiv(x, x, a, [1.7]) + (O.affineApplyDiv(x, a, kaff) if kaff else 0)
    assert rel(K.applyDivergence(x, a), ref) < TOL['float64']
    ref = O.applyDiv(y, x, a, [1.7]) + (O.affineApplyDiv(x, a, kaff) if kaff else 0)
    assert rel(K.applyDivergence(x, a, firstVar=y), ref) < TOL['float64']


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("kind", [None, "affine"])
def test_direct_evolution_with_jacobian(env, dtype, kind):
    pe, Kernel = env["pe"], env["Kernel"]
    rng = np.random.default_rng(5)
    n, T = 800, 6
    x0 = rng.normal(size=(n, 3)) * 3
    at = rng.normal(size=(T, n, 3)) * 0.3 / np.sqrt(n)
    K = Kernel(name='gauss', sigma=[1.5])
    c = np.array([0.5, 0.0, -0.2])
    if kind:
        K.setAffine(kind, w1=0.3, w2=0.8, center=c)
    K.pk_dtype = dtype
    kaff = (kind, 0.3, 0.8, c) if kind else None
    out = pe.landmarkDirectEvolutionEuler(x0, at, K, options={'withJacobian': True})
    xr = O.landmarkDirectEvolutionEuler(x0, at, [1.5], kaff=kaff)
    Jr = O.jacobian(xr, at, [1.5], kaff=kaff)
    assert out['Jt'].shape == (T + 1, n, 1)
    assert np.all(out['Jt'][0] == 0)
    assert rel(out['xt'], xr) < TOL[dtype]
    assert rel(out['Jt'], Jr) < TOL[dtype]
    assert pe.landmarkDirectEvolutionEuler(x0, at, K)['Jt'] is None


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_geodesic_shooting_with_jacobian(env, dtype):
    pe, Kernel = env["pe"], env["Kernel"]
    rng = np.random.default_rng(6)
    n = 600
    x0 = rng.normal(size=(n, 3)) * 3
    a0 = rng.normal(size=(n, 3)) * 0.05
    K = Kernel(name='gauss', sigma=[2.0])
    K.pk_dtype = dtype
    out = pe.secondOrderEvolution(x0, a0, K, 0.1, options={'withJacobian': True})
    xr, ar = O.secondOrderEvolution(x0, a0, [2.0], 0.1)
    Jr = O.jacobian(xr, ar[:10], [2.0])
    assert out['Jt'].shape == (11, n, 1)
    assert rel(out['Jt'], Jr) < TOL[dtype]
    xt, at, Jt = pe.landmarkEPDiff(10, x0, a0, K, withJacobian=True)
    assert at.shape == (10, n, 3)
    assert rel(Jt, Jr) < TOL[dtype]


def test_matching_objective_with_jacobian(env):
    """objectiveFunDef(withJacobian=True) carries Jt in the state
    (base/surfaceMatching.py:468-565 with solveStateEquation's options)."""
    torch = env["torch"]
    from lddmm_amd import synth
    from lddmm_amd.surfaces import Surface
    from lddmm_amd.surfaceMatching import SurfaceMatching
    from lddmm_amd.pointSets import PointSet
    from lddmm_amd.secondOrderPointSetMatching import SecondOrderPointSetMatching
    Kernel = env["Kernel"]
    v, f = synth.icosphere(2, 5.0)
    rng = np.random.default_rng(2)
    sm = SurfaceMatching(Template=Surface((f, v)), Target=Surface((f, v * 1.1)),
                         options={'KparDiff': Kernel('gauss', sigma=2.5), 'KparDist': Kernel('gauss', sigma=1.0),
                                  'errorType': 'varifold', 'mode': 'quiet', 'timeStep': 0.2,
                                  'pk_dtype': 'float64'})
    at = rng.normal(size=(5, v.shape[0], 3)) * 0.05
    ctl = {'at': torch.tensor(at, device='cuda')}
    obj, st = sm.objectiveFunDef(ctl, withTrajectory=False, withJacobian=True)
    xr = O.landmarkDirectEvolutionEuler(v, at, [2.5])
    r = rel(st['Jt'].cpu().numpy(), O.jacobian(xr, at, [2.5]))
    assert r < TOL['float64'], r

    a0 = rng.normal(size=(v.shape[0], 3)) * 0.05
    m = SecondOrderPointSetMatching(Template=PointSet(v), Target=PointSet(v * 1.1),
                                    options={'KparDiff': Kernel('gauss', sigma=2.0),
                                             'KparDist': Kernel('gauss', sigma=1.0), 'errorType': 'measure',
                                             'timeStep': 0.1, 'mode': 'quiet', 'pk_dtype': 'float64'})
    _, st = m.objectiveFunDef({'a0': torch.tensor(a0, device='cuda')}, withJacobian=True)
    xr, ar = O.secondOrderEvolution(v, a0, [2.0], 0.1)
    r = rel(st['Jt'].cpu().numpy(), O.jacobian(xr, ar[:10], [2.0]))
    assert r < TOL['float64'], r


def _vtk_field(path, kind, name, n):
    lines = open(path).read().splitlines()
    if kind == "SCALARS":
        i = lines.index(f"SCALARS {name} float 1") + 2
        return np.array([float(t) for t in lines[i:i + n]])
    i = lines.index(f"VECTORS {name} float") + 1
    return np.array([[float(u) for u in t.split()] for t in lines[i:i + n]])


def test_registration_writes_the_evolution_with_its_jacobian(env, tmp_path):
    """writeOutput: every saveRate iterations and at the end, one legacy VTK
    file per time step (base/surfaceMatching.py:1219-1290, :1370-1432); the
    last files hold exp(J_t), the vertex-area ratio and the velocity of the
    final control, and the point trajectories (saveTrajectories); the result
    file holds the datasets of the reference's saveHdf5 under its paths."""
    from lddmm_amd import synth
    from lddmm_amd.surfaces import Surface
    from lddmm_amd.surfaceMatching import SurfaceMatching
    Kernel = env["Kernel"]
    v, f = synth.icosphere(2, 5.0)
    sm = SurfaceMatching(Template=Surface((f, v)), Target=Surface((f, v * 1.15)),
                         options={'KparDiff': Kernel('gauss', sigma=2.5), 'KparDist': Kernel('gauss', sigma=1.0),
                                  'errorType': 'varifold', 'mode': 'quiet', 'timeStep': 0.2,
                                  'pk_dtype': 'float64', 'maxIter': 4, 'saveRate': 2, 'writeOutput': True,
                                  'outputDir': str(tmp_path / "out"), 'saveTrajectories': True})
    sm.optimizeMatching()
    names = sorted(p.name for p in (tmp_path / "out").iterdir())
    assert names == ["evolution%03d.vtk" % k for k in range(6)] + ["evolutioncurves.vtk", "output.h5.npz"]
    at = sm.control['at'].cpu().numpy()
    xr = O.landmarkDirectEvolutionEuler(v, at, [2.5])
    Jr = O.jacobian(xr, at, [2.5])
    nv = v.shape[0]
    for k in (0, 3, 5):
        path = str(tmp_path / "out" / ("evolution%03d.vtk" % k))
        J = _vtk_field(path, "SCALARS", "Jacobian", nv)
        assert rel(J, np.exp(Jr[k, :, 0])) < 1e-12
        JT = _vtk_field(path, "SCALARS", "Jacobian_T", nv)
        S = Surface((f, xr[k]))
        assert rel(JT, S.computeVertexArea()[0] / Surface((f, v)).computeVertexArea()[0]) < 1e-10
        assert rel(_vtk_field(path, "SCALARS", "Jacobian_N", nv), J / JT) < 1e-12
        vel = _vtk_field(path, "VECTORS", "velocity", nv)
        kk = min(k, 4)
        assert rel(vel, O.applyK(xr[kk], xr[kk], at[kk], [2.5])) < 1e-11
    assert np.all(_vtk_field(str(tmp_path / "out" / "evolution000.vtk"), "SCALARS", "displacement", nv) == 0)
    # the template grows toward the larger target: the mean Jacobian exceeds 1
    assert np.mean(_vtk_field(str(tmp_path / "out" / "evolution005.vtk"), "SCALARS", "Jacobian", nv)) > 1.0
    # the result file (base/surfaceMatching.py:1506-1555 layout, npz archive)
    res = np.load(str(tmp_path / "out" / "output.h5.npz"))
    r = 'LDDMM Results/'
    assert int(res[r + 'parameters/Time steps']) == 5
    assert str(res[r + 'parameters/Deformation Kernel type']) == 'gauss'
    assert float(res[r + 'parameters/Deformation Kernel width'][0]) == 2.5
    assert float(res[r + 'parameters/Spatial Varifold width'][0]) == 1.0
    np.testing.assert_array_equal(res[r + 'template/vertices'], v)
    np.testing.assert_array_equal(res[r + 'template/faces'], f)
    np.testing.assert_array_equal(res[r + 'target/vertices'], v * 1.15)
    np.testing.assert_array_equal(res[r + 'variables/alpha'], at)
    assert str(res[r + 'variables/affine']) == 'None'
    assert rel(res[r + 'deformedTemplate/vertices'], xr[-1]) < 1e-12
    assert rel(res[r + 'descriptors/Jacobian'], Jr[-1]) < 1e-12
    assert rel(res[r + 'descriptors/Surface Jacobian'],
               Surface((f, xr[-1])).computeVertexArea()[0] / Surface((f, v)).computeVertexArea()[0]) < 1e-10
    assert rel(res[r + 'descriptors/Displacement'], xr[-1] - xr[0]) < 1e-12
