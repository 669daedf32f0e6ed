"""GPU parity of the laplacian (Matern) kernel family against the FP64 oracle.

Every kernel sum, both shooting loops, a batched population shoot and the
SurfaceMatching objective/gradient with a laplacian deformation kernel are
compared with the oracle evaluated with kernel=('laplacian', order).  Bounds:
FP64 <= 1e-11 relative (norm-wise); FP32 compensated <= 1e-5 for the sums
and 2e-5 for trajectories and adjoints (5x that for the second-order
adjoint, as for the Gaussian).  Every call goes through the C ABI.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL_SUM = {'float64': 1e-11, 'float32': 1e-5}
TOL_SHOOT = {'float64': 1e-11, 'float32': 2e-5}


@pytest.fixture(scope="module")
def env():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lddmm_amd import kernelFunctions_util as ku, pointEvolution as pe, surfaceDistances as sd
    from lddmm_amd import synth, _backend as B
    from lddmm_amd.kernelFunctions import Kernel
    from lddmm_amd.surfaces import Surface
    from lddmm_amd.surfaceMatching import SurfaceMatching
    from lddmm_amd.population import Population
    return dict(torch=torch, ku=ku, pe=pe, sd=sd, synth=synth, B=B, Kernel=Kernel, Surface=Surface,
                SurfaceMatching=SurfaceMatching, Population=Population)


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    if nb == 0:
        return np.linalg.norm(a)
    return np.linalg.norm(a - b) / nb


def pts(rng, n, dim=3, scale=3.0, shift=50.0):
    return rng.normal(size=(n, dim)) * scale + shift


ORDERS = [0, 1, 2, 3, 4]
SIGS = [[1.5], [0.8, 2.5]]


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("sig", SIGS)
@pytest.mark.parametrize("order", ORDERS)
def test_applyK_and_diffKT(env, dtype, sig, order):
    ku = env["ku"]
    kern = ('laplacian', order)
    rng = np.random.default_rng(10 * order + len(sig))
    n, m = 700, 900
    y, x = pts(rng, n), pts(rng, m)
    a, p = rng.normal(size=(m, 3)), rng.normal(size=(n, 3))
    got = ku.applyK(y, x, a, 'laplacian', sig, order, dtype=dtype)
    assert rel(got, O.applyK(y, x, a, sig, kernel=kern)) < TOL_SUM[dtype]
    got = ku.applyDiffKT(y, x, p, a, 'laplacian', sig, order, dtype=dtype)
    assert rel(got, O.applyDiffKT(y, x, p, a, sig, kernel=kern)) < TOL_SUM[dtype]
    for kw in ({'sym': True}, {'lddmm': True, 'regweight': 0.6}):
        got = ku.applyDiffKT(y, y, p, p, 'laplacian', sig, order, dtype=dtype, **kw)
        assert rel(got, O.applyDiffKT(y, y, p, p, sig, kernel=kern, **kw)) < TOL_SUM[dtype]


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("order", ORDERS)
def test_diffK_ddiffK_ktensor(env, dtype, order):
    ku = env["ku"]
    kern = ('laplacian', order)
    rng = np.random.default_rng(100 + order)
    n = 800
    x = pts(rng, n)
    a1, a2, p = rng.normal(size=(n, 3)), rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    sig = [1.2, 3.0]
    for opt in ('1', '2', '1and2'):
        got = ku.applyDiffK(x, a1, a2, 'laplacian', sig, order, dtype=dtype, option=opt)
        assert rel(got, O.applyDiffK(x, a1, a2, sig, option=opt, kernel=kern)) < TOL_SUM[dtype], opt
    for opt in ('11', '12', '11and12'):
        got = ku.applyDDiffK(x, a1, a2, p, 'laplacian', sig, order, dtype=dtype, option=opt)
        assert rel(got, O.applyDDiffK(x, a1, a2, p, sig, option=opt, kernel=kern)) < TOL_SUM[dtype], opt
    y = pts(rng, 500)
    ay, ax, by, bx = rng.normal(size=500), rng.normal(size=n), rng.normal(size=(500, 3)), a1
    got = ku.applyktensor(y, x, ay, ax, by, bx, 'laplacian', sig, order, dtype=dtype)
    assert rel(got, O.applyktensor(y, x, ay, ax, by, bx, sig, kernel=kern)) < TOL_SUM[dtype]


@pytest.mark.parametrize("order", [0, 3])
def test_dense_weight_sums(env, order):
    ku = env["ku"]
    kern = ('laplacian', order)
    rng = np.random.default_rng(7 + order)
    y, x = pts(rng, 60), pts(rng, 90)
    beta = rng.normal(size=(60, 90, 2))
    got = ku.applykmat(y, x, beta, 'laplacian', [1.3, 2.0], order)
    assert rel(got, O.applykmat(y, x, beta, [1.3, 2.0], kernel=kern)) < 1e-12
    b2 = rng.normal(size=(60, 90))
    got = ku.applykdiffmat(y, x, b2, 'laplacian', [1.3], order)
    assert rel(got, O.applykdiffmat(y, x, b2, [1.3], kernel=kern)) < 1e-12


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("order,sig", [(1, [2.0]), (3, [1.5]), (4, [1.0, 3.0])])
def test_shooting_forward_adjoint(env, dtype, order, sig):
    pe, Kernel = env["pe"], env["Kernel"]
    kern = ('laplacian', order)
    rng = np.random.default_rng(order)
    n, T = 600, 5
    x0 = rng.normal(size=(n, 3)) * 4
    at = rng.normal(size=(T, n, 3)) * 0.3
    px1 = rng.normal(size=(n, 3))
    K = Kernel(name='laplacian', sigma=sig, order=order)
    K.pk_dtype = dtype
    xt = pe.landmarkDirectEvolutionEuler(x0, at, K)['xt']
    xr = O.landmarkDirectEvolutionEuler(x0, at, sig, kernel=kern)
    assert rel(xt, xr) < TOL_SHOOT[dtype]
    dat, _ = pe.landmarkHamiltonianGradient(x0, at, px1, K, 0.7)
    dr, _, _ = O.landmarkHamiltonianGradient(x0, at, px1, sig, 0.7, kernel=kern)
    assert rel(dat, dr) < TOL_SHOOT[dtype]
    xt_d, e = pe.shoot(x0, at, K, regweight=0.7, energy=True)
    er = sum(0.7 / T * (at[t] * O.applyK(xr[t], xr[t], at[t], sig, kernel=kern)).sum() for t in range(T))
    assert abs(e.item() - er) / abs(er) < TOL_SUM[dtype]
    g, _ = pe.shoot_adjoint(xt_d, at, px1, K, 0.7, euclidean=True)
    gr = np.stack([O.applyK(xr[t], xr[t], dr[t], sig, kernel=kern) for t in range(T)]) / T
    assert rel(g.cpu().numpy(), gr) < 5 * TOL_SHOOT[dtype]


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("order", [2, 3])
def test_epdiff_forward_adjoint(env, dtype, order):
    pe, Kernel = env["pe"], env["Kernel"]
    kern = ('laplacian', order)
    rng = np.random.default_rng(30 + order)
    n, T = 500, 5
    x0 = rng.normal(size=(n, 3)) * 3
    a0 = rng.normal(size=(n, 3)) * 0.3
    px1, pa1 = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    K = Kernel(name='laplacian', sigma=[1.3], order=order)
    K.pk_dtype = dtype
    st = pe.secondOrderEvolution(x0, a0, K, 1.0 / T)
    xr, ar = O.secondOrderEvolution(x0, a0, [1.3], 1.0 / T, kernel=kern)
    assert rel(st['xt'], xr) < TOL_SHOOT[dtype]
    assert rel(st['at'], ar) < TOL_SHOOT[dtype]
    res = pe.secondOrderGradient(x0, a0, px1, pa1, K, 1.0 / T)
    da0, _, _, pxr, par = O.secondOrderGradient(x0, a0, px1, pa1, [1.3], 1.0 / T, kernel=kern)
    assert rel(res['pxt'], pxr) < 5 * TOL_SHOOT[dtype]
    assert rel(res['pat'], par) < 5 * TOL_SHOOT[dtype]
    assert rel(res['da0'], da0) < 5 * TOL_SHOOT[dtype]


def test_population_shoot_matches_oracle(env):
    torch, B, S = env["torch"], env["B"], env["Surface"]
    prs = env["synth"].population_pairs(4, seed=9, nmin=80, nmax=900)
    pop = env["Population"]([S((f, v)) for (v, f), _ in prs], [S((f2, v2)) for _, (v2, f2) in prs])
    try:
        kd = B.kernel_for('laplacian', [3.0], 3, -1, 'float64')
        T = 4
        rng = np.random.default_rng(2)
        at = torch.tensor(rng.normal(size=(T, pop.V, 3)) * 0.3, device='cuda')
        xt = pop.x0.expand(T + 1, -1, -1).clone()
        en = torch.zeros(pop.P, dtype=torch.float64, device='cuda')
        pop.shoot(kd, at, xt, 0.7, en)
        for p in range(pop.P):
            r0, r1 = int(pop.voff[p]), int(pop.voff[p + 1])
            xr = O.landmarkDirectEvolutionEuler(pop.x0[r0:r1].cpu().numpy(), at[:, r0:r1].cpu().numpy(),
                                                [3.0], kernel=('laplacian', 3))
            assert rel(xt[:, r0:r1].cpu(), xr) < 1e-11
    finally:
        pop.close()


def test_surface_matching_objective_and_gradient(env):
    torch, Kernel, S = env["torch"], env["Kernel"], env["Surface"]
    (v, f), (v2, _) = env["synth"].plain_matching_pair(seed=7, subdiv=3, radius=10.0, nbumps=20, width=3.0)
    opts = {'KparDiff': Kernel(name='laplacian', sigma=2.5, order=3), 'KparDist': Kernel(name='gauss', sigma=1.0),
            'errorType': 'varifold', 'sigmaError': 1.0, 'timeStep': 0.1, 'pk_dtype': 'float64',
            'regWeight': 1.0, 'mode': 'quiet'}
    sm = env["SurfaceMatching"](Template=S((f, v)), Target=S((f, v2)), options=opts)
    at = np.random.default_rng(11).normal(size=(10, v.shape[0], 3)) * 0.05
    sm.control['at'] = torch.tensor(at, device='cuda')
    sm.obj = None
    kern = ('laplacian', 3)
    ref, _ = O.matching_objective(v, f, at, (v2, f), [2.5], [1.0], kernel_diff=kern)
    assert abs(sm.objectiveFun() - ref) / abs(ref) < 1e-10
    g = sm.getGradient()['at'].cpu().numpy()
    assert rel(g, O.matching_gradient(v, f, at, (v2, f), [2.5], [1.0], kernel_diff=kern)) < 1e-10


def test_varifold_data_term_rejects_laplacian(env):
    sd, Kernel, S = env["sd"], env["Kernel"], env["Surface"]
    (v, f), (v2, _) = env["synth"].plain_matching_pair(seed=1, subdiv=2, radius=10.0)
    with pytest.raises(Exception, match="gauss"):
        sd.varifoldNormDef(S((f, v)), S((f, v2)), Kernel(name='laplacian', sigma=1.0))
