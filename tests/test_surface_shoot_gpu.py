"""GPU parity of the surface data terms and the shooting loops against the oracle.

FP64: <= 1e-11 relative (norm-wise).  FP32 compensated: energies <= 1e-5
relative; gradients and trajectories <= 2e-5 norm-wise (the varifold
gradient is a difference of two kernel blocks, which loses a few bits).
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lddmm_amd import surfaceDistances as sd, pointEvolution as pe, synth
    from lddmm_amd.surfaces import Surface
    from lddmm_amd.kernelFunctions import Kernel
    return sd, pe, synth, Surface, Kernel


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


TOL = {'float64': 1e-11, 'float32': 2e-5}


def surfaces(synth, Surface, subdiv, radius=10.0, nb=8, width=3.0, amp=1.0, seed=0):
    (v, f), (v2, _) = synth.plain_matching_pair(seed=seed, subdiv=subdiv, radius=radius, nbumps=nb,
                                                width=width, amp_sd=amp)
    return Surface((f, v)), Surface((f, v2)), (v, f), (v2, f)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("subdiv,sig", [(2, [1.0]), (3, [1.0]), (3, [0.7, 2.0]), (4, [1.0])])
def test_varifold_energy_and_gradient(mods, dtype, subdiv, sig):
    sd, pe, synth, Surface, Kernel = mods
    S0, S1, A, Bt = surfaces(synth, Surface, subdiv)
    K = Kernel(name='gauss', sigma=sig)
    K.pk_dtype = dtype
    e0 = sd.varifoldNorm0(S1, K)
    eD = sd.varifoldNormDef(S0, S1, K)
    r0, rD = O.varifoldNorm0(Bt, sig), O.varifoldNormDef(A, Bt, sig)
    tol = 1e-11 if dtype == 'float64' else 1e-5
    assert abs(e0 - r0) / abs(r0) < tol
    assert abs(eD - rD) / abs(rD) < tol
    g = sd.varifoldNormGradient(S0, S1, K)
    gr = O.varifoldNormGradient(A, Bt, sig)
    assert g.shape == gr.shape
    assert rel(g, gr) < TOL[dtype]


def test_varifold_self_distance_and_weights(mods):
    sd, pe, synth, Surface, Kernel = mods
    S0, S1, A, Bt = surfaces(synth, Surface, 2)
    K = Kernel(name='gauss', sigma=1.0)
    d = sd.varifoldNorm(S0, S0, K)
    assert abs(d) < 1e-10 * sd.varifoldNorm0(S0, K)
    # vertex weights enter through the face weights
    w = np.linspace(0.5, 1.5, S0.vertices.shape[0])
    Sw = Surface((S0.faces, S0.vertices), weights=w)
    ref = O.varifoldNormDef((A[0], A[1], w), Bt, [1.0])
    assert abs(sd.varifoldNormDef(Sw, S1, K) - ref) / abs(ref) < 1e-11
    assert rel(sd.varifoldNormGradient(Sw, S1, K), O.varifoldNormGradient((A[0], A[1], w), Bt, [1.0])) < 1e-11


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_current_terms(mods, dtype):
    sd, pe, synth, Surface, Kernel = mods
    S0, S1, A, Bt = surfaces(synth, Surface, 3)
    K = Kernel(name='gauss', sigma=1.5)
    K.pk_dtype = dtype
    tol = 1e-11 if dtype == 'float64' else 1e-5
    assert abs(sd.currentNorm0(S1, K) - O.currentNorm0(Bt, [1.5])) / O.currentNorm0(Bt, [1.5]) < tol
    rD = O.currentNormDef(A, Bt, [1.5])
    assert abs(sd.currentNormDef(S0, S1, K) - rD) / abs(rD) < tol
    assert rel(sd.currentNormGradient(S0, S1, K), O.currentNormGradient(A, Bt, [1.5])) < TOL[dtype]
    assert rel(sd.currentMagnitudeGradient(S0, K), O.currentMagnitudeGradient(A, [1.5])) < TOL[dtype]


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("n,T,sig", [(1, 3, [1.0]), (300, 5, [1.5]), (2562, 10, [2.5]), (700, 4, [1.0, 3.0])])
def test_shooting_forward_adjoint(mods, dtype, n, T, sig):
    sd, pe, synth, Surface, Kernel = mods
    rng = np.random.default_rng(n + T)
    x0 = rng.normal(size=(n, 3)) * 4
    at = rng.normal(size=(T, n, 3)) * 0.3
    px1 = rng.normal(size=(n, 3))
    K = Kernel(name='gauss', sigma=sig)
    K.pk_dtype = dtype
    xt = pe.landmarkDirectEvolutionEuler(x0, at, K)['xt']
    xr = O.landmarkDirectEvolutionEuler(x0, at, sig)
    assert rel(xt, xr) < TOL[dtype]
    pxt, _ = pe.landmarkHamiltonianCovector(x0, at, px1, K, 0.7)
    pr, _ = O.landmarkHamiltonianCovector(x0, at, px1, sig, 0.7)
    assert rel(pxt, pr) < TOL[dtype]
    dat, _ = pe.landmarkHamiltonianGradient(x0, at, px1, K, 0.7)
    dr, _, _ = O.landmarkHamiltonianGradient(x0, at, px1, sig, 0.7)
    assert rel(dat, dr) < TOL[dtype]
    # energy + Euclidean gradient of the device loops
    xt_d, e = pe.shoot(x0, at, K, regweight=0.7, energy=True)
    er = sum(0.7 / T * (at[t] * O.applyK(xr[t], xr[t], at[t], sig)).sum() for t in range(T))
    assert abs(e.item() - er) / abs(er) < (1e-11 if dtype == 'float64' else 1e-5)
    g, _ = pe.shoot_adjoint(xt_d, at, px1, K, 0.7, euclidean=True)
    gr = np.stack([O.applyK(xr[t], xr[t], dr[t], sig) for t in range(T)]) / T
    assert rel(g.cpu().numpy(), gr) < TOL[dtype]


def test_zero_control_identity(mods):
    sd, pe, synth, Surface, Kernel = mods
    x0 = np.random.default_rng(1).normal(size=(50, 3))
    xt = pe.landmarkDirectEvolutionEuler(x0, np.zeros((4, 50, 3)), Kernel(name='gauss', sigma=1.0))['xt']
    np.testing.assert_array_equal(xt, np.broadcast_to(x0, xt.shape))


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("n,T,sig", [(2, 3, [1.0]), (400, 5, [1.3]), (900, 10, [0.8, 2.0])])
def test_epdiff_forward_adjoint(mods, dtype, n, T, sig):
    sd, pe, synth, Surface, Kernel = mods
    rng = np.random.default_rng(3 * n + T)
    x0 = rng.normal(size=(n, 3)) * 3
    a0 = rng.normal(size=(n, 3)) * 0.3
    px1, pa1 = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    K = Kernel(name='gauss', sigma=sig)
    K.pk_dtype = dtype
    st = pe.secondOrderEvolution(x0, a0, K, 1.0 / T)
    xr, ar = O.secondOrderEvolution(x0, a0, sig, 1.0 / T)
    assert rel(st['xt'], xr) < TOL[dtype]
    assert rel(st['at'], ar) < TOL[dtype]
    xe, ae = pe.landmarkEPDiff(T, x0, a0, K)
    xe_r, ae_r = O.landmarkEPDiff(T, x0, a0, sig)
    assert rel(xe, xe_r) < TOL[dtype] and rel(ae, ae_r) < TOL[dtype]
    res = pe.secondOrderGradient(x0, a0, px1, pa1, K, 1.0 / T)
    da0, _, _, pxr, par = O.secondOrderGradient(x0, a0, px1, pa1, sig, 1.0 / T)
    assert rel(res['pxt'], pxr) < TOL[dtype] * 5
    assert rel(res['pat'], par) < TOL[dtype] * 5
    assert rel(res['da0'], da0) < TOL[dtype] * 5
    H = pe.secondOrderHamiltonian(x0, a0, px1, pa1, K)
    Hr = O.secondOrderHamiltonian(x0, a0, px1, pa1, sig)
    assert abs(H - Hr) / abs(Hr) < (1e-11 if dtype == 'float64' else 1e-5)


_LARGE_EPD = {}


def _large_epdiff_case():
    """n = 4,200 points, T = 2: above the one-launch step limit (4,096), so
    the chunked pair kernels and their reduce kernels run; the oracle's
    forward and adjoint are computed once for both precisions."""
    if not _LARGE_EPD:
        n, T, sig = 4200, 2, [1.3]
        rng = np.random.default_rng(4200)
        x0 = rng.normal(size=(n, 3)) * 5
        a0 = rng.normal(size=(n, 3)) * 0.3
        px1, pa1 = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
        xr, ar = O.secondOrderEvolution(x0, a0, sig, 1.0 / T)
        da0, _, _, pxr, par = O.secondOrderGradient(x0, a0, px1, pa1, sig, 1.0 / T)
        _LARGE_EPD.update(T=T, sig=sig, inputs=(x0, a0, px1, pa1), fwd=(xr, ar), adj=(da0, pxr, par))
    return _LARGE_EPD


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_epdiff_chunked_path_above_step_limit(mods, dtype):
    sd, pe, synth, Surface, Kernel = mods
    case = _large_epdiff_case()
    x0, a0, px1, pa1 = case['inputs']
    K = Kernel(name='gauss', sigma=case['sig'])
    K.pk_dtype = dtype
    dt = 1.0 / case['T']
    st = pe.secondOrderEvolution(x0, a0, K, dt)
    xr, ar = case['fwd']
    assert rel(st['xt'], xr) < TOL[dtype]
    assert rel(st['at'], ar) < TOL[dtype]
    res = pe.secondOrderGradient(x0, a0, px1, pa1, K, dt)
    da0, pxr, par = case['adj']
    assert rel(res['pxt'], pxr) < TOL[dtype] * 5
    assert rel(res['pat'], par) < TOL[dtype] * 5
    assert rel(res['da0'], da0) < TOL[dtype] * 5


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_epdiff_forward_adjoint_2d(mods, dtype):
    """Planar points (dim 2) through the geodesic forward and adjoint."""
    sd, pe, synth, Surface, Kernel = mods
    rng = np.random.default_rng(21)
    n, T, sig = 300, 6, [1.1]
    x0 = rng.normal(size=(n, 2)) * 3
    a0 = rng.normal(size=(n, 2)) * 0.3
    px1, pa1 = rng.normal(size=(n, 2)), rng.normal(size=(n, 2))
    K = Kernel(name='gauss', sigma=sig)
    K.pk_dtype = dtype
    st = pe.secondOrderEvolution(x0, a0, K, 1.0 / T)
    xr, ar = O.secondOrderEvolution(x0, a0, sig, 1.0 / T)
    assert rel(st['xt'], xr) < TOL[dtype]
    assert rel(st['at'], ar) < TOL[dtype]
    res = pe.secondOrderGradient(x0, a0, px1, pa1, K, 1.0 / T)
    da0, _, _, pxr, par = O.secondOrderGradient(x0, a0, px1, pa1, sig, 1.0 / T)
    assert rel(res['pxt'], pxr) < TOL[dtype] * 5
    assert rel(res['pat'], par) < TOL[dtype] * 5
    assert rel(res['da0'], da0) < TOL[dtype] * 5


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_geodesic_hamiltonian_conservation(mods, dtype):
    """H = 1/2 a.K(x)a along EPDiff, evaluated in float64 on the device: equals
    the oracle's H of the same trajectory, and its drift shrinks like dt."""
    sd, pe, synth, Surface, Kernel = mods
    rng = np.random.default_rng(17)
    x0 = rng.normal(size=(500, 3)) * 3
    a0 = rng.normal(size=(500, 3)) * 0.3
    K = Kernel(name='gauss', sigma=1.5)
    K.pk_dtype = dtype
    drifts = []
    for T in (10, 40):
        xt, at = pe.landmarkEPDiff(T, x0, a0, K)
        H, drift = pe.geodesicHamiltonian(xt, at, K)
        Hr = np.array([0.5 * (at[t] * O.applyK(xt[t], xt[t], at[t], [1.5])).sum() for t in range(T)])
        assert np.max(np.abs(H - Hr) / np.abs(Hr)) < 1e-12
        drifts.append(drift)
    assert drifts[1] < 0.5 * drifts[0]
    assert drifts[0] < 0.05
