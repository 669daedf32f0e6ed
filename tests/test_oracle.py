"""Pins the FP64 oracle (oracle/oracle.py) without retyping its formulas.

Each check is chosen so that a plausible mistake fails it: closed forms catch
a dropped 1/2 or a wrong sigma power, finite differences catch wrong signs
and missing terms of the derivative sums, symmetry identities catch swapped
operands, and invariances catch orientation / frame mistakes in the data
terms.  The reference's own finite-difference self-checks are mirrored where
they exist (Kernel.testDiffKT, base/kernelFunctions.py:629-635;
testEndpointGradient, base/surfaceMatching.py:678-698; the BFGS gradient
test, base/bfgs.py:167-176).
"""
import numpy as np
import pytest

from oracle import oracle as O
from lddmm_amd import synth as S

RNG = np.random.default_rng(1234)


def rnd(*shape, scale=1.0):
    return RNG.normal(size=shape) * scale


# ----------------------------------------------------------------- applyK
def test_applyK_single_point_returns_weight():
    x = rnd(1, 3)
    a = rnd(1, 3)
    np.testing.assert_allclose(O.applyK(x, x, a, [1.7]), a, rtol=1e-15)


def test_applyK_two_points_closed_form():
    s = 1.3
    y = np.array([[0., 0., 0.]])
    x = np.array([[0.4, -1.2, 0.7]])
    a = np.array([[2.0, -1.0, 0.5]])
    r2 = (x ** 2).sum()
    np.testing.assert_allclose(O.applyK(y, x, a, [s]), np.exp(-r2 / (2 * s * s)) * a, rtol=1e-14)


def test_applyK_multiscale_weights():
    # w_s = s^-1 / sum s^-1 ; checked on a two-point pair against the closed form
    y = np.zeros((1, 3))
    x = np.array([[1.0, 0.5, -0.3]])
    a = np.ones((1, 1))
    s = np.array([0.8, 2.0, 5.0])
    r2 = (x ** 2).sum()
    w = (1 / s) / (1 / s).sum()
    expect = (w * np.exp(-r2 / (2 * s ** 2))).sum()
    np.testing.assert_allclose(O.applyK(y, x, a, s)[0, 0], expect, rtol=1e-14)
    # a repeated sigma is the same kernel
    np.testing.assert_allclose(O.applyK(y, x, a, [2.0, 2.0]), O.applyK(y, x, a, [2.0]), rtol=1e-15)


def test_applyK_large_sigma_limit_is_column_sum():
    y, x, a = rnd(5, 3), rnd(7, 3), rnd(7, 2)
    np.testing.assert_allclose(O.applyK(y, x, a, [1e7]), np.tile(a.sum(0), (5, 1)), rtol=1e-10)


def test_applyK_adjoint_symmetry():
    y, x = rnd(40, 3), rnd(60, 3)
    p, a = rnd(40, 3), rnd(60, 3)
    lhs = (p * O.applyK(y, x, a, [1.5])).sum()
    rhs = (a * O.applyK(x, y, p, [1.5])).sum()
    np.testing.assert_allclose(lhs, rhs, rtol=1e-13)


def test_applyK_translation_rotation_invariance():
    y, x, a = rnd(20, 3), rnd(30, 3), rnd(30, 1)
    Q, _ = np.linalg.qr(rnd(3, 3))
    t = rnd(1, 3) * 10
    np.testing.assert_allclose(O.applyK(y @ Q.T + t, x @ Q.T + t, a, [1.1]),
                               O.applyK(y, x, a, [1.1]), rtol=1e-12)


# ------------------------------------------------------------ applyDiffKT
def _fd(f, x, dx, eps=1e-6):
    return (f(x + eps * dx) - f(x - eps * dx)) / (2 * eps)


@pytest.mark.parametrize("sig", [[1.2], [0.9, 2.1]])
def test_diffKT_plain_is_gradient_wrt_targets(sig):
    y, x, p, a = rnd(15, 3), rnd(20, 3), rnd(15, 3), rnd(20, 3)
    dy = rnd(15, 3)
    fd = _fd(lambda yy: (p * O.applyK(yy, x, a, sig)).sum(), y, dy)
    g = O.applyDiffKT(y, x, p, a, sig)
    np.testing.assert_allclose((g * dy).sum(), fd, rtol=1e-7)


def test_diffKT_sym_is_gradient_of_pKa():
    x, p, a = rnd(18, 3), rnd(18, 3), rnd(18, 3)
    dx = rnd(18, 3)
    fd = _fd(lambda xx: (p * O.applyK(xx, xx, a, [1.4])).sum(), x, dx)
    g = O.applyDiffKT(x, x, p, a, [1.4], sym=True)
    np.testing.assert_allclose((g * dx).sum(), fd, rtol=1e-7)


def test_diffKT_lddmm_is_gradient_of_pKa_minus_raKa():
    x, p, a = rnd(18, 3), rnd(18, 3), rnd(18, 3)
    dx = rnd(18, 3)
    r = 0.7
    f = lambda xx: (p * O.applyK(xx, xx, a, [1.4])).sum() - r * (a * O.applyK(xx, xx, a, [1.4])).sum()
    g = O.applyDiffKT(x, x, p, a, [1.4], regweight=r, lddmm=True)
    np.testing.assert_allclose((g * dx).sum(), _fd(f, x, dx), rtol=1e-7)


# ------------------------------------------------------------ applyktensor
def test_ktensor_single_pair_closed_form():
    y = np.zeros((1, 3))
    x = np.array([[0.3, 0.2, -0.1]])
    by, bx = np.array([[1.0, 2.0, 0.5]]), np.array([[-0.5, 1.0, 2.0]])
    ay, ax = np.array([1.5]), np.array([0.7])
    K = np.exp(-(x ** 2).sum() / 2 / 0.8 ** 2)
    expect = K * (1.5 * 0.7 + (by * bx).sum() ** 2)
    np.testing.assert_allclose(O.applyktensor(y, x, ay, ax, by, bx, [0.8]), [expect], rtol=1e-14)


def test_ktensor_symmetry():
    y, x = rnd(25, 3), rnd(30, 3)
    ay, ax, by, bx = rnd(25), rnd(30), rnd(25, 3), rnd(30, 3)
    np.testing.assert_allclose(O.applyktensor(y, x, ay, ax, by, bx, [1.0]).sum(),
                               O.applyktensor(x, y, ax, ay, bx, by, [1.0]).sum(), rtol=1e-13)


# --------------------------------------------------- dense-weight helpers
def test_kmat_reduces_to_applyK():
    y, x, a = rnd(12, 3), rnd(17, 3), rnd(17, 3)
    beta = np.broadcast_to(a[None], (12, 17, 3))
    np.testing.assert_allclose(O.applykmat(y, x, beta, [1.3, 2.0]), O.applyK(y, x, a, [1.3, 2.0]),
                               rtol=1e-13)


def test_kdiffmat_reduces_to_diffKT_plain():
    y, x, p, a = rnd(12, 3), rnd(17, 3), rnd(12, 3), rnd(17, 3)
    beta = p @ a.T
    np.testing.assert_allclose(O.applykdiffmat(y, x, beta, [1.3]), O.applyDiffKT(y, x, p, a, [1.3]),
                               rtol=1e-12)


# ------------------------------------------- first / second derivatives
@pytest.mark.parametrize("sig", [[1.1], [0.8, 1.9]])
def test_diffK_1and2_is_directional_derivative_of_applyK(sig):
    x, a1, a2 = rnd(16, 3), rnd(16, 3), rnd(16, 2)
    fd = _fd(lambda t: O.applyK(x + t * a1, x + t * a1, a2, sig), 0.0, 1.0)
    np.testing.assert_allclose(O.applyDiffK(x, a1, a2, sig, option='1and2'), fd, rtol=1e-6, atol=1e-9)


def test_diffK_options_1_plus_2_make_1and2():
    x, a1, a2 = rnd(16, 3), rnd(16, 3), rnd(16, 3)
    tot = O.applyDiffK(x, a1, a2, [1.2], option='1') + O.applyDiffK(x, a1, a2, [1.2], option='2')
    np.testing.assert_allclose(tot, O.applyDiffK(x, a1, a2, [1.2], option='1and2'), rtol=1e-12)
    # option 1 only moves the target point: derivative of sum_l K(x_i + t a1_i, x_l) a2_l
    i = 3
    def row(t):
        xx = x.copy()
        yi = x[i:i + 1] + t * a1[i:i + 1]
        return O.applyK(yi, xx, a2, [1.2])[0]
    np.testing.assert_allclose(O.applyDiffK(x, a1, a2, [1.2], option='1')[i], _fd(row, 0.0, 1.0),
                               rtol=1e-6, atol=1e-10)


@pytest.mark.parametrize("sig", [[1.3], [0.9, 2.2]])
def test_ddiffK_11and12_is_derivative_of_diffKT_plain(sig):
    x, a1, a2, p = rnd(14, 3), rnd(14, 3), rnd(14, 3), rnd(14, 3)
    fd = _fd(lambda t: O.applyDiffKT(x + t * p, x + t * p, a1, a2, sig), 0.0, 1.0)
    np.testing.assert_allclose(O.applyDDiffK(x, a1, a2, p, sig, option='11and12'), fd,
                               rtol=1e-6, atol=1e-9)


def test_ddiffK_options_11_plus_12_make_11and12():
    x, a1, a2, p = rnd(14, 3), rnd(14, 3), rnd(14, 3), rnd(14, 3)
    tot = O.applyDDiffK(x, a1, a2, p, [1.2], option='11') + O.applyDDiffK(x, a1, a2, p, [1.2], option='12')
    np.testing.assert_allclose(tot, O.applyDDiffK(x, a1, a2, p, [1.2], option='11and12'), rtol=1e-12)


# --------------------------------------------------------- surface terms
@pytest.fixture(scope="module")
def spheres():
    v, f = S.icosphere(2, radius=3.0)
    v2 = S.bump_deform(v, f, nbumps=6, width=1.0, amp_sd=0.3, seed=3)
    return (v, f), (v2, f)


def test_measure_terms(spheres):
    """Surface measure: |mu - mu| = 0, face-flip invariance, sigma -> inf limit
    (total area)^2, FD gradient; the point-set form with the target's face
    centres / areas equals the surface form; point-set measure FD gradient."""
    A, Bs = spheres
    v, f = A
    v2, f2 = Bs
    assert abs(O.measureNormDef(A, A, [0.8]) + O.measureNorm0(A, [0.8])) < 1e-9 * O.measureNorm0(A, [0.8])
    flip = (v2, f2[:, [0, 2, 1]])
    np.testing.assert_allclose(O.measureNormDef(A, flip, [0.8]), O.measureNormDef(A, Bs, [0.8]), rtol=1e-12)
    area = np.linalg.norm(np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]]), axis=1).sum() / 2
    np.testing.assert_allclose(O.measureNorm0(A, [1e6]), area ** 2, rtol=1e-8)
    dv = rnd(*v.shape)
    fd = _fd(lambda vv: O.measureNormDef((vv, f), Bs, [0.8]), v, dv)
    np.testing.assert_allclose((O.measureNormGradient(A, Bs, [0.8]) * dv).sum(), fd, rtol=1e-6)
    c2, S2, _ = O.surface_geometry(v2, f2)
    a2 = np.sqrt((S2 ** 2).sum(1) + 1e-10)
    np.testing.assert_allclose(O.measureNormPSDef(A, c2, a2, [0.8]), O.measureNormDef(A, Bs, [0.8]), rtol=1e-12)
    np.testing.assert_allclose(O.measureNormPS0(c2, a2, [0.8]), O.measureNorm0(Bs, [0.8]), rtol=1e-12)
    fd = _fd(lambda vv: O.measureNormPSDef((vv, f), c2, a2, [0.8]), v, dv)
    np.testing.assert_allclose((O.measureNormPSGradient(A, c2, a2, [0.8]) * dv).sum(), fd, rtol=1e-6)
    # weighted template: surface-form gradient stays the FD derivative
    w = np.random.default_rng(2).uniform(0.5, 1.5, size=v.shape[0])
    fd = _fd(lambda vv: O.measureNormDef((vv, f, w), Bs, [0.8]), v, dv)
    np.testing.assert_allclose((O.measureNormGradient((v, f, w), Bs, [0.8]) * dv).sum(), fd, rtol=1e-6)
    # point sets
    x, y = rnd(30, 3), rnd(25, 3)
    wx, wy = np.abs(rnd(30)) + 0.1, np.abs(rnd(25)) + 0.1
    np.testing.assert_allclose(O.pointMeasureNorm0(x, wx, [1e6]), wx.sum() ** 2, rtol=1e-8)
    dx = rnd(30, 3)
    fd = _fd(lambda xx: O.pointMeasureNormDef(xx, wx, y, wy, [1.3]), x, dx)
    np.testing.assert_allclose((O.pointMeasureNormGradient(x, wx, y, wy, [1.3]) * dx).sum(), fd, rtol=1e-6)


def test_point_set_matching_gradient_fd():
    """Point-set measure matching: Euclidean gradient = FD of the objective."""
    x0 = rnd(40, 3) * 2
    w0 = np.abs(rnd(40)) + 0.2
    y, wy = rnd(35, 3) * 2 + 0.3, np.abs(rnd(35)) + 0.2
    at = 0.05 * rnd(4, 40, 3)
    d = rnd(*at.shape)
    g = O.matching_gradient(x0, w0, at, (y, wy), [1.5], [0.9], kind='point_measure')
    fd = _fd(lambda a_: O.matching_objective(x0, w0, a_, (y, wy), [1.5], [0.9], kind='point_measure')[0], at, d)
    np.testing.assert_allclose((g * d).sum(), fd, rtol=1e-6)


def test_icosphere_closed_forms():
    v, f = S.icosphere(3, radius=2.0)
    assert v.shape == (642, 3) and f.shape == (1280, 3)
    np.testing.assert_allclose(np.linalg.norm(v, axis=1), 2.0, rtol=1e-14)
    vol = O.surface_volume(v, f)
    assert vol > 0  # outward orientation
    assert abs(vol - 4 / 3 * np.pi * 8) / (4 / 3 * np.pi * 8) < 0.02
    c, Sf, fw = O.surface_geometry(v, f)
    area = np.linalg.norm(Sf, axis=1).sum()
    assert abs(area - 4 * np.pi * 4) / (4 * np.pi * 4) < 0.02


def test_varifold_self_distance_zero(spheres):
    A, _ = spheres
    np.testing.assert_allclose(O.varifoldNorm(A, A, [1.0]), 0.0, atol=1e-9 * O.varifoldNorm0(A, [1.0]))


def test_varifold_rigid_and_flip_invariance(spheres):
    A, Bs = spheres
    d0 = O.varifoldNorm(A, Bs, [0.8])
    Q, _ = np.linalg.qr(rnd(3, 3))
    t = rnd(1, 3)
    A2 = (A[0] @ Q.T + t, A[1])
    B2 = (Bs[0] @ Q.T + t, Bs[1])
    np.testing.assert_allclose(O.varifoldNorm(A2, B2, [0.8]), d0, rtol=1e-11)
    Bflip = (Bs[0], Bs[1][:, [0, 2, 1]])
    np.testing.assert_allclose(O.varifoldNorm(A, Bflip, [0.8]), d0, rtol=1e-12)
    assert d0 > 0


def test_varifold_single_triangle_pair_closed_form():
    f = np.array([[0, 1, 2]])
    A = (np.array([[0., 0, 0], [1, 0, 0], [0, 1, 0]]), f)
    B = (np.array([[0., 0, 1], [1, 0, 1], [0, 0, 2]]), f)
    cA, SA, _ = O.surface_geometry(*A)
    cB, SB, _ = O.surface_geometry(*B)
    aA, aB = np.linalg.norm(SA), np.linalg.norm(SB)
    cos = (SA[0] @ SB[0]) / (aA * aB)
    K = np.exp(-((cA - cB) ** 2).sum() / 2)
    cross = aA * aB * K * (1 + cos ** 2)
    expect = 2 * aA ** 2 + 2 * aB ** 2 - 2 * cross   # E(A,A) = 2 a^2 (cos = 1)
    np.testing.assert_allclose(O.varifoldNorm(A, B, [1.0]), expect, rtol=1e-9)


@pytest.mark.parametrize("sig", [[0.7], [0.5, 1.5]])
def test_varifold_gradient_fd(spheres, sig):
    A, Bs = spheres
    v, f = A
    dv = rnd(*v.shape)
    fd = _fd(lambda vv: O.varifoldNorm((vv, f), Bs, sig), v, dv, eps=1e-6)
    g = O.varifoldNormGradient(A, Bs, sig)
    np.testing.assert_allclose((g * dv).sum(), fd, rtol=1e-6)


def test_varifold_gradient_vertex_rows_equal_full(spheres):
    A, Bs = spheres
    rows = np.array([0, 5, 17, 40, 41, A[0].shape[0] - 1])
    full = O.varifoldNormGradient(A, Bs, [0.7])
    np.testing.assert_array_equal(O.varifoldNormGradient(A, Bs, [0.7], vertex_rows=rows), full[rows])


@pytest.mark.parametrize("mode", ["plain", "sym", "lddmm"])
def test_diffKT_target_rows_equal_full(mode):
    x, p, a = rnd(40, 3), rnd(40, 3), rnd(40, 3)
    kw = dict(sym=mode == 'sym', lddmm=mode == 'lddmm', regweight=0.7)
    rows = np.array([3, 0, 39, 11])
    full = O.applyDiffKT(x, x, p, a, [1.1, 2.0], **kw)
    np.testing.assert_allclose(O.applyDiffKT(x, x, p, a, [1.1, 2.0], rows=rows, **kw), full[rows],
                               rtol=1e-14, atol=1e-15)


def _weighted(spheres, rng_seed=5):
    (v, f), (v2, f2) = spheres
    r = np.random.default_rng(rng_seed)
    return v, f, r.uniform(0.5, 1.5, v.shape[0]), v2, f2, r.uniform(0.5, 1.5, v2.shape[0])


@pytest.mark.parametrize("kind", ["current", "measure"])
def test_vertex_weight_gradients_fd(spheres, kind):
    """current / measure: the weight part is the derivative of the distance
    in the template's vertex weights (base/surfaceDistances.py:130-137, :269-276)."""
    v, f, w, v2, f2, w2 = _weighted(spheres)
    norm = {'current': O.currentNorm, 'measure': O.measureNorm}[kind]
    grad = {'current': O.currentNormGradient, 'measure': O.measureNormGradient}[kind]
    px, pxw = grad((v, f, w), (v2, f2, w2), [0.9], with_weights=True)
    np.testing.assert_array_equal(px, grad((v, f, w), (v2, f2, w2), [0.9]))
    dw = rnd(v.shape[0])
    fd = _fd(lambda ww: norm((v, f, ww), (v2, f2, w2), [0.9]), w, dw, eps=1e-6)
    np.testing.assert_allclose((pxw * dw).sum(), fd, rtol=1e-6)
    pxm, pxwm = O.currentMagnitudeGradient((v, f, w), [0.9], with_weights=True)
    fdm = _fd(lambda ww: O.currentNorm0((v, f, ww), [0.9]), w, dw, eps=1e-6)
    np.testing.assert_allclose((pxwm * dw).sum(), fdm, rtol=1e-6)


@pytest.mark.parametrize("c", [1.0, 1.7])
def test_varifold_vertex_weight_gradient(spheres, c):
    """The varifold weight part keeps the face weight fw_f in its per-face
    factor (base/surfaceDistances.py:455-462): with uniform template weights c
    it is c times the derivative of the distance in the vertex weights."""
    (v, f), (v2, f2) = spheres
    w = np.full(v.shape[0], c)
    w2 = np.random.default_rng(8).uniform(0.5, 1.5, v2.shape[0])
    px, pxw = O.varifoldNormGradient((v, f, w), (v2, f2, w2), [0.8], with_weights=True)
    np.testing.assert_array_equal(px, O.varifoldNormGradient((v, f, w), (v2, f2, w2), [0.8]))
    dw = rnd(v.shape[0])
    fd = _fd(lambda ww: O.varifoldNorm((v, f, ww), (v2, f2, w2), [0.8]), w, dw, eps=1e-6)
    np.testing.assert_allclose((pxw * dw).sum(), c * fd, rtol=1e-6)
    rows = np.array([1, 7, 30])
    pr, pwr = O.varifoldNormGradient((v, f, w), (v2, f2, w2), [0.8], vertex_rows=rows, with_weights=True)
    np.testing.assert_array_equal(pwr, pxw[rows])


def test_current_orientation_and_gradient(spheres):
    A, Bs = spheres
    v, f = A
    flip = (Bs[0], Bs[1][:, [0, 2, 1]])
    # flipping the target reverses the cross term only
    cross = -2 * (O.currentNormDef(A, Bs, [0.8]) - O.currentNorm0(A, [0.8]))
    cross_f = -2 * (O.currentNormDef(A, flip, [0.8]) - O.currentNorm0(A, [0.8]))
    np.testing.assert_allclose(cross_f, -cross, rtol=1e-12)
    np.testing.assert_allclose(O.currentNorm(A, A, [0.8]), 0.0, atol=1e-10 * O.currentNorm0(A, [0.8]))
    dv = rnd(*v.shape)
    fd = _fd(lambda vv: O.currentNorm((vv, f), Bs, [0.8]), v, dv)
    np.testing.assert_allclose((O.currentNormGradient(A, Bs, [0.8]) * dv).sum(), fd, rtol=1e-6)
    fd = _fd(lambda vv: O.currentMagnitude((vv, f), [0.8]), v, dv)
    np.testing.assert_allclose((O.currentMagnitudeGradient(A, [0.8]) * dv).sum(), fd, rtol=1e-6)


# ----------------------------------------------------------- curve terms
@pytest.fixture(scope="module")
def curves():
    t = np.linspace(0, 2 * np.pi, 41)[:-1]
    v1 = np.stack([np.cos(t), 0.7 * np.sin(t)], 1) * 2
    v2 = np.stack([np.cos(t) + 0.1 * np.cos(3 * t), 0.6 * np.sin(t)], 1) * 2.1
    f = np.stack([np.arange(40), (np.arange(40) + 1) % 40], 1)
    return (v1, f), (v2, f)


def test_curve_varifold(curves):
    A, Bc = curves
    v, f = A
    assert abs(O.curveVarifoldNormDef(A, A, [0.5]) + O.curveVarifoldNorm0(A, [0.5])) < 1e-9
    dv = rnd(*v.shape)
    fd = _fd(lambda vv: O.curveVarifoldNormDef((vv, f), Bc, [0.5]), v, dv)
    np.testing.assert_allclose((O.curveVarifoldNormGradient(A, Bc, [0.5]) * dv).sum(), fd, rtol=1e-6)
    flip = (Bc[0], Bc[1][:, [1, 0]])
    np.testing.assert_allclose(O.curveVarifoldNormDef(A, flip, [0.5]), O.curveVarifoldNormDef(A, Bc, [0.5]),
                               rtol=1e-12)


def test_curve_varifold_weight_and_3d():
    """Orientation weight d: FD gradient for d = 0.5 on a weighted 3-D curve; d = 0
    reduces to the length measure sum_ij K(c_i, c_j) |l_i| |l_j| (brute force)."""
    rng = np.random.default_rng(4)
    t = np.linspace(0, 2 * np.pi, 31)[:-1]
    v1 = np.stack([np.cos(t), np.sin(t), 0.3 * np.sin(2 * t)], 1) * 2
    v2 = v1 * 1.1 + 0.05 * rng.normal(size=v1.shape)
    f = np.stack([np.arange(30), (np.arange(30) + 1) % 30], 1)
    w = rng.uniform(0.5, 1.5, size=30)
    A, Bc = (v1, f, w), (v2, f)
    dv = rnd(*v1.shape)
    fd = _fd(lambda vv: O.curveVarifoldNormDef((vv, f, w), Bc, [0.7], weight=0.5), v1, dv)
    np.testing.assert_allclose((O.curveVarifoldNormGradient(A, Bc, [0.7], weight=0.5) * dv).sum(), fd,
                               rtol=1e-6)
    c = (v1[f[:, 0]] + v1[f[:, 1]]) / 2
    lw = (w[f[:, 0]] + w[f[:, 1]]) / 2
    a = np.sqrt(((v1[f[:, 1]] - v1[f[:, 0]]) ** 2).sum(1) * lw ** 2 + 1e-10)
    d2 = ((c[:, None] - c[None]) ** 2).sum(-1)
    brute = (np.exp(-d2 / (2 * 0.7 ** 2)) * a[:, None] * a[None]).sum()
    np.testing.assert_allclose(O.curveVarifoldNorm0(A, [0.7], weight=0.), brute, rtol=1e-12)


def test_curve_matching_gradient_fd(curves):
    """Curve-matching gradient as BFGS sees it, K(x_t)(2 r a_t - p_{t+1}) / T,
    is the exact derivative of the discrete objective (FD), for both data terms."""
    A, Bc = curves
    v, f = A
    at = 0.05 * rnd(5, *v.shape)
    dat = rnd(*at.shape)
    for kind in ('curve_varifold', 'curve_current'):
        g = O.matching_gradient(v, f, at, Bc, [1.5], [0.5], kind=kind, regweight=0.7)
        fd = _fd(lambda a_: O.matching_objective(v, f, a_, Bc, [1.5], [0.5], kind=kind, regweight=0.7)[0],
                 at, dat)
        np.testing.assert_allclose((g * dat).sum(), fd, rtol=1e-6)


def test_curve_current(curves):
    A, Bc = curves
    v, f = A
    dv = rnd(*v.shape)
    fd = _fd(lambda vv: O.curveCurrentNormDef((vv, f), Bc, [0.5]), v, dv)
    np.testing.assert_allclose((O.curveCurrentNormGradient(A, Bc, [0.5]) * dv).sum(), fd, rtol=1e-6)


# ------------------------------------------------------------- shooting
def test_zero_control_is_identity():
    x0 = rnd(10, 3)
    xt = O.landmarkDirectEvolutionEuler(x0, np.zeros((5, 10, 3)), [1.0])
    np.testing.assert_array_equal(xt, np.broadcast_to(x0, xt.shape))


def test_one_step_closed_form():
    x0, a = rnd(6, 3), rnd(1, 6, 3)
    xt = O.landmarkDirectEvolutionEuler(x0, a, [1.2])
    np.testing.assert_allclose(xt[1], x0 + O.applyK(x0, x0, a[0], [1.2]), rtol=1e-14)


def test_matching_gradient_fd(spheres):
    A, Bs = spheres
    v, f = A
    T = 4
    at = rnd(T, v.shape[0], 3, scale=0.05)
    d = rnd(*at.shape)
    obj = lambda aa: O.matching_objective(v, f, aa, Bs, [1.5], [0.8], regweight=0.5)[0]
    fd = _fd(obj, at, d)
    # the Euclidean gradient K(x_t)(2 r a_t - p_{t+1}) / T is the exact discrete adjoint
    g = O.matching_gradient(v, f, at, Bs, [1.5], [0.8], regweight=0.5, euclidean=True)
    np.testing.assert_allclose((g * d).sum(), fd, rtol=1e-6)


def test_matching_gradient_current_fd(spheres):
    A, Bs = spheres
    v, f = A
    at = rnd(3, v.shape[0], 3, scale=0.05)
    d = rnd(*at.shape)
    obj = lambda aa: O.matching_objective(v, f, aa, Bs, [1.5], [0.8], kind='current')[0]
    g = O.matching_gradient(v, f, at, Bs, [1.5], [0.8], kind='current', euclidean=True)
    np.testing.assert_allclose((g * d).sum(), _fd(obj, at, d), rtol=1e-6)


# ------------------------------------------------------ geodesic shooting
def test_epdiff_matches_second_order_evolution():
    x0, a0 = rnd(12, 3), rnd(12, 3, scale=0.3)
    xe, ae = O.landmarkEPDiff(8, x0, a0, [1.3])
    xs, as_ = O.secondOrderEvolution(x0, a0, [1.3], 1 / 8)
    np.testing.assert_allclose(xe, xs, rtol=1e-14)
    np.testing.assert_allclose(ae, as_[:8], rtol=1e-14)


def test_hamiltonian_drift_shrinks_with_dt():
    x0, a0 = rnd(12, 3), rnd(12, 3, scale=0.4)
    H = lambda x, a: 0.5 * (a * O.applyK(x, x, a, [1.3])).sum()
    drift = []
    for T in (10, 40):
        xt, at = O.secondOrderEvolution(x0, a0, [1.3], 1 / T)
        drift.append(abs(H(xt[-1], at[-1]) - H(x0, a0)))
    assert drift[1] < drift[0] / 2.5   # first-order scheme: O(dt)


def test_second_order_gradient_fd():
    x0, a0 = rnd(10, 3), rnd(10, 3, scale=0.3)
    tgt = rnd(10, 3)
    T = 5
    # J(a0) = 1/2 a0.K a0 + |x_T - tgt|^2 ; px1 = -2(x_T - tgt), pa1 = 0
    def J(a):
        xt, at = O.secondOrderEvolution(x0, a, [1.2], 1 / T)
        return 0.5 * (a * O.applyK(x0, x0, a, [1.2])).sum() + ((xt[-1] - tgt) ** 2).sum()
    xt, _ = O.secondOrderEvolution(x0, a0, [1.2], 1 / T)
    px1 = -2 * (xt[-1] - tgt)
    da0 = O.secondOrderGradient(x0, a0, px1, np.zeros_like(a0), [1.2], 1 / T)[0]
    d = rnd(10, 3)
    np.testing.assert_allclose((da0 * d).sum(), _fd(J, a0, d), rtol=1e-6)


@pytest.mark.parametrize("aff", ["affine", "euclidean"])
@pytest.mark.parametrize("dim", [3, 2])
def test_second_order_gradient_affine_kernel_fd(aff, dim):
    """The geodesic adjoint with the kernel's affine part (Kernel.applyK,
    applyDiffKT, applyDDiffK11and12 and applyDiffK1and2 each add theirs):
    J(a0) = 1/2 a0.K a0 + |x_T - tgt|^2 + w.a_T, so px1 = -2 (x_T - tgt) and
    pa1 = -w (every adjoint term is exercised); da0 = FD of J."""
    x0, a0 = rnd(10, dim), rnd(10, dim, scale=0.3)
    tgt, w = rnd(10, dim), rnd(10, dim)
    kaff = (aff, 0.7, 0.4, np.full(dim, 0.2))
    T = 5

    def J(a):
        xt, at = O.secondOrderEvolution(x0, a, [1.2], 1 / T, kaff=kaff)
        Ka = O.applyK(x0, x0, a, [1.2]) + O.affineApplyK(x0, x0, a, kaff)
        return 0.5 * (a * Ka).sum() + ((xt[-1] - tgt) ** 2).sum() + (w * at[-1]).sum()
    xt, at = O.secondOrderEvolution(x0, a0, [1.2], 1 / T, kaff=kaff)
    da0 = O.secondOrderGradient(x0, a0, -2 * (xt[-1] - tgt), -w, [1.2], 1 / T, kaff=kaff)[0]
    d = rnd(10, dim)
    np.testing.assert_allclose((da0 * d).sum(), _fd(J, a0, d), rtol=1e-6)
    # the affine part moves the trajectory
    xt0, _ = O.secondOrderEvolution(x0, a0, [1.2], 1 / T)
    assert np.abs(xt[-1] - xt0[-1]).max() > 1e-3


@pytest.mark.parametrize("aff", ["affine", "euclidean"])
def test_second_order_matching_affine_kernel_gradient_fd(aff):
    """Second-order point-set matching with the kernel's affine part
    (affineKernel): da0 = FD of the objective in a0."""
    x0 = rnd(20, 3) * 2
    w0 = np.abs(rnd(20)) + 0.2
    tgt = (rnd(18, 3) * 2 + 0.3, np.abs(rnd(18)) + 0.2)
    a0 = 0.1 * rnd(20, 3)
    d = rnd(20, 3)
    kaff = (aff, 0.5, 0.3, x0.mean(0))
    kw = dict(kind='point_measure', sigma_error=0.7, kaff=kaff)
    g = O.second_order_matching_gradient(x0, w0, a0, tgt, [1.5], [0.9], 0.2, **kw)
    fd = _fd(lambda a_: O.second_order_matching_objective(x0, w0, a_, tgt, [1.5], [0.9], 0.2, **kw)[0], a0, d)
    np.testing.assert_allclose((g * d).sum(), fd, rtol=1e-6)


@pytest.mark.parametrize("kind", ["point_measure", "L2"])
def test_second_order_matching_gradient_fd(kind):
    """Second-order point-set matching: da0 = FD of the objective in a0."""
    x0 = rnd(30, 3) * 2
    w0 = np.abs(rnd(30)) + 0.2
    if kind == 'L2':
        tgt = x0 + 0.3 * rnd(30, 3)
    else:
        tgt = (rnd(25, 3) * 2 + 0.3, np.abs(rnd(25)) + 0.2)
    a0 = 0.1 * rnd(30, 3)
    d = rnd(30, 3)
    g = O.second_order_matching_gradient(x0, w0, a0, tgt, [1.5], [0.9], 0.2, kind=kind, sigma_error=0.7)
    fd = _fd(lambda a_: O.second_order_matching_objective(x0, w0, a_, tgt, [1.5], [0.9], 0.2, kind=kind,
                                                          sigma_error=0.7)[0], a0, d)
    np.testing.assert_allclose((g * d).sum(), fd, rtol=1e-6)


@pytest.mark.parametrize("kind", ["varifold", "current", "measure"])
def test_second_order_surface_matching_gradient_fd(spheres, kind):
    """Second-order surface matching (base/secondOrderSurfaceMatching.py with the
    surface data terms): da0 = FD of the objective in a0 (the template's
    faces take the place of the point weights)."""
    (v, f), (v2, f2) = spheres
    a0 = 0.05 * rnd(*v.shape)
    d = rnd(*v.shape)
    kw = dict(kind=kind, sigma_error=0.7)
    g = O.second_order_matching_gradient(v, f, a0, (v2, f2), [1.5], [0.8], 0.2, **kw)
    fd = _fd(lambda a_: O.second_order_matching_objective(v, f, a_, (v2, f2), [1.5], [0.8], 0.2, **kw)[0],
             a0, d)
    np.testing.assert_allclose((g * d).sum(), fd, rtol=1e-6)


# ------------------------------------------------------ laplacian family
def _matern(order, r):
    """Matern nu = order + 1/2 without its e^-r: order!/(2 order)! sum_i
    (order + i)! / (i! (order - i)!) (2 r)^(order - i)."""
    from math import factorial as f
    k = order
    return f(k) / f(2 * k) * sum(f(k + i) / (f(i) * f(k - i)) * (2 * r) ** (k - i) for i in range(k + 1))


@pytest.mark.parametrize("order", [0, 1, 2, 3, 4])
def test_laplacian_profile_is_matern(order):
    kern = ('laplacian', order)
    y = np.zeros((1, 3))
    u = rnd(1, 3)
    u /= np.linalg.norm(u)
    a = np.array([[1.5, -0.5]])
    for r in (0.0, 0.3, 1.7, 4.2):
        x = u * r * 0.8
        got = O.applyK(y, x, a, [0.8], kernel=kern)
        np.testing.assert_allclose(got, _matern(order, r) * np.exp(-r) * a, rtol=1e-13, atol=1e-300)
    # two scales: weights s^-1 / sum s^-1
    x = u * 1.1
    w = np.array([1 / 0.8, 1 / 2.0]) / (1 / 0.8 + 1 / 2.0)
    expect = sum(wi * _matern(order, 1.1 / s) * np.exp(-1.1 / s) for wi, s in zip(w, (0.8, 2.0)))
    np.testing.assert_allclose(O.applyK(y, x, a, [0.8, 2.0], kernel=kern), expect * a, rtol=1e-13)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("sig", [[1.2], [0.9, 2.1]])
def test_laplacian_diffKT_is_gradient_wrt_targets(order, sig):
    kern = ('laplacian', order)
    y, x, p, a = rnd(15, 3), rnd(20, 3), rnd(15, 3), rnd(20, 3)
    dy = rnd(15, 3)
    fd = _fd(lambda yy: (p * O.applyK(yy, x, a, sig, kernel=kern)).sum(), y, dy)
    g = O.applyDiffKT(y, x, p, a, sig, kernel=kern)
    np.testing.assert_allclose((g * dy).sum(), fd, rtol=1e-7)
    r = 0.6
    a = a[:15]
    f = lambda xx: ((p * O.applyK(xx, xx, a, sig, kernel=kern)).sum()
                    - r * (a * O.applyK(xx, xx, a, sig, kernel=kern)).sum())
    g = O.applyDiffKT(y, y, p, a, sig, regweight=r, lddmm=True, kernel=kern)
    np.testing.assert_allclose((g * dy).sum(), _fd(f, y, dy), rtol=1e-7)


def test_laplacian_order0_derivative_sums_vanish():
    """The reference's first-derivative row for order 0 is zero
    (base/kernelFunctions_util.py:27), so every derivative sum is zero."""
    kern = ('laplacian', 0)
    x, p, a = rnd(10, 3), rnd(10, 3), rnd(10, 3)
    assert np.all(O.applyDiffKT(x, x, p, a, [1.0], kernel=kern) == 0)
    assert np.all(O.applyDiffK(x, p, a, [1.0], option='1and2', kernel=kern) == 0)
    assert np.all(O.applyDDiffK(x, p, a, p, [1.0], kernel=kern) == 0)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_laplacian_diffK_1and2_is_directional_derivative(order):
    kern = ('laplacian', order)
    x, a1, a2 = rnd(16, 3), rnd(16, 3), rnd(16, 2)
    fd = _fd(lambda t: O.applyK(x + t * a1, x + t * a1, a2, [1.1], kernel=kern), 0.0, 1.0)
    np.testing.assert_allclose(O.applyDiffK(x, a1, a2, [1.1], option='1and2', kernel=kern), fd,
                               rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("order", [2, 3, 4])
@pytest.mark.parametrize("sig", [[1.3], [0.9, 2.2]])
def test_laplacian_ddiffK_is_derivative_of_diffKT(order, sig):
    kern = ('laplacian', order)
    x, a1, a2, p = rnd(14, 3), rnd(14, 3), rnd(14, 3), rnd(14, 3)
    fd = _fd(lambda t: O.applyDiffKT(x + t * p, x + t * p, a1, a2, sig, kernel=kern), 0.0, 1.0)
    np.testing.assert_allclose(O.applyDDiffK(x, a1, a2, p, sig, option='11and12', kernel=kern), fd,
                               rtol=1e-6, atol=1e-9)


def test_laplacian_order1_ddiffK_drops_the_second_derivative_term():
    """Order 1: Q = 1, so the exact second-derivative factor is 1/r, which the
    reference's table replaces by 0 (base/kernelFunctions_util.py:34).  The
    finite derivative minus the oracle's sum is then exactly the dropped
    term sum_l (d.q) d / r e^-r (a1_i . a2_l) s^(KP-2) / W (d scaled by 1/s,
    s^(KP-2) / W = s^-2 for one scale)."""
    kern = ('laplacian', 1)
    x, a1, a2, p = rnd(9, 3), rnd(9, 3), rnd(9, 3), rnd(9, 3)
    s = 1.3
    fd = _fd(lambda t: O.applyDiffKT(x + t * p, x + t * p, a1, a2, [s], kernel=kern), 0.0, 1.0)
    got = O.applyDDiffK(x, a1, a2, p, [s], option='11and12', kernel=kern)
    dropped = np.zeros_like(x)
    for i in range(9):
        for l in range(9):
            if i == l:
                continue
            d = (x[i] - x[l]) / s
            q = p[i] - p[l]
            r = np.linalg.norm(d)
            dropped[i] += (d @ q) * d / r * np.exp(-r) * (a1[i] @ a2[l]) * s ** (-2)
    np.testing.assert_allclose(fd - got, dropped, rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("order", [0, 3])
def test_laplacian_dense_helpers_and_tensor(order):
    kern = ('laplacian', order)
    y, x, a, p = rnd(12, 3), rnd(17, 3), rnd(17, 3), rnd(12, 3)
    beta = np.broadcast_to(a[None], (12, 17, 3))
    np.testing.assert_allclose(O.applykmat(y, x, beta, [1.3, 2.0], kernel=kern),
                               O.applyK(y, x, a, [1.3, 2.0], kernel=kern), rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(O.applykdiffmat(y, x, p @ a.T, [1.3], kernel=kern),
                               O.applyDiffKT(y, x, p, a, [1.3], kernel=kern), rtol=1e-12, atol=1e-300)
    y1, x1 = np.zeros((1, 3)), np.array([[0.3, 0.2, -0.1]])
    by, bx = np.array([[1.0, 2.0, 0.5]]), np.array([[-0.5, 1.0, 2.0]])
    r = np.linalg.norm(x1) / 0.8
    expect = _matern(order, r) * np.exp(-r) * (1.5 * 0.7 + (by * bx).sum() ** 2)
    np.testing.assert_allclose(O.applyktensor(y1, x1, [1.5], [0.7], by, bx, [0.8], kernel=kern), [expect],
                               rtol=1e-13)


def test_laplacian_matching_gradient_fd(spheres):
    """Surface matching with a laplacian deformation kernel (the reference's
    examples pass Kernel(name='laplacian') as KparDiff): the Euclidean
    gradient through the discrete adjoint matches finite differences."""
    A, Bs = spheres
    v, f = A
    kern = ('laplacian', 3)
    at = rnd(3, v.shape[0], 3, scale=0.05)
    d = rnd(*at.shape)
    obj = lambda aa: O.matching_objective(v, f, aa, Bs, [1.5], [0.8], regweight=0.5, kernel_diff=kern)[0]
    g = O.matching_gradient(v, f, at, Bs, [1.5], [0.8], regweight=0.5, euclidean=True, kernel_diff=kern)
    np.testing.assert_allclose((g * d).sum(), _fd(obj, at, d), rtol=1e-6)


@pytest.mark.parametrize("order", [2, 4])
def test_laplacian_second_order_gradient_fd(order):
    """Geodesic (EPDiff) shooting with a laplacian kernel (orders >= 2, where
    every table row is the true derivative): J(a0) = 1/2 a0.K a0 + |x_T - tgt|^2."""
    kern = ('laplacian', order)
    x0, a0, tgt = rnd(10, 3), rnd(10, 3, scale=0.3), rnd(10, 3)
    T = 4

    def J(a):
        xt, _ = O.secondOrderEvolution(x0, a, [1.4], 1 / T, kernel=kern)
        return 0.5 * (a * O.applyK(x0, x0, a, [1.4], kernel=kern)).sum() + ((xt[-1] - tgt) ** 2).sum()
    xt, _ = O.secondOrderEvolution(x0, a0, [1.4], 1 / T, kernel=kern)
    da0 = O.secondOrderGradient(x0, a0, -2 * (xt[-1] - tgt), np.zeros_like(a0), [1.4], 1 / T,
                                kernel=kern)[0]
    d = rnd(10, 3)
    np.testing.assert_allclose((da0 * d).sum(), _fd(J, a0, d), rtol=1e-6)


# ------------------------------------------------------- affine kernel part
KAFF = [('affine', 0.7, 0.4, np.array([0.3, -0.2, 0.5])), ('euclidean', 1.3, 0.6, np.array([0.1, 0.4, -0.3]))]


@pytest.mark.parametrize("kaff", KAFF, ids=['affine', 'euclidean'])
def test_affine_diffKT_modes_are_gradients(kaff):
    """plain: d/dy <p, A(y, x) a>; sym: d/dx <p, A(x, x) a>; lddmm: d/dx of
    <p, A(x, x) a> - r <a, A(x, x) a>, A the affine part (all by finite
    differences)."""
    x, p, a, dx = rnd(14, 3), rnd(14, 3), rnd(14, 3), rnd(14, 3)
    y = rnd(14, 3)
    fd = _fd(lambda yy: (p * O.affineApplyK(yy, x, a, kaff)).sum(), y, dx)
    # plain mode: the reference evaluates xx at the sources only, so the
    # gradient wrt targets is p_i (a^T xx) (independent of y)
    np.testing.assert_allclose((O.affineApplyDiffKT(x, p, a, kaff) * dx).sum(), fd, rtol=1e-7)
    fd = _fd(lambda xx: (p * O.affineApplyK(xx, xx, a, kaff)).sum(), x, dx)
    np.testing.assert_allclose((O.affineApplyDiffKT(x, p, a, kaff, sym=True) * dx).sum(), fd, rtol=1e-7)
    r = 0.35
    f = lambda xx: (p * O.affineApplyK(xx, xx, a, kaff)).sum() - r * (a * O.affineApplyK(xx, xx, a, kaff)).sum()
    np.testing.assert_allclose((O.affineApplyDiffKT(x, p, a, kaff, regweight=r, lddmm=True) * dx).sum(),
                               _fd(f, x, dx), rtol=1e-7)


def test_affine_parts_closed_forms():
    """'affine' is linear in y - c plus the column sum; 'euclidean' is an
    infinitesimal rotation (antisymmetric linear map) plus the column sum."""
    x, a, y = rnd(20, 3), rnd(20, 3), rnd(9, 3)
    c = np.array([0.2, -0.1, 0.3])
    za = O.affineApplyK(y, x, a, ('affine', 1.0, 0.0, c))
    Mf = (x - c).T @ a
    np.testing.assert_allclose(za, (y - c) @ Mf, rtol=1e-13)
    ze = O.affineApplyK(y, x, a, ('euclidean', 1.0, 0.0, c))
    L = np.linalg.lstsq(y - c, ze, rcond=None)[0]       # ze = (y - c) L
    np.testing.assert_allclose(L, -L.T, atol=1e-12)
    np.testing.assert_allclose(L, (Mf - Mf.T) / 2, atol=1e-12)
    zt = O.affineApplyK(y, x, a, ('euclidean', 0.0, 2.5, c))
    np.testing.assert_allclose(zt, np.broadcast_to(2.5 * a.sum(0), zt.shape), rtol=1e-14)
    # 2-D Euclidean basis: one rotation generator
    x2, a2, y2 = rnd(10, 2), rnd(10, 2), rnd(5, 2)
    z2 = O.affineApplyK(y2, x2, a2, ('euclidean', 1.0, 0.0, np.zeros(2)))
    M2 = x2.T @ a2
    np.testing.assert_allclose(z2, y2 @ ((M2 - M2.T) / 2), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("kaff", KAFF, ids=['affine', 'euclidean'])
def test_affine_kernel_matching_gradient_fd(spheres, kaff):
    A, Bs = spheres
    v, f = A
    at = rnd(3, v.shape[0], 3, scale=0.05)
    d = rnd(*at.shape)
    kaff = (kaff[0], kaff[1] * 1e-3, kaff[2] * 1e-3, kaff[3])
    obj = lambda aa: O.matching_objective(v, f, aa, Bs, [1.5], [0.8], regweight=0.5, kaff=kaff)[0]
    g = O.matching_gradient(v, f, at, Bs, [1.5], [0.8], regweight=0.5, euclidean=True, kaff=kaff)
    np.testing.assert_allclose((g * d).sum(), _fd(obj, at, d), rtol=1e-6)


# ------------------------------------------------------------ internal costs
@pytest.mark.parametrize("kind", ["h1", "elastic", "displacement"])
def test_internal_cost_derivatives_fd(spheres, kind):
    A, _ = spheres
    v, f = A
    v = v + rnd(*v.shape, scale=0.05)              # irregular triangles
    phi = rnd(*v.shape)
    dphi, dx = rnd(*v.shape), rnd(*v.shape)
    c, gphi, gx = O.internal_cost(kind, v, f, phi)
    fd_phi = _fd(lambda pp: O.internal_cost(kind, v, f, pp)[0], phi, dphi)
    fd_x = _fd(lambda xx: O.internal_cost(kind, xx, f, phi)[0], v, dx)
    np.testing.assert_allclose((gphi * dphi).sum(), fd_phi, rtol=1e-6)
    np.testing.assert_allclose((gx * dx).sum(), fd_x, rtol=1e-6)


def test_internal_cost_closed_forms():
    v = np.array([[0., 0, 0], [1, 0, 0], [0, 1, 0], [0.3, 0.2, 0.9]])
    f = np.array([[0, 1, 2], [0, 1, 3]])
    area = 0.5 + 0.5 * np.linalg.norm(np.cross(v[1] - v[0], v[3] - v[0]))
    # h1 of a linear field phi = x G^T is sum over faces of area * |G P_f|^2
    G = rnd(3, 3)
    expect = 0.0
    for tri in f:
        e1, e2 = v[tri[1]] - v[tri[0]], v[tri[2]] - v[tri[0]]
        Q, _ = np.linalg.qr(np.stack([e1, e2], 1))          # orthonormal basis of the face plane
        P = Q @ Q.T
        expect += 0.5 * np.linalg.norm(np.cross(e1, e2)) * np.linalg.norm(G @ P) ** 2
    np.testing.assert_allclose(O.internal_cost('h1', v, f, v @ G.T)[0], expect, rtol=1e-12)
    # elastic vanishes on a rigid field (translation + in-plane rotation) of a flat triangle
    flat = f[:1]
    W = np.array([[0., -1, 0], [1, 0, 0], [0, 0, 0]])     # rotation about z
    rigid = v @ W.T + np.array([0.3, -0.7, 0.0])
    assert abs(O.internal_cost('elastic', v, flat, rigid)[0]) < 1e-14
    assert O.internal_cost('h1', v, flat, rigid)[0] > 0.1      # h1 does not vanish there
    # displacement of a constant field is |c|^2 * area
    cst = np.broadcast_to(np.array([1.0, 2.0, -0.5]), v.shape)
    np.testing.assert_allclose(O.internal_cost('displacement', v, f, cst)[0], 5.25 * area, rtol=1e-13)


def test_internal_cost_matching_gradient_fd(spheres):
    A, Bs = spheres
    v, f = A
    at = rnd(3, v.shape[0], 3, scale=0.05)
    d = rnd(*at.shape)
    ic = [('elastic', 5.0), ('h1', 0.5), ('displacement', 0.2)]
    obj = lambda aa: O.matching_objective(v, f, aa, Bs, [1.5], [0.8], regweight=0.5, icost=ic)[0]
    g = O.matching_gradient(v, f, at, Bs, [1.5], [0.8], regweight=0.5, euclidean=True, icost=ic)
    np.testing.assert_allclose((g * d).sum(), _fd(obj, at, d), rtol=1e-6)


# ------------------------------------------------------------ unreduced
def test_semi_reduced_flow_reduces_to_landmark_flow():
    """With c_t equal to the trajectory itself the semi-reduced flow is the
    reduced one (one step at a time, since c_t is given up front)."""
    x0, at = rnd(12, 3), rnd(1, 12, 3, scale=0.3)
    xr = O.landmarkDirectEvolutionEuler(x0, at, [1.3])
    xs = O.semiReducedEvolution(x0, x0[None], at, [1.3], fidelity_weight=0.7)
    np.testing.assert_allclose(xs, xr, rtol=1e-14)


@pytest.mark.parametrize("variant", ["plain", "fidelity+costs+affine"])
def test_unreduced_matching_gradient_fd(spheres, variant):
    A, Bs = spheres
    v, f = A
    T = 3
    ct = np.broadcast_to(v, (T,) + v.shape) + rnd(T, *v.shape, scale=0.1)
    at = rnd(T, v.shape[0], 3, scale=0.05)
    kw = dict(regweight=0.5, unreduced_weight=0.8)
    if variant != "plain":
        kw.update(fidelity_weight=0.3, icost=[('elastic', 2.0), ('displacement', 0.1)],
                  kaff=('euclidean', 1e-3, 2e-3, v.mean(0)), kernel_diff=('laplacian', 3))
    obj = lambda cc, aa: O.matching_objective_unreduced(v, f, cc, aa, Bs, [1.5], [0.8], **kw)[0]
    gc, ga = O.matching_gradient_unreduced(v, f, ct, at, Bs, [1.5], [0.8], **kw)
    dc, da = rnd(*ct.shape), rnd(*at.shape)
    np.testing.assert_allclose((ga * da).sum(), _fd(lambda aa: obj(ct, aa), at, da), rtol=1e-6)
    np.testing.assert_allclose((gc * dc).sum(), _fd(lambda cc: obj(cc, at), ct, dc), rtol=1e-6)


# ------------------------------------------------------------- divergence
def _fd_div(f, y, eps=1e-6):
    """Trace of the central-difference Jacobian of f: [N, d] -> [N, d]."""
    n, dim = y.shape
    div = np.zeros(n)
    for k in range(dim):
        e = np.zeros((n, dim))
        e[:, k] = 1.0
        div += ((f(y + eps * e) - f(y - eps * e)) / (2 * eps))[:, k]
    return div


@pytest.mark.parametrize("kern,sig,dim", [(O.GAUSS, [1.2], 3), (O.GAUSS, [0.9, 2.1], 3), (O.GAUSS, [1.0], 2),
                                          (('laplacian', 1), [1.1], 3), (('laplacian', 3), [0.9, 2.0], 3),
                                          (('laplacian', 4), [1.3], 2)])
def test_applyDiv_is_trace_of_the_jacobian_of_applyK(kern, sig, dim):
    rng = np.random.default_rng(len(sig) + 10 * dim)
    y, x, a = rng.normal(size=(12, dim)), rng.normal(size=(17, dim)), rng.normal(size=(17, dim))
    fd = _fd_div(lambda yy: O.applyK(yy, x, a, sig, kernel=kern), y)
    got = O.applyDiv(y, x, a, sig, kernel=kern)
    assert got.shape == (12, 1)
    # the sum cancels (terms ~1, divergence ~1e-3): the difference quotient's
    # rounding (~1e-9) sets the absolute floor; a wrong term or power is ~1e-2
    np.testing.assert_allclose(got[:, 0], fd, rtol=1e-7, atol=1e-8)


def test_applyDiv_laplacian_order0_vanishes():
    """Order 0 has a zero first-derivative row in the reference's table
    (base/kernelFunctions_util.py:27): the divergence sum is zero."""
    x, a = rnd(9, 3), rnd(9, 3)
    assert np.all(O.applyDiv(x, x, a, [1.0], kernel=('laplacian', 0)) == 0)


@pytest.mark.parametrize("kaff", KAFF, ids=['affine', 'euclidean'])
def test_affine_divergence_is_trace_of_the_jacobian(kaff):
    rng = np.random.default_rng(77)
    y, x, a = rng.normal(size=(11, 3)), rng.normal(size=(11, 3)), rng.normal(size=(11, 3))
    fd = _fd_div(lambda yy: O.affineApplyK(yy, x, a, kaff), y)
    got = O.affineApplyDiv(x, a, kaff)
    np.testing.assert_allclose(np.full(11, got), fd, rtol=1e-7, atol=1e-9)
    if kaff[0] == 'euclidean':          # rotations preserve volume
        assert got == 0


def _flow_logdet(x0, at, sig, kaff, h=1e-5):
    """log det of the central-difference Jacobian of the Euler flow map at x0
    (the perturbed copies ride along as passenger points)."""
    xt = O.landmarkDirectEvolutionEuler(x0, at, sig, kaff=kaff)
    T = at.shape[0]
    n, dim = x0.shape
    ys = np.concatenate([x0 + s * h * np.eye(dim)[k] for k in range(dim) for s in (1, -1)])
    for t in range(T):
        v = O.applyK(ys, xt[t], at[t], sig)
        if kaff is not None:
            v = v + O.affineApplyK(ys, xt[t], at[t], kaff)
        ys = ys + v / T
    ys = ys.reshape(dim, 2, n, dim)
    D = np.stack([(ys[k, 0] - ys[k, 1]) / (2 * h) for k in range(dim)], axis=-1)   # [n, dim, dim]
    return np.log(np.linalg.det(D))


@pytest.mark.parametrize("kaff", [None, KAFF[0]], ids=['none', 'affine'])
def test_jacobian_tracks_log_det_of_the_flow(kaff):
    """J_T = sum_t dt div v_t(x_t) differs from log det D(phi_T) by
    sum_t [log det(I + dt Dv_t) - dt tr Dv_t] = O(dt): halving dt halves it."""
    rng = np.random.default_rng(78)
    x0 = rng.normal(size=(7, 3))
    # the affine field grows with x_t: a smaller momentum keeps the flow bounded
    a = (0.4 if kaff is None else 0.06) * rng.normal(size=(7, 3))
    errs = []
    for T in (20, 40, 80):
        at = np.repeat(a[None], T, axis=0)
        xt = O.landmarkDirectEvolutionEuler(x0, at, [1.3], kaff=kaff)
        J = O.jacobian(xt, at, [1.3], kaff=kaff)
        assert J.shape == (T + 1, 7, 1) and np.all(J[0] == 0)
        ld = _flow_logdet(x0, at, [1.3], kaff)
        errs.append(np.abs(J[-1, :, 0] - ld).max())
        assert errs[-1] < 0.05 * np.abs(ld).max()
    assert 1.7 < errs[0] / errs[1] < 2.3 and 1.7 < errs[1] / errs[2] < 2.3


# ------------------------------------------- C restatement (CPU baseline)
@pytest.mark.parametrize("sig", [[1.3], [0.8, 2.5]])
def test_cpu_reference_matches_numpy_oracle(sig):
    """oracle/cpu_ref.c (the timed CPU baseline) against the numpy oracle:
    N != M, one and two scales, the three applyDiffKT modes, k = 3 and 1."""
    from oracle import cpu
    r = np.random.default_rng(12)
    y, x = r.normal(size=(37, 3)) * 2, r.normal(size=(53, 3)) * 2
    a, p = r.normal(size=(53, 3)), r.normal(size=(37, 3))
    np.testing.assert_allclose(cpu.applyK(y, x, a, sig), O.applyK(y, x, a, sig), rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(cpu.applyK(y, x, a[:, :1], sig), O.applyK(y, x, a[:, :1], sig),
                               rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(cpu.applyDiffKT(y, x, p, a, sig), O.applyDiffKT(y, x, p, a, sig),
                               rtol=1e-13, atol=1e-14)
    xs, ps, as_ = r.normal(size=(41, 3)), r.normal(size=(41, 3)), r.normal(size=(41, 3))
    for kw in ({'sym': True}, {'lddmm': True, 'regweight': 0.6}):
        np.testing.assert_allclose(cpu.applyDiffKT(xs, xs, ps, as_, sig, **kw),
                                   O.applyDiffKT(xs, xs, ps, as_, sig, **kw), rtol=1e-13, atol=1e-14)


# ------------------------------------------------- Kernel.solve (linearcg)
def _gram(x, s):
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    return np.exp(-d2 / (2 * s * s))


def _solve_problem(n=40, seed=5):
    r = np.random.default_rng(seed)
    x = r.uniform(0, 6, size=(n, 3))
    return x, r.normal(size=(n, 3)), _gram(x, 1.0)


def test_linearcg_textbook_iterates_minimise_over_krylov_space():
    """CG's k-th iterate minimises a.Ga/2 - b.a over span{b, Gb, .., G^(k-1) b}
    (whole [n, 3] arrays as one vector); computed here by a dense projection."""
    x, b, G = _solve_problem()
    V = [b]
    for _ in range(5):
        V.append(G @ V[-1])
    for k in range(1, 6):
        B = np.stack([v.ravel() for v in V[:k]], 1)
        GB = np.stack([(G @ v).ravel() for v in V[:k]], 1)
        c = np.linalg.solve(B.T @ GB, B.T @ b.ravel())
        expect = (B @ c).reshape(b.shape)
        got, products = O.kernel_solve(x, b, [1.0], iter_max=k, sequence='textbook')
        assert products == k
        np.testing.assert_allclose(got, expect, rtol=1e-8, atol=1e-10 * np.abs(expect).max())


def test_linearcg_converged_solve_is_dense_solve():
    x, b, G = _solve_problem(n=30, seed=6)
    expect = np.linalg.solve(G, b)
    for seq in ('textbook', 'reference'):
        got, _ = O.kernel_solve(x, b, [1.0], iter_max=3000, sequence=seq)
        assert np.linalg.norm(G @ got - b) / np.linalg.norm(b) < 1e-7, seq
        assert np.linalg.norm(got - expect) / np.linalg.norm(expect) < 1e-4, seq


def _aliased_cg(op, b, iters):
    """The reference loop's statements with numpy's own semantics: the first
    direction is bound to the residual array and both are updated in place."""
    sol = np.zeros(b.shape)
    zz = op(sol) - 2 * b
    res = -zz - b
    e_prev = (zz * sol).sum()
    m = 0.
    d = None
    for k in range(iters):
        if k == 0:
            m = (res * res).sum()
            if m < 1e-10:
                return sol
            d = res
        else:
            m_prev = m
            m = (res * res).sum()
            d = res + (m / m_prev) * d
        qq = op(d)
        step = m / (d * qq).sum()
        sol += step * d
        res -= step * qq
        zz += step * qq
        e = (zz * sol).sum()
        if step * np.mean(np.fabs(d)) < 1e-10:
            return sol
        if np.fabs(e - e_prev) < 1e-20 * np.fabs(e_prev):
            return sol
        e_prev = e
    return sol


@pytest.mark.parametrize("iters", [1, 2, 3, 10, 60])
def test_linearcg_reference_sequence_follows_the_aliased_loop(iters):
    x, b, G = _solve_problem(n=50, seed=7)
    got, _ = O.kernel_solve(x, b, [1.0], iter_max=iters, sequence='reference')
    expect = _aliased_cg(lambda v: G @ v, b, iters)
    np.testing.assert_allclose(got, expect, rtol=1e-9, atol=1e-12 * np.abs(expect).max())
    tb, _ = O.kernel_solve(x, b, [1.0], iter_max=iters, sequence='textbook')
    if iters == 1:
        np.testing.assert_allclose(got, tb, rtol=1e-13)        # first step: steepest descent in both
    else:
        assert np.linalg.norm(got - tb) > 1e-3 * np.linalg.norm(tb)


def test_linearcg_zero_rhs_and_stop_counts():
    x, b, G = _solve_problem(n=20, seed=8)
    for seq in ('textbook', 'reference'):
        a, products = O.kernel_solve(x, 1e-7 * b, [1.0], iter_max=10, sequence=seq)
        assert products == 0 and np.all(a == 0)
    # the energy test stops a converged textbook solve well before the cap
    _, products = O.kernel_solve(x, b, [1.0], iter_max=5000, sequence='textbook')
    assert products < 5000
