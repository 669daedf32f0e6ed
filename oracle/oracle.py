"""FP64 numpy oracle for the py-lddmm point-set hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package (`lddmm_amd`)
imports this module; only `tests/`, `__graft_entry__.smoke()` and the
`cpu_baseline` leg of `bench.py` use it, always as the checker.

Each function restates one reference function (paths relative to
`py-lddmm/`), in float64, evaluated over blocks of target rows so that the
O(N*M) pair sums never materialise more than `block * M` pairs at once.
Index conventions follow the reference: `y` are targets (rows, N), `x` are
sources (columns, M), weights `a` are [M, k], scales are an array of sigmas
and the multi-scale weight is s**KP / sum(s**KP) with KP = -1.

How each function is pinned (tests/test_oracle.py):
  applyK          closed forms (one point, two points), sigma->inf limit,
                  duplicated-sigma identity, adjoint symmetry
  applyDiffKT     central finite differences of <p, K(y,x) a> (plain),
                  of p^T K(x) a (sym) and p^T K a - r a^T K a (lddmm)
  applyDiv        trace of the central-difference Jacobian (in y) of applyK;
                  the affine part likewise from affineApplyK
  jacobian        log det of the finite-difference Jacobian of the Euler flow
                  map carried alongside, error shrinking like dt
  applyktensor    single pair closed form, adjoint symmetry, rigid invariance
  applykmat       reduces to applyK when beta[i,j,:] = a[j,:]
  applykdiffmat   reduces to applyDiffKT(plain) when beta[i,j] = p_i.a_j
  applyDiffK      finite differences of applyK along the point set
  applyDDiffK     finite differences of applyDiffKT(plain) along the point set
  linearcg        'textbook': the k-th iterate minimises the quadratic energy
                  over the Krylov space of b (dense projection), and the
                  converged solve equals a dense LU solve; 'reference': the
                  reference loop's own statements run with numpy's in-place
                  updates on the aliased direction/residual arrays
  surface / curve varifold and current norms: ||S - S|| = 0, rigid-motion
                  invariance, face-flip invariance (varifold), single pair
                  closed forms, finite-difference gradients
  shooting        zero control -> identity, finite differences of the full
                  objective through the discrete adjoint
  EPDiff          trajectory identity with landmarkEPDiff, Hamiltonian drift
                  O(dt), finite differences of the second-order gradient
  affine parts    finite differences of <p, affine K(y, x) a> for the plain,
                  sym and lddmm covectors (both kinds); the Euclidean part is
                  a rotation field (antisymmetric linear map) plus a
                  translation; shooting with them: FD of the matching objective
  internal costs  finite differences of each cost in phi and in the vertices
                  (h1, elastic, displacement); h1 of a linear field on a
                  triangle = area * |gradient|^2; elastic vanishes on rigid
                  (translation + in-plane rotation) fields; displacement of a
                  constant field = |c|^2 * area; FD of the matching objective
                  with internal costs through the extra-term adjoint
  unreduced       FD of the unreduced objective in both controls (ct, at),
                  with fidelity, internal costs and affine kernel parts;
                  c_t = x_t reduces the flow to landmarkDirectEvolutionEuler
  laplacian       (every kernel sum above takes kernel=('laplacian', order))
                  profile = the Matern nu = order + 1/2 closed form written
                  with factorials; first-derivative sums = finite differences
                  of applyK for orders 1-4, second-derivative sums = finite
                  differences of applyDiffKT for orders 2-4; the reference's
                  table rows that are not derivatives (Q_0 = 0; R_0 = R_1 = 0)
                  are pinned by the zero they produce (order 0) and by the
                  difference from the finite derivative being exactly the
                  dropped R term (order 1)
"""
import numpy as np

__all__ = [
    "applyK", "applyDiffKT", "applyktensor", "applykmat", "applykdiffmat",
    "applyDiffK", "applyDDiffK", "applyDiv", "affineApplyK", "affineApplyDiffKT", "affineApplyDiv",
    "jacobian", "internal_cost", "internal_costs",
    "semiReducedEvolution", "matching_objective_unreduced", "matching_gradient_unreduced",
    "surface_geometry", "surface_volume",
    "varifoldNorm0", "varifoldNormDef", "varifoldNorm", "varifoldNormGradient",
    "currentNorm0", "currentNormDef", "currentNorm", "currentNormGradient",
    "currentMagnitude", "currentMagnitudeGradient",
    "curve_geometry", "curveVarifoldNorm0", "curveVarifoldNormDef",
    "curveVarifoldNormGradient", "curveCurrentNorm0", "curveCurrentNormDef",
    "curveCurrentNormGradient",
    "landmarkDirectEvolutionEuler", "landmarkHamiltonianCovector",
    "landmarkHamiltonianGradient", "matching_objective", "matching_gradient",
    "landmarkEPDiff", "secondOrderEvolution", "secondOrderHamiltonian",
    "secondOrderCovector", "secondOrderGradient", "linearcg", "kernel_solve",
]

_BLOCK_PAIRS = 1 << 22   # pairs held in memory per block (~100 MB of f64 temporaries)


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _scales(scale, KP):
    s = np.atleast_1d(np.asarray(scale, dtype=np.float64))
    sKP = s ** KP
    return s, sKP, sKP.sum()


def _row_blocks(n, m):
    bs = max(1, min(n, _BLOCK_PAIRS // max(m, 1)))
    for i0 in range(0, n, bs):
        yield i0, min(n, i0 + bs)


def _as2d(a):
    a = _f64(a)
    return a[:, None] if a.ndim == 1 else a


# ---------------------------------------------------------------------------
# Radial profiles.  Gaussian: K = exp(-rho^2/2) and both derivative factors
# equal K.  Laplacian (Matern) of order k (base/kernelFunctions_util.py:20-52,
# 92-95, 197-201, 238-243, 309-314): with rho = |d| (scaled distance),
# K = P_k(rho) e^-rho, first-derivative factor Q_k(rho) e^-rho, second
# R_k(rho) e^-rho, where Q = (P - P')/rho and R = (Q - Q')/rho for the orders
# where these are polynomials (the reference's tables set Q_0 = 0, R_0 = R_1 = 0).
# ---------------------------------------------------------------------------
LAP_P = [[1, 0, 0, 0, 0], [1, 1, 0, 0, 0], [1, 1, 1 / 3, 0, 0], [1, 1, 0.4, 1 / 15, 0],
         [1, 1, 3 / 7, 2 / 21, 1 / 105]]
LAP_Q = [[0, 0, 0, 0], [1, 0, 0, 0], [1 / 3, 1 / 3, 0, 0], [1 / 5, 1 / 5, 1 / 15, 0],
         [1 / 7, 1 / 7, 2 / 35, 1 / 105]]
LAP_R = [[0, 0, 0], [0, 0, 0], [1 / 3, 0, 0], [1 / 15, 1 / 15, 0], [1 / 35, 1 / 35, 1 / 105]]


def _poly(c, u):
    return sum(ci * u ** i for i, ci in enumerate(c))


def _profile(d, kernel, which):
    """which 0: K, 1: first-derivative factor, 2: second-derivative factor, for
    scaled differences d [..., dim]; kernel = ('gauss', _) or ('laplacian', order)."""
    name, order = kernel
    if 'gauss' in name:
        return np.exp(-(d ** 2).sum(-1) / 2)
    if 'lap' in name:
        r = np.sqrt((d ** 2).sum(-1))
        c = (LAP_P, LAP_Q, LAP_R)[which][order]
        return _poly(c, r) * np.exp(-r)
    raise ValueError(name)


GAUSS = ('gauss', 0)


# ---------------------------------------------------------------------------
# L1 kernel sums
# ---------------------------------------------------------------------------

def applyK(y, x, a, scale, KP=-1, kernel=GAUSS):
    """out[i] = sum_s w_s sum_j exp(-|y_i/s - x_j/s|^2 / 2) a[j].

    Follows base/kernelFunctions_util.py:451-466 (applyK_numba) with the
    pair function gauss_fun at :86-89; w_s = s^KP / sum(s^KP) (:454-456, :465).
    """
    y, x, a = _f64(y), _f64(x), _as2d(a)
    s, sKP, W = _scales(scale, KP)
    res = np.zeros((y.shape[0], a.shape[1]))
    for si, sc in enumerate(s):
        ys, xs = y / sc, x / sc
        for i0, i1 in _row_blocks(y.shape[0], x.shape[0]):
            d = ys[i0:i1, None, :] - xs[None, :, :]
            K = _profile(d, kernel, 0)
            res[i0:i1] += (K @ a) * sKP[si]
    return res / W


def applyDiffKT(y, x, p, a, scale, regweight=1., sym=False, lddmm=False, KP=-1, kernel=GAUSS, rows=None):
    """out[i] = sum_s s^(KP-1)/W sum_j -(y_i - x_j)/s * K_s(y_i, x_j) * c_ij.

    c_ij = p_i.a_j (plain), + a_i.p_j (sym), + a_i.p_j - 2 r a_i.a_j (lddmm).
    Follows base/kernelFunctions_util.py:773-793 (applyDiffKT_numba) with
    gauss_fun_diffT at :232-236.  sym / lddmm need y and x of the same size.
    rows (index array) returns only those target rows (y_i, p_i and, for
    sym / lddmm, a_i taken at the rows; every source j kept): out[rows].
    """
    y, x, p, a = _f64(y), _f64(x), _as2d(p), _as2d(a)
    if rows is None:
        yt, pt, at = y, p, a
    else:
        rows = np.asarray(rows)
        yt, pt = y[rows], p[rows]
        at = a[rows] if (sym or lddmm) else None
    s, sKP, W = _scales(scale, KP)
    sKP1 = s ** (KP - 1)
    res = np.zeros(yt.shape)
    for si, sc in enumerate(s):
        ys, xs = yt / sc, x / sc
        for i0, i1 in _row_blocks(yt.shape[0], x.shape[0]):
            d = ys[i0:i1, None, :] - xs[None, :, :]
            K = _profile(d, kernel, 1)
            c = pt[i0:i1] @ a.T
            if lddmm:
                c = c + at[i0:i1] @ p.T - 2 * regweight * (at[i0:i1] @ a.T)
            elif sym:
                c = c + at[i0:i1] @ p.T
            res[i0:i1] += -(d * (K * c)[..., None]).sum(1) * sKP1[si]
    return res / W


def applyDiv(y, x, a, scale, KP=-1, kernel=GAUSS):
    """out[i, 0] = sum_s s^(KP-1)/W sum_j -((y_i - x_j)/s).a_j K_s(y_i, x_j), the
    divergence at y_i of the field K(., x) a.

    Follows base/kernelFunctions_util.py:606-620 (applyDiv_numba) with
    gauss_fun_diff / lap_fun_diff at :191-201.
    """
    y, x, a = _f64(y), _f64(x), _as2d(a)
    s, sKP, W = _scales(scale, KP)
    sKP1 = s ** (KP - 1)
    res = np.zeros((y.shape[0], 1))
    for si, sc in enumerate(s):
        ys, xs = y / sc, x / sc
        for i0, i1 in _row_blocks(y.shape[0], x.shape[0]):
            d = ys[i0:i1, None, :] - xs[None, :, :]
            K = _profile(d, kernel, 1)
            res[i0:i1, 0] += -(K * (d * a[None, :, :]).sum(-1)).sum(1) * sKP1[si]
    return res / W


def applyktensor(y, x, ay, ax, betay, betax, scale, KP=-1, kernel=GAUSS):
    """out[i] = sum_s w_s sum_j K_s(y_i, x_j) (ay_i ax_j + (betay_i . betax_j)^2).

    Follows base/kernelFunctions_util.py:1475-1506 (applyktensor_numba, gauss
    branch :1490-1496).
    """
    y, x, ay, ax = _f64(y), _f64(x), _f64(ay).ravel(), _f64(ax).ravel()
    by, bx = _as2d(betay), _as2d(betax)
    s, sKP, W = _scales(scale, KP)
    res = np.zeros(y.shape[0])
    for si, sc in enumerate(s):
        ys, xs = y / sc, x / sc
        for i0, i1 in _row_blocks(y.shape[0], x.shape[0]):
            d = ys[i0:i1, None, :] - xs[None, :, :]
            K = _profile(d, kernel, 0)
            g = by[i0:i1] @ bx.T
            res[i0:i1] += (K * (ay[i0:i1, None] * ax[None, :] + g * g)).sum(1) * sKP[si]
    return res / W


def applykmat(y, x, beta, scale, KP=-1, kernel=GAUSS):
    """out[i] = sum_s w_s sum_j K_s(y_i, x_j) beta[i, j, :] (dense matrix weights).

    Follows base/kernelFunctions_util.py:1509-1540 (applykmat, gauss branch).
    """
    y, x, beta = _f64(y), _f64(x), _f64(beta)
    if beta.ndim == 2:
        beta = beta[..., None]
    s, sKP, W = _scales(scale, KP)
    res = np.zeros((y.shape[0], beta.shape[2]))
    for si, sc in enumerate(s):
        ys, xs = y / sc, x / sc
        for i0, i1 in _row_blocks(y.shape[0], x.shape[0]):
            d = ys[i0:i1, None, :] - xs[None, :, :]
            K = _profile(d, kernel, 0)
            res[i0:i1] += (K[..., None] * beta[i0:i1]).sum(1) * sKP[si]
    return res / W


def applykdiffmat(y, x, beta, scale, KP=-1, kernel=GAUSS):
    """out[i] = -sum_s s^(KP-1)/W sum_j K_s(y_i, x_j) (y_i - x_j)/s beta[i, j].

    Follows base/kernelFunctions_util.py:1638-1669 (applykdiffmat_numba).
    """
    y, x, beta = _f64(y), _f64(x), _f64(beta)
    s, sKP, W = _scales(scale, KP)
    sKP1 = s ** (KP - 1)
    res = np.zeros(y.shape)
    for si, sc in enumerate(s):
        ys, xs = y / sc, x / sc
        for i0, i1 in _row_blocks(y.shape[0], x.shape[0]):
            d = ys[i0:i1, None, :] - xs[None, :, :]
            K = _profile(d, kernel, 1)
            res[i0:i1] -= (d * (K * beta[i0:i1])[..., None]).sum(1) * sKP1[si]
    return res / W


def applyDiffK(x, a1, a2, scale, KP=-1, option='1and2', kernel=GAUSS):
    """First derivative sums on one point set (y = x).

    option '1':     out[k] = sum_l -K (d_kl . a1_k) a2_l
    option '2':     out[k] = -sum_l -K (d_kl . a1_l) a2_l
    option '1and2': out[k] = sum_l -K (d_kl . (a1_k - a1_l)) a2_l
    with d_kl = (x_k - x_l)/s and the scale weight s^(KP-1)/W.
    Follows base/kernelFunctions_util.py:655-680 (applyDiffK_numba) with
    gauss_fun_diff at :190-194.  Kernel.applyDiffK1and2 reaches it through
    applykdiff1and2 (:954-955), which always passes KP = -1.
    """
    x, a1, a2 = _f64(x), _as2d(a1), _as2d(a2)
    s, sKP, W = _scales(scale, KP)
    sKP1 = s ** (KP - 1)
    # the reference allocates x.shape, i.e. a2 has d columns; any width is accepted here
    res = np.zeros((x.shape[0], a2.shape[1]))
    for si, sc in enumerate(s):
        xs = x / sc
        for i0, i1 in _row_blocks(x.shape[0], x.shape[0]):
            d = xs[i0:i1, None, :] - xs[None, :, :]
            K = _profile(d, kernel, 1)
            if option == '1':
                da = (d * a1[i0:i1, None, :]).sum(-1)
                res[i0:i1] += ((-K * da) @ a2) * sKP1[si]
            elif option == '2':
                da = (d * a1[None, :, :]).sum(-1)
                res[i0:i1] -= ((-K * da) @ a2) * sKP1[si]
            elif option == '1and2':
                da = (d * (a1[i0:i1, None, :] - a1[None, :, :])).sum(-1)
                res[i0:i1] += ((-K * da) @ a2) * sKP1[si]
            else:
                raise ValueError(option)
    return res / W


def applyDDiffK(x, a1, a2, p, scale, KP=-1, option='11and12', kernel=GAUSS):
    """Second derivative sums on one point set.

    out[k] = sum_s s^(KP-2)/W sum_l ((d.q) d - q) K_s(x_k, x_l) (a1_k . a2_l)
    with d = (x_k - x_l)/s and q = p_k (option '11'), -p_l ('12'),
    p_k - p_l ('11and12').  Follows base/kernelFunctions_util.py:1092-1129
    (applyDDiffK_numba) with gauss_fun_ddiff at :298-306.
    """
    x, a1, a2, p = _f64(x), _as2d(a1), _as2d(a2), _as2d(p)
    s, sKP, W = _scales(scale, KP)
    res = np.zeros(x.shape)
    for sc in s:
        xs = x / sc
        KS = sc ** (KP - 2)
        for i0, i1 in _row_blocks(x.shape[0], x.shape[0]):
            d = xs[i0:i1, None, :] - xs[None, :, :]
            K = _profile(d, kernel, 1)
            KDD = _profile(d, kernel, 2)
            aa = a1[i0:i1] @ a2.T
            if option == '11':
                q = np.broadcast_to(p[i0:i1, None, :], d.shape)
                sign = 1.
            elif option == '12':
                q = np.broadcast_to(p[None, :, :], d.shape)
                sign = -1.
            elif option == '11and12':
                q = p[i0:i1, None, :] - p[None, :, :]
                sign = 1.
            else:
                raise ValueError(option)
            dq = (d * q).sum(-1)
            t = ((dq * KDD)[..., None] * d - K[..., None] * q) * aa[..., None]
            res[i0:i1] += sign * t.sum(1) * KS
    return res / W


# ---------------------------------------------------------------------------
# Affine part of a kernel (KernelSpec.setAffine, base/kernelFunctions.py:393-412;
# Kernel.applyK :476-505; Kernel.applyDiffKT :601-625).  kaff = (kind, w1, w2,
# center), kind 'affine' or 'euclidean'; the Euclidean basis is the
# reference's list of three (3-D) or one (2-D) antisymmetric matrices / sqrt 2.
# ---------------------------------------------------------------------------

def _euclid_basis(dim):
    s2 = np.sqrt(2.0)
    if dim == 3:
        return [np.array([[0, 1, 0], [-1, 0, 0], [0, 0, 0]]) / s2,
                np.array([[0, 0, 1], [0, 0, 0], [-1, 0, 0]]) / s2,
                np.array([[0, 0, 0], [0, 0, 1], [0, -1, 0]]) / s2]
    return [np.array([[0, 1], [-1, 0]]) / s2]


def affineApplyK(y, x, a, kaff):
    """Affine part of Kernel.applyK(x, a, firstVar=y) (base/kernelFunctions.py:476-505)."""
    kind, w1, w2, c = kaff
    y, x, a = _f64(y), _f64(x), _as2d(a)
    xx, yy = x - c, y - c
    if kind == 'affine':
        return w1 * (yy @ (xx.T @ a)) + w2 * a.sum(axis=0)
    z = np.zeros((y.shape[0], a.shape[1])) + w2 * a.sum(axis=0)
    for E in _euclid_basis(x.shape[1]):
        z += w1 * ((xx @ E.T) * a).sum() * (yy @ E.T)
    return z


def affineApplyDiffKT(x, p, a, kaff, regweight=1., sym=False, lddmm=False):
    """Affine part of Kernel.applyDiffKT(x, p, a) (base/kernelFunctions.py:601-625)."""
    kind, w1, w2, c = kaff
    x, p, a = _f64(x), _as2d(p), _as2d(a)
    xx = x - c
    if kind == 'affine':
        z = w1 * (p @ (a.T @ xx))
        if lddmm:
            z += w1 * (a @ (p.T @ xx) - 2 * regweight * a @ (a.T @ xx))
        elif sym:
            z += w1 * (a @ (p.T @ xx))
        return z
    z = np.zeros(x.shape)
    for E in _euclid_basis(x.shape[1]):
        ETp, ETa = p @ E, a @ E
        z += w1 * (xx * ETa).sum() * ETp
        if lddmm:
            z += w1 * ((xx * ETp).sum() * ETa - 2 * regweight * (xx * ETa).sum() * ETa)
        elif sym:
            z += w1 * (xx * ETp).sum() * ETa
    return z


def affineApplyDDiffK11and12(x, a, p, kaff):
    """Affine part of Kernel.applyDDiffK11and12(x, n, a, p) (base/kernelFunctions.py:665-674;
    n does not enter it)."""
    kind, w1, w2, c = kaff
    a, p = _as2d(a), _as2d(p)
    if kind == 'affine':
        return w1 * (a @ (a.T @ p))
    z = np.zeros(a.shape)
    for E in _euclid_basis(a.shape[1]):
        ETa = a @ E
        z += w1 * (p * ETa).sum() * ETa
    return z


def affineApplyDiffK1and2(x, a1, a2, kaff):
    """Affine part of Kernel.applyDiffK1and2(x, a1, a2) (base/kernelFunctions.py:549-561)."""
    kind, w1, w2, c = kaff
    x, a1, a2 = _f64(x), _as2d(a1), _as2d(a2)
    xx = x - c
    if kind == 'affine':
        return w1 * (a1 @ (xx.T @ a2) + (xx @ (a1.T @ a2)))
    z = np.zeros(x.shape)
    for E in _euclid_basis(x.shape[1]):
        Exx, Ea1, ETa2 = xx @ E.T, a1 @ E.T, a2 @ E
        z += w1 * ((xx * ETa2).sum() * Ea1 + (a1 * ETa2).sum() * Exx)
    return z


def affineApplyDiv(x, a, kaff):
    """Affine part of Kernel.applyDivergence(x, a) (base/kernelFunctions.py:653-659),
    the same at every target: w1 sum_j (x_j - c).a_j for 'affine'; the
    Euclidean basis is skew (trace 0).  Returns a float."""
    kind, w1, w2, c = kaff
    x, a = _f64(x), _as2d(a)
    if kind == 'affine':
        return float(w1 * ((x - c) * a).sum())
    z = 0.0
    for E in _euclid_basis(x.shape[1]):
        z += w1 * ((x - c) * (a @ E)).sum() * np.trace(E)
    return float(z)


# ---------------------------------------------------------------------------
# Shapes
# ---------------------------------------------------------------------------

def surface_geometry(vertices, faces, weights=None):
    """(centers, surfel, face_weights) of a triangulated surface.

    Follows base/surfaces.py:326-335 (computeCentersAreas); unit vertex
    weights by default (base/surfaces.py:238-240).
    """
    v = _f64(vertices)
    f = np.asarray(faces, dtype=np.int64)
    x1, x2, x3 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    centers = (x1 + x2 + x3) / 3
    surfel = np.cross(x2 - x1, x3 - x1) / 2
    if weights is None:
        fw = np.ones(f.shape[0])
    else:
        w = _f64(weights)
        fw = (w[f[:, 0]] + w[f[:, 1]] + w[f[:, 2]]) / 3
    return centers, surfel, fw


def surface_volume(vertices, faces):
    """Signed volume sum_f det(v[f]) / 6 (base/surfaces.py:1645-1651)."""
    v = _f64(vertices)
    f = np.asarray(faces, dtype=np.int64)
    return np.linalg.det(v[f]).sum() / 6


def _surf(fv):
    """Accept (vertices, faces) or (vertices, faces, weights)."""
    v, f = fv[0], fv[1]
    w = fv[2] if len(fv) > 2 else None
    c, S, fw = surface_geometry(v, f, w)
    return _f64(v), np.asarray(f, dtype=np.int64), c, S, fw


def _varifold_features(S, fw):
    a = np.sqrt((S ** 2).sum(axis=1) + 1e-10)
    n = S / a[:, None]
    return a, n, a * fw


def varifoldNorm0(fv1, sigma, KP=-1):
    """base/surfaceDistances.py:293-302 (varifoldNorm0_numpy)."""
    _, _, c2, S2, fw2 = _surf(fv1)
    _, n2, a2 = _varifold_features(S2, fw2)
    b2 = n2 * np.sqrt(a2[:, None])
    return applyktensor(c2, c2, a2, a2, b2, b2, sigma, KP).sum()


def varifoldNormDef(fvDef, fv1, sigma, KP=-1):
    """base/surfaceDistances.py:335-353 (varifoldNormDef_numpy)."""
    _, _, c1, S1, fw1 = _surf(fvDef)
    _, _, c2, S2, fw2 = _surf(fv1)
    _, n1, a1 = _varifold_features(S1, fw1)
    _, n2, a2 = _varifold_features(S2, fw2)
    b1 = n1 * np.sqrt(a1[:, None])
    b2 = n2 * np.sqrt(a2[:, None])
    obj1 = applyktensor(c1, c1, a1, a1, b1, b1, sigma, KP).sum()
    obj2 = applyktensor(c1, c2, a1, a2, b1, b2, sigma, KP).sum()
    return obj1 - 2 * obj2


def varifoldNorm(fvDef, fv1, sigma, KP=-1):
    """base/surfaceDistances.py:386-390."""
    return varifoldNormDef(fvDef, fv1, sigma, KP) + varifoldNorm0(fv1, sigma, KP)


def _scatter_faces(nv, faces, per_face_vertex):
    px = np.zeros((nv, per_face_vertex[0].shape[1]))
    for j in range(faces.shape[1]):
        np.add.at(px, faces[:, j], per_face_vertex[j])
    return px


def _vertex_sums(nv, faces, per_face):
    """sum over the faces of each vertex of a per-face scalar."""
    out = np.zeros(nv)
    for j in range(faces.shape[1]):
        np.add.at(out, faces[:, j], per_face)
    return out


def varifoldNormGradient(fvDef, fv1, sigma, KP=-1, vertex_rows=None, with_weights=False):
    """Gradient of |fvDef - fv1|^2 (varifold, f0(u) = 1 + u^2) wrt fvDef vertices.

    Follows base/surfaceDistances.py:403-468 (varifoldNormGradient_numpy),
    with the dense beta/u temporaries evaluated per row block.  vertex_rows
    (index array) returns only those vertices' rows, computing the face sums
    of their incident faces only.  with_weights also returns the weight part
    of base/surfaceDistances.py:455-467: per vertex the sum over its faces of
    (2/3) fw_f a_f (sum_j K f0 a1w_j - sum_j K f0 a2w_j) (the reference keeps
    the face weight fw_f in this factor).
    """
    xDef, f, c1, S1, fw1 = _surf(fvDef)
    _, _, c2, S2, fw2 = _surf(fv1)
    _, n1, a1w = _varifold_features(S1, fw1)
    _, n2, a2w = _varifold_features(S2, fw2)
    s, sKP, W = _scales(sigma, KP)
    sKP1 = s ** (KP - 1)
    if vertex_rows is None:
        sel = np.arange(c1.shape[0])
    else:
        vertex_rows = np.asarray(vertex_rows)
        sel = np.nonzero(np.isin(f, vertex_rows).any(axis=1))[0]
    F = sel.size
    z1 = np.zeros((F, 3))
    dz1 = np.zeros((F, 3))
    ct, nt, at = c1[sel], n1[sel], a1w[sel]
    for (cx, nx, ax, sign) in ((c1, n1, a1w, 1.), (c2, n2, a2w, -1.)):
        for i0, i1 in _row_blocks(F, cx.shape[0]):
            cc = nt[i0:i1] @ nx.T                                 # cr1cr_
            beta = at[i0:i1, None] * ax[None, :] * (1 + cc ** 2)
            u = (2 * cc[..., None] * (nx[None, :, :] - cc[..., None] * nt[i0:i1, None, :])
                 + (1 + cc ** 2)[..., None] * nt[i0:i1, None, :]) * ax[None, :, None]
            for si, sc in enumerate(s):
                d = ct[i0:i1, None, :] / sc - cx[None, :, :] / sc
                K = np.exp(-(d ** 2).sum(-1) / 2)
                z1[i0:i1] += sign * (K[..., None] * u).sum(1) * sKP[si] / W
                dz1[i0:i1] -= sign * (d * (K * beta)[..., None]).sum(1) * sKP1[si] / W
    z1 *= fw1[sel, None]
    dz1 *= 2. / 3.
    fs = f[sel]
    x1, x2, x3 = xDef[fs[:, 0]], xDef[fs[:, 1]], xDef[fs[:, 2]]
    px = _scatter_faces(xDef.shape[0], fs, (dz1 - np.cross(x3 - x2, z1),
                                           dz1 - np.cross(x1 - x3, z1),
                                           dz1 - np.cross(x2 - x1, z1)))
    if vertex_rows is not None:
        px = px[vertex_rows]
    if not with_weights:
        return px
    a1 = np.sqrt((S1 ** 2).sum(axis=1) + 1e-10)
    ones = np.ones(c1.shape[0])
    s11 = applyktensor(c1, c1, ones, a1w, n1, n1 * np.sqrt(a1w)[:, None], sigma, KP)
    s12 = applyktensor(c1, c2, ones, a2w, n1, n2 * np.sqrt(a2w)[:, None], sigma, KP)
    pxw = _vertex_sums(xDef.shape[0], f, (2 / 3) * fw1 * a1 * (s11 - s12))
    return px, (pxw if vertex_rows is None else pxw[vertex_rows])


def currentNorm0(fv1, sigma, KP=-1):
    """base/surfaceDistances.py:21-25 (currentMagnitude) via :67-68."""
    _, _, c2, S2, fw2 = _surf(fv1)
    cr2 = S2 * fw2[:, None]
    return (cr2 * applyK(c2, c2, cr2, sigma, KP)).sum()


currentMagnitude = currentNorm0


def currentNormDef(fvDef, fv1, sigma, KP=-1):
    """base/surfaceDistances.py:72-78."""
    _, _, c1, S1, fw1 = _surf(fvDef)
    _, _, c2, S2, fw2 = _surf(fv1)
    cr1 = fw1[:, None] * S1
    cr2 = fw2[:, None] * S2
    return (cr1 * applyK(c1, c1, cr1, sigma, KP)).sum() - 2 * (cr1 * applyK(c1, c2, cr2, sigma, KP)).sum()


def currentNorm(fvDef, fv1, sigma, KP=-1):
    """base/surfaceDistances.py:82-83."""
    return currentNormDef(fvDef, fv1, sigma, KP) + currentNorm0(fv1, sigma, KP)


def currentNormGradient(fvDef, fv1, sigma, KP=-1, with_weights=False):
    """base/surfaceDistances.py:87-139 (currentNormGradient); with_weights also
    returns the vertex-weight gradient (:130-137): per vertex the sum over its
    faces of (2/3) S_f . z1_f, S the unweighted surfel."""
    xDef, f, c1, S1, w1 = _surf(fvDef)
    _, _, c2, S2, w2 = _surf(fv1)
    wcr1 = w1[:, None] * S1
    wcr2 = w2[:, None] * S2
    z1 = applyK(c1, c1, wcr1, sigma, KP) - applyK(c1, c2, wcr2, sigma, KP)
    wz1 = w1[:, None] * z1
    dz1 = (2 / 3) * (applyDiffKT(c1, c1, wcr1, wcr1, sigma, KP=KP) -
                     applyDiffKT(c1, c2, wcr1, wcr2, sigma, KP=KP))
    x1, x2, x3 = xDef[f[:, 0]], xDef[f[:, 1]], xDef[f[:, 2]]
    px = _scatter_faces(xDef.shape[0], f, (dz1 - np.cross(x3 - x2, wz1),
                                          dz1 - np.cross(x1 - x3, wz1),
                                          dz1 - np.cross(x2 - x1, wz1)))
    if not with_weights:
        return px
    return px, _vertex_sums(xDef.shape[0], f, (2 / 3) * (S1 * z1).sum(axis=1))


def currentMagnitudeGradient(fvDef, sigma, KP=-1, with_weights=False):
    """base/surfaceDistances.py:28-63 (currentMagnitudeGradient); with_weights
    as currentNormGradient (:61-67)."""
    xDef, f, c1, S1, w1 = _surf(fvDef)
    wcr1 = w1[:, None] * S1
    z1 = applyK(c1, c1, wcr1, sigma, KP)
    wz1 = w1[:, None] * z1
    dz1 = (2 / 3) * applyDiffKT(c1, c1, wcr1, wcr1, sigma, KP=KP)
    x1, x2, x3 = xDef[f[:, 0]], xDef[f[:, 1]], xDef[f[:, 2]]
    px = _scatter_faces(xDef.shape[0], f, (dz1 - np.cross(x3 - x2, wz1),
                                          dz1 - np.cross(x1 - x3, wz1),
                                          dz1 - np.cross(x2 - x1, wz1)))
    if not with_weights:
        return px
    return px, _vertex_sums(xDef.shape[0], f, (2 / 3) * (S1 * z1).sum(axis=1))


# Measures ------------------------------------------------------------------

def _measure_features(S, fw):
    a = np.sqrt((S ** 2).sum(axis=1) + 1e-10)
    return a, S / a[:, None], a * fw


def measureNorm0(fv1, sigma, KP=-1):
    """base/surfaceDistances.py:208-211."""
    _, _, c2, S2, fw2 = _surf(fv1)
    _, _, cr2 = _measure_features(S2, fw2)
    return (cr2 * applyK(c2, c2, cr2[:, None], sigma, KP)[:, 0]).sum()


def measureNormDef(fvDef, fv1, sigma, KP=-1):
    """base/surfaceDistances.py:215-223."""
    _, _, c1, S1, fw1 = _surf(fvDef)
    _, _, c2, S2, fw2 = _surf(fv1)
    cr1 = _measure_features(S1, fw1)[2][:, None]
    cr2 = _measure_features(S2, fw2)[2][:, None]
    return (cr1 * applyK(c1, c1, cr1, sigma, KP)).sum() - 2 * (cr1 * applyK(c1, c2, cr2, sigma, KP)).sum()


def measureNorm(fvDef, fv1, sigma, KP=-1):
    """base/surfaceDistances.py:226-227."""
    return measureNormDef(fvDef, fv1, sigma, KP) + measureNorm0(fv1, sigma, KP)


def measureNormGradient(fvDef, fv1, sigma, KP=-1, with_weights=False):
    """base/surfaceDistances.py:231-278; with_weights also returns the
    vertex-weight gradient (:269-276): per vertex the sum over its faces of
    (2/3) a_f z1_f."""
    xDef, f, c1, S1, w1 = _surf(fvDef)
    _, _, c2, S2, w2 = _surf(fv1)
    a1, cr1, a1w = _measure_features(S1, w1)
    _, _, a2w = _measure_features(S2, w2)
    z1 = applyK(c1, c1, a1w[:, None], sigma, KP) - applyK(c1, c2, a2w[:, None], sigma, KP)
    wz1 = w1[:, None] * z1 * cr1
    dz1 = (2 / 3) * (applyDiffKT(c1, c1, a1w[:, None], a1w[:, None], sigma, KP=KP) -
                     applyDiffKT(c1, c2, a1w[:, None], a2w[:, None], sigma, KP=KP))
    x1, x2, x3 = xDef[f[:, 0]], xDef[f[:, 1]], xDef[f[:, 2]]
    px = _scatter_faces(xDef.shape[0], f, (dz1 - np.cross(x3 - x2, wz1),
                                          dz1 - np.cross(x1 - x3, wz1),
                                          dz1 - np.cross(x2 - x1, wz1)))
    if not with_weights:
        return px
    return px, _vertex_sums(xDef.shape[0], f, (2 / 3) * a1 * z1[:, 0])


def measureNormPS0(points, weights, sigma, KP=-1):
    """Point-set target measure |mu_1|^2 (base/surfaceDistances.py:141-144)."""
    p, w = _f64(points), _f64(weights)[:, None]
    return (w * applyK(p, p, w, sigma, KP)).sum()


def measureNormPSDef(fvDef, points, weights, sigma, KP=-1):
    """Surface vs point set (base/surfaceDistances.py:148-155)."""
    _, _, c1, S1, fw1 = _surf(fvDef)
    cr1 = _measure_features(S1, fw1)[2][:, None]
    p, w = _f64(points), _f64(weights)[:, None]
    return (cr1 * applyK(c1, c1, cr1, sigma, KP)).sum() - 2 * (cr1 * applyK(c1, p, w, sigma, KP)).sum()


def measureNormPSGradient(fvDef, points, weights, sigma, KP=-1):
    """base/surfaceDistances.py:163-200: the surface side uses the unweighted
    face measure a = |S| (the reference ignores face weights here)."""
    xDef, f, c1, S1, _ = _surf(fvDef)
    a1 = np.sqrt((S1 ** 2).sum(axis=1) + 1e-10)
    cr1 = S1 / a1[:, None]
    p, w = _f64(points), _f64(weights)
    z1 = applyK(c1, c1, a1[:, None], sigma, KP) - applyK(c1, p, w[:, None], sigma, KP)
    z1 = z1 * cr1
    dz1 = (2 / 3) * (applyDiffKT(c1, c1, a1[:, None], a1[:, None], sigma, KP=KP) -
                     applyDiffKT(c1, p, a1[:, None], w[:, None], sigma, KP=KP))
    x1, x2, x3 = xDef[f[:, 0]], xDef[f[:, 1]], xDef[f[:, 2]]
    return _scatter_faces(xDef.shape[0], f, (dz1 - np.cross(x3 - x2, z1),
                                            dz1 - np.cross(x1 - x3, z1),
                                            dz1 - np.cross(x2 - x1, z1)))


def pointMeasureNorm0(x1, w1, sigma, KP=-1):
    """Point sets (base/pointset_distances.py:14-16)."""
    x1, w = _f64(x1), _f64(w1)[:, None]
    return (w * applyK(x1, x1, w, sigma, KP)).sum()


def pointMeasureNormDef(xDef, wDef, x1, w1, sigma, KP=-1):
    """base/pointset_distances.py:20-25."""
    xd, wd, x1, w = _f64(xDef), _f64(wDef)[:, None], _f64(x1), _f64(w1)[:, None]
    return (wd * applyK(xd, xd, wd, sigma, KP)).sum() - 2 * (wd * applyK(xd, x1, w, sigma, KP)).sum()


def pointMeasureNormGradient(xDef, wDef, x1, w1, sigma, KP=-1):
    """base/pointset_distances.py:34-41."""
    xd, wd, x1, w = _f64(xDef), _f64(wDef)[:, None], _f64(x1), _f64(w1)[:, None]
    return 2 * (applyDiffKT(xd, xd, wd, wd, sigma, KP=KP) - applyDiffKT(xd, x1, wd, w, sigma, KP=KP))


# Curves --------------------------------------------------------------------

def curve_geometry(vertices, faces, weights=None):
    """(centers, linel, line_weights); base/curves.py:117-125."""
    v = _f64(vertices)
    f = np.asarray(faces, dtype=np.int64)
    x1, x2 = v[f[:, 0]], v[f[:, 1]]
    w = np.ones(v.shape[0]) if weights is None else _f64(weights)
    return (x1 + x2) / 2, x2 - x1, (w[f[:, 0]] + w[f[:, 1]]) / 2


def _curve(fv):
    v, f = fv[0], fv[1]
    w = fv[2] if len(fv) > 2 else None
    c, l, lw = curve_geometry(v, f, w)
    return _f64(v), np.asarray(f, dtype=np.int64), c, l, lw


def _curve_vfeat(cr):
    a = np.sqrt((cr ** 2).sum(axis=1) + 1e-10)
    return a, cr / a[:, None]


def curveVarifoldNorm0(fv1, sigma, weight=1., KP=-1):
    """base/curveDistances.py:186-200 (_varifoldNorm0 / varifoldNorm0)."""
    _, _, c2, l2, w2 = _curve(fv1)
    a2, n2 = _curve_vfeat(w2[:, None] * l2)
    tot = 0.
    for i0, i1 in _row_blocks(c2.shape[0], c2.shape[0]):
        cc = n2[i0:i1] @ n2.T
        beta = (1 + weight * cc ** 2) * a2[i0:i1, None] * a2[None, :]
        tot += applykmat(c2[i0:i1], c2, beta, sigma, KP).sum()
    return tot


def curveVarifoldNormDef(fvDef, fv1, sigma, weight=1., KP=-1):
    """base/curveDistances.py:218-242 (_varifoldNormDef / varifoldNormDef)."""
    _, _, c1, l1, w1 = _curve(fvDef)
    _, _, c2, l2, w2 = _curve(fv1)
    a1, n1 = _curve_vfeat(w1[:, None] * l1)
    a2, n2 = _curve_vfeat(w2[:, None] * l2)
    tot = 0.
    for (cx, nx, ax, sgn) in ((c1, n1, a1, 1.), (c2, n2, a2, -2.)):
        for i0, i1 in _row_blocks(c1.shape[0], cx.shape[0]):
            cc = n1[i0:i1] @ nx.T
            beta = (1 + weight * cc ** 2) * a1[i0:i1, None] * ax[None, :]
            tot += sgn * applykmat(c1[i0:i1], cx, beta, sigma, KP).sum()
    return tot


def curveVarifoldNormGradient(fvDef, fv1, sigma, weight=1., KP=-1):
    """base/curveDistances.py:280-336 (_varifoldNormGradient / varifoldNormGradient)."""
    xDef, f, c1, l1, w1 = _curve(fvDef)
    _, _, c2, l2, w2 = _curve(fv1)
    d = weight
    a1, n1 = _curve_vfeat(l1)
    a2, n2 = _curve_vfeat(l2)
    a1w, a2w = a1 * w1, a2 * w2
    F = c1.shape[0]
    z1 = np.zeros((F, xDef.shape[1]))
    dz1 = np.zeros((F, xDef.shape[1]))
    for (cx, nx, ax, sgn) in ((c1, n1, a1w, 1.), (c2, n2, a2w, -1.)):
        for i0, i1 in _row_blocks(F, cx.shape[0]):
            cc = n1[i0:i1] @ nx.T
            beta = a1w[i0:i1, None] * ax[None, :] * (1 + d * cc ** 2)
            u = (2 * d * cc[..., None] * nx[None, ...]
                 - d * (cc ** 2)[..., None] * n1[i0:i1, None, :]
                 + n1[i0:i1, None, :]) * ax[None, :, None]
            z1[i0:i1] += sgn * 2 * applykmat(c1[i0:i1], cx, u, sigma, KP)
            dz1[i0:i1] += sgn * applykdiffmat(c1[i0:i1], cx, beta, sigma, KP)
    wz1 = z1 * w1[:, None]
    return _scatter_faces(xDef.shape[0], f, (dz1 - wz1, dz1 + wz1))


def curveCurrentNorm0(fv1, sigma, KP=-1):
    """base/curveDistances.py:30-37 (currentNorm0, weight=None)."""
    _, _, c2, l2, w2 = _curve(fv1)
    cr2 = l2 * w2[:, None]
    return (cr2 * applyK(c2, c2, cr2, sigma, KP)).sum()


def curveCurrentNormDef(fvDef, fv1, sigma, KP=-1):
    """base/curveDistances.py:41-52 (currentNormDef, weight=None)."""
    _, _, c1, l1, w1 = _curve(fvDef)
    _, _, c2, l2, w2 = _curve(fv1)
    cr1 = l1 * w1[:, None]
    cr2 = l2 * w2[:, None]
    return (cr1 * applyK(c1, c1, cr1, sigma, KP)).sum() - 2 * (cr1 * applyK(c1, c2, cr2, sigma, KP)).sum()


def curveCurrentNormGradient(fvDef, fv1, sigma, KP=-1):
    """base/curveDistances.py:61-110 (currentNormGradient, weight=None)."""
    xDef, f, c1, l1, w1 = _curve(fvDef)
    _, _, c2, l2, w2 = _curve(fv1)
    wcr1 = w1[:, None] * l1
    wcr2 = w2[:, None] * l2
    z1 = 2 * (applyK(c1, c1, wcr1, sigma, KP) - applyK(c1, c2, wcr2, sigma, KP))
    wz1 = w1[:, None] * z1
    dz1 = applyDiffKT(c1, c1, wcr1, wcr1, sigma, KP=KP) - applyDiffKT(c1, c2, wcr1, wcr2, sigma, KP=KP)
    return _scatter_faces(xDef.shape[0], f, (dz1 - wz1, dz1 + wz1))


# ---------------------------------------------------------------------------
# Kernel.solve: conjugate gradients on K(x, x)
# ---------------------------------------------------------------------------

def linearcg(op, b, iter_max=100, sequence='reference'):
    """Solve op(a) = b from a = 0 (base/conjugateGradient.py:351-408, no x0 /
    param).  The whole array is one vector (dot products over all entries).
    Stops: |b|^2 < 1e-10 -> zero; alpha * mean|p| < 1e-10; energy a.op(a) - 2 b.a
    changing by less than 1e-20 of its previous value; iter_max products.

    sequence 'reference' is linearcg as numpy executes it: at the first
    iteration `p = r` binds the direction to the residual array, and the
    in-place `r -= alpha * q` moves both, so that iteration's stopping test
    reads the updated residual and the second direction is r1 + beta r1.
    'textbook' keeps the direction and the residual apart.  Returns
    (a, products)."""
    b = _f64(b)
    a = np.zeros(b.shape)
    z = -2.0 * b                       # op(a) - 2 b at a = 0
    r = b.copy()
    ener_old = 0.0
    mu = 0.0
    p = None
    products = 0
    for it in range(iter_max):
        mu_new = float((r * r).sum())
        if it == 0:
            if mu_new < 1e-10:
                break
            p = r.copy()
        else:
            p = r + (mu_new / mu) * p
        mu = mu_new
        q = op(p)
        products += 1
        alpha = mu / float((p * q).sum())
        a = a + alpha * p
        r = r - alpha * q
        z = z + alpha * q
        if it == 0 and sequence == 'reference':
            p = r.copy()               # the direction is the residual array
        ener = float((z * a).sum())
        if alpha * float(np.abs(p).mean()) < 1e-10:
            break
        if abs(ener - ener_old) < 1e-20 * abs(ener_old):
            break
        ener_old = ener
    return a, products


def kernel_solve(x, b, scale, iter_max=100, sequence='reference', KP=-1, kernel=GAUSS):
    """Kernel.solve(x, b) (base/kernelFunctions.py:676-680): linearcg with the
    operator a -> K(x, x) a.  Returns (a, products)."""
    x = _f64(x)
    return linearcg(lambda v: applyK(x, x, v, scale, KP, kernel=kernel), b, iter_max, sequence)


# ---------------------------------------------------------------------------
# Landmark shooting (first order) and its discrete adjoint
# ---------------------------------------------------------------------------

def _kv(x, a, sigma, KP, kernel, kaff):
    """Kernel.applyK(x, a): the kernel sum plus the affine part, if any."""
    v = applyK(x, x, a, sigma, KP, kernel=kernel)
    return v if kaff is None else v + affineApplyK(x, x, a, kaff)


def landmarkDirectEvolutionEuler(x0, at, sigma, KP=-1, kernel=GAUSS, kaff=None):
    """x_{t+1} = x_t + (1/T) K(x_t, x_t) a_t; base/pointEvolution.py:88-164 (no
    affine control; kaff = the kernel's affine part)."""
    x0, at = _f64(x0), _f64(at)
    T = at.shape[0]
    dt = 1.0 / T
    xt = np.zeros((T + 1,) + x0.shape)
    xt[0] = x0
    for t in range(T):
        xt[t + 1] = xt[t] + dt * _kv(xt[t], at[t], sigma, KP, kernel, kaff)
    return xt


def jacobian(xt, at, sigma, KP=-1, kernel=GAUSS, kaff=None):
    """J_0 = 0, J_{t+1} = J_t + dt div v_t(x_t), v_t = K(., x_t) a_t, dt = 1/T with
    T = len(at): the withJacobian output of base/pointEvolution.py:149-152
    (direct evolution), :366-367 (landmarkEPDiff) and :459-463
    (secondOrderEvolution); with xt = the control points c_t, the semi-reduced
    flow's (base/pointEvolutionSemiReduced.py:78-81)."""
    xt, at = _f64(xt), _f64(at)
    T = at.shape[0]
    dt = 1.0 / T
    Jt = np.zeros((T + 1, xt.shape[1], 1))
    for t in range(T):
        d = applyDiv(xt[t], xt[t], at[t], sigma, KP, kernel=kernel)
        if kaff is not None:
            d = d + affineApplyDiv(xt[t], at[t], kaff)
        Jt[t + 1] = Jt[t] + dt * d
    return Jt


def landmarkHamiltonianCovector(x0, at, px1, sigma, regweight, KP=-1, kernel=GAUSS, kaff=None):
    """base/pointEvolution.py:167-201 (no affine control, no extra term)."""
    xt = landmarkDirectEvolutionEuler(x0, at, sigma, KP, kernel=kernel, kaff=kaff)
    T = at.shape[0]
    dt = 1.0 / T
    pxt = np.zeros(xt.shape)
    pxt[T] = px1
    for t in range(T):
        px = pxt[T - t]
        z = xt[T - t - 1]
        a = at[T - t - 1]
        zpx = applyDiffKT(z, z, px, a, sigma, regweight=regweight, lddmm=True, KP=KP, kernel=kernel)
        if kaff is not None:
            zpx += affineApplyDiffKT(z, px, a, kaff, regweight=regweight, lddmm=True)
        pxt[T - t - 1] = px + dt * zpx
    return pxt, xt


def landmarkHamiltonianGradient(x0, at, px1, sigma, regweight, KP=-1, kernel=GAUSS, kaff=None):
    """dat_t = 2 r a_t - p_{t+1}; base/pointEvolution.py:205-236 (no affine control)."""
    pxt, xt = landmarkHamiltonianCovector(x0, at, px1, sigma, regweight, KP, kernel=kernel, kaff=kaff)
    at = _f64(at)
    dat = 2 * regweight * at - pxt[1:]
    return dat, xt, pxt


# ---------------------------------------------------------------------------
# Internal (regularity) costs of a velocity field phi on a triangulated
# surface (base/surfaceDistances.py:631-872): 'h1' (normGrad / diffNormGrad),
# 'elastic' (elasticNorm / diffElasticNorm), 'displacement'
# (normDisplacement / diffNormDisplacement).  Each returns the cost, the
# derivative in phi and the derivative in the vertices, in the reference's
# formulas.
# ---------------------------------------------------------------------------

def _ic_face_terms(x, f, phi):
    v1, v2, v3 = x[f[:, 0]], x[f[:, 1]], x[f[:, 2]]
    p1, p2, p3 = phi[f[:, 0]], phi[f[:, 1]], phi[f[:, 2]]
    S = np.cross(v2 - v1, v3 - v1) / 2
    A = np.sqrt((S ** 2).sum(1))
    l1 = ((v2 - v3) ** 2).sum(1)
    l2 = ((v1 - v3) ** 2).sum(1)
    l3 = ((v1 - v2) ** 2).sum(1)
    return v1, v2, v3, p1, p2, p3, S, A, l1, l2, l3


def _dot(a, b):
    return (a * b).sum(1)


def internal_cost(kind, vertices, faces, phi):
    """(cost, d cost / d phi, d cost / d vertices) for kind in 'h1', 'elastic',
    'displacement' (base/surfaceDistances.py:631-645, :680-720, :731-777,
    :779-872)."""
    x, f, phi = _f64(vertices), np.asarray(faces), _f64(phi)
    v1, v2, v3, p1, p2, p3, S, A, l1, l2, l3 = _ic_face_terms(x, f, phi)
    nv = x.shape[0]
    if kind == 'displacement':
        pc = (p1 + p2 + p3) / 3
        cost = ((pc ** 2).sum(1) * A).sum()
        g = pc * A[:, None] * (2 / 3)
        gphi = _scatter_faces(nv, f, (g, g, g))
        nphi = ((pc ** 2).sum(1) / A)[:, None]
        gx = _scatter_faces(nv, f, (np.cross(v2 - v3, S) * nphi, np.cross(v3 - v1, S) * nphi,
                                    np.cross(v1 - v2, S) * nphi)) / 2
        return cost, gphi, gx
    if kind not in ('h1', 'elastic'):
        raise ValueError(kind)
    q1 = _dot(p2 - p1, p3 - p1)          # (phi2 - phi1).(phi3 - phi1)
    q2 = _dot(p3 - p2, p1 - p2)
    q3 = _dot(p1 - p3, p2 - p3)
    u = l1 * q1 + l2 * q2 + l3 * q3
    a = 2 * A
    a2 = 2 * a
    elastic = kind == 'elastic'
    corr = _dot(v1 - v3, p1 - p2) - _dot(v1 - v2, p1 - p3) if elastic else np.zeros_like(u)
    ue = u - 0.5 * corr ** 2
    cost = (ue / (4 * A)).sum()
    r1 = (l1[:, None] * (p2 + p3 - 2 * p1) + (l2 - l3)[:, None] * (p2 - p3)) / a2[:, None]
    r2 = (l2[:, None] * (p1 + p3 - 2 * p2) + (l1 - l3)[:, None] * (p1 - p3)) / a2[:, None]
    r3 = (l3[:, None] * (p1 + p2 - 2 * p3) + (l2 - l1)[:, None] * (p2 - p1)) / a2[:, None]
    if elastic:
        c_a = (corr / a)[:, None]
        r1 = r1 + 0.5 * (v2 - v3) * c_a
        r2 = r2 + 0.5 * (v3 - v1) * c_a
        r3 = r3 + 0.5 * (v1 - v2) * c_a
    gphi = -_scatter_faces(nv, f, (r1, r2, r3))
    uu = (ue / a ** 3)[:, None]
    x1 = -uu * np.cross(v2 - v3, S) + 2 * ((v1 - v3) * q2[:, None] + (v1 - v2) * q3[:, None]) / a2[:, None]
    x2 = -uu * np.cross(v3 - v1, S) + 2 * ((v2 - v1) * q3[:, None] + (v2 - v3) * q1[:, None]) / a2[:, None]
    x3 = -uu * np.cross(v1 - v2, S) + 2 * ((v3 - v2) * q1[:, None] + (v3 - v1) * q2[:, None]) / a2[:, None]
    if elastic:
        c_a = (corr / a)[:, None]
        x1 = x1 - 0.5 * (p3 - p2) * c_a
        x2 = x2 - 0.5 * (p1 - p3) * c_a
        x3 = x3 - 0.5 * (p2 - p1) * c_a
    gx = _scatter_faces(nv, f, (x1, x2, x3))
    return cost, gphi, gx


def internal_costs(icost, vertices, faces, phi):
    """sum_k w_k C_k over icost = [(kind, w_k), ...] (SurfaceMatching.internalCost__ /
    internalCostGrad__, base/surfaceMatching.py:156-173)."""
    c, gp, gx = 0.0, np.zeros(np.shape(phi)), np.zeros(np.shape(vertices))
    for kind, w in icost:
        ck, gpk, gxk = internal_cost(kind, vertices, faces, phi)
        c += w * ck
        gp += w * gpk
        gx += w * gxk
    return c, gp, gx


def _data_term(kind, fvDef, fv1, sigma_dist):
    if kind == 'varifold':
        return varifoldNormDef(fvDef, fv1, sigma_dist), varifoldNorm0(fv1, sigma_dist)
    if kind == 'current':
        return currentNormDef(fvDef, fv1, sigma_dist), currentNorm0(fv1, sigma_dist)
    if kind == 'measure':
        return measureNormDef(fvDef, fv1, sigma_dist), measureNorm0(fv1, sigma_dist)
    if kind == 'point_measure':             # fvDef = (x, w) (weights in place of faces), fv1 = (y, w1)
        return (pointMeasureNormDef(fvDef[0], fvDef[1], fv1[0], fv1[1], sigma_dist),
                pointMeasureNorm0(fv1[0], fv1[1], sigma_dist))
    if kind == 'curve_varifold':
        return curveVarifoldNormDef(fvDef, fv1, sigma_dist), curveVarifoldNorm0(fv1, sigma_dist)
    if kind == 'curve_current':
        return curveCurrentNormDef(fvDef, fv1, sigma_dist), curveCurrentNorm0(fv1, sigma_dist)
    raise ValueError(kind)


def _data_grad(kind, fvDef, fv1, sigma_dist):
    if kind == 'varifold':
        return varifoldNormGradient(fvDef, fv1, sigma_dist)
    if kind == 'current':
        return currentNormGradient(fvDef, fv1, sigma_dist)
    if kind == 'measure':
        return measureNormGradient(fvDef, fv1, sigma_dist)
    if kind == 'point_measure':
        return pointMeasureNormGradient(fvDef[0], fvDef[1], fv1[0], fv1[1], sigma_dist)
    if kind == 'curve_varifold':
        return curveVarifoldNormGradient(fvDef, fv1, sigma_dist)
    if kind == 'curve_current':
        return curveCurrentNormGradient(fvDef, fv1, sigma_dist)
    raise ValueError(kind)


# ---------------------------------------------------------------------------
# Unreduced (semi-reduced) matching: the velocity at time t is generated by
# control points c_t (control 'ct' beside 'at'): v_t(x) = K(x, c_t) a_t
# (base/pointEvolutionSemiReduced.py:7-90, base/surfaceMatching.py:468-565,
# :728-908 unreduced branches).
# ---------------------------------------------------------------------------

def _kyx(y, x, a, sigma, kernel, kaff):
    """Kernel.applyK(x, a, firstVar=y) with the affine part."""
    v = applyK(y, x, a, sigma, -1, kernel=kernel)
    return v if kaff is None else v + affineApplyK(y, x, a, kaff)


def _dkt_yx(y, x, p, a, sigma, kernel, kaff):
    """Kernel.applyDiffKT(x, p, a, firstVar=y) (plain) with the affine part."""
    z = applyDiffKT(y, x, p, a, sigma, kernel=kernel)
    return z if kaff is None else z + affineApplyDiffKT(x, p, a, kaff)


def semiReducedEvolution(x0, ct, at, sigma, fidelity_weight=0., kernel=GAUSS, kaff=None):
    """x_{t+1} = x_t + dt (K(x_t, c_t) a_t - fw (x_t - c_t)) (the fidelity term
    when c_t has as many rows as x_t); base/pointEvolutionSemiReduced.py:7-90."""
    x0, ct, at = _f64(x0), _f64(ct), _f64(at)
    T = at.shape[0]
    dt = 1.0 / T
    fid = ct.shape[1] == x0.shape[0]
    xt = np.zeros((T + 1,) + x0.shape)
    xt[0] = x0
    for t in range(T):
        dx = _kyx(xt[t], ct[t], at[t], sigma, kernel, kaff)
        if fid:
            dx = dx - fidelity_weight * (xt[t] - ct[t])
        xt[t + 1] = xt[t] + dt * dx
    return xt


def matching_objective_unreduced(x0, faces, ct, at, target, sigma_diff, sigma_dist, kind='varifold',
                                 regweight=1., sigma_error=1., unreduced_weight=1., fidelity_weight=0.,
                                 kernel_diff=GAUSS, kaff=None, icost=None):
    """sum_t dt [r a_t.K(c_t, c_t) a_t + uw |c_t - x_t|^2 + C(x_t, K(x_t, c_t) a_t)]
    + data term (base/surfaceMatching.py:536-553); `unreduced_weight` is the
    option after its 1000 / N scaling (:93)."""
    ct, at = _f64(ct), _f64(at)
    T = at.shape[0]
    dt = 1.0 / T
    xt = semiReducedEvolution(x0, ct, at, sigma_diff, fidelity_weight, kernel_diff, kaff)
    obj = 0.
    for t in range(T):
        z, a, c = xt[t], at[t], ct[t]
        obj += regweight * dt * (a * _kyx(c, c, a, sigma_diff, kernel_diff, kaff)).sum()
        obj += unreduced_weight * dt * ((c - z) ** 2).sum()
        if icost:
            obj += internal_costs(icost, z, faces, _kyx(z, c, a, sigma_diff, kernel_diff, kaff))[0] * dt
    dterm, obj0 = _data_term(kind, (xt[-1], faces), target, sigma_dist)
    return obj + (obj0 + dterm) / sigma_error ** 2, xt


def matching_gradient_unreduced(x0, faces, ct, at, target, sigma_diff, sigma_dist, kind='varifold',
                                regweight=1., sigma_error=1., unreduced_weight=1., fidelity_weight=0.,
                                kernel_diff=GAUSS, kaff=None, icost=None):
    """(grad ct, grad at) of the unreduced objective as BFGS sees it
    (base/surfaceMatching.py:782-813 covector, :861-887 gradient, :1082-1083 / T)."""
    ct, at = _f64(ct), _f64(at)
    T = at.shape[0]
    dt = 1.0 / T
    fid = ct.shape[1] == np.shape(x0)[0]
    xt = semiReducedEvolution(x0, ct, at, sigma_diff, fidelity_weight, kernel_diff, kaff)
    pxt = np.zeros(xt.shape)
    pxt[T] = -_data_grad(kind, (xt[-1], faces), target, sigma_dist) / sigma_error ** 2
    dat, dct = np.zeros(at.shape), np.zeros(ct.shape)
    for t in range(T - 1, -1, -1):
        z, a, c, px = xt[t], at[t], ct[t], pxt[t + 1]
        pp, DLv = px, 0.
        if icost:
            v = _kyx(z, c, a, sigma_diff, kernel_diff, kaff)
            _, Lv, DLv = internal_costs(icost, z, faces, v)
            pp = px - Lv
        zpx = _dkt_yx(z, c, pp, a, sigma_diff, kernel_diff, kaff) - DLv - 2 * unreduced_weight * (z - c)
        if fid:
            zpx = zpx - fidelity_weight * px
        pxt[t] = px + dt * zpx
        dat[t] = 2 * regweight * _kyx(c, c, a, sigma_diff, kernel_diff, kaff) - _kyx(c, z, pp, sigma_diff,
                                                                                     kernel_diff, kaff)
        dct[t] = (2 * regweight * _dkt_yx(c, c, a, a, sigma_diff, kernel_diff, kaff)
                  - _dkt_yx(c, z, a, pp, sigma_diff, kernel_diff, kaff) + 2 * unreduced_weight * (c - z))
        if fid:
            dct[t] -= fidelity_weight * px
    return dct / T, dat / T


def _internal_cost_gradient(xt, at, px1, faces, sigma, regweight, kernel, kaff, icost):
    """Covector and Hamiltonian gradient with internal costs: with
    v = K(x_t) a_t, Lv / DLv the cost's derivatives in phi / x at (x_t, v),
    p_t = p_{t+1} + dt (applyDiffKT(x_t, p_{t+1} - Lv, a_t, lddmm) - DLv) and
    dat_t = 2 r a_t - p_{t+1} + Lv (base/surfaceMatching.py:787-804, :860-885)."""
    T = at.shape[0]
    dt = 1.0 / T
    pxt = np.zeros(xt.shape)
    pxt[T] = px1
    dat = np.zeros(at.shape)
    for t in range(T - 1, -1, -1):
        z, a, px = xt[t], at[t], pxt[t + 1]
        v = _kv(z, a, sigma, -1, kernel, kaff)
        _, Lv, DLv = internal_costs(icost, z, faces, v)
        pe = px - Lv
        zpx = applyDiffKT(z, z, pe, a, sigma, regweight=regweight, lddmm=True, kernel=kernel)
        if kaff is not None:
            zpx += affineApplyDiffKT(z, pe, a, kaff, regweight=regweight, lddmm=True)
        pxt[t] = px + dt * (zpx - DLv)
        dat[t] = 2 * regweight * a - px + Lv
    return dat, pxt


def matching_objective(x0, faces, at, target, sigma_diff, sigma_dist, kind='varifold',
                       regweight=1., sigma_error=1., kernel_diff=GAUSS, kaff=None, icost=None):
    """Surface- or curve-matching objective at control `at` (plain, no affine, ds = 1).
    kind: 'varifold' / 'current' / 'measure' (surfaces), 'curve_varifold' /
    'curve_current' (curves, base/curveMatching.py:272-313 has the same
    structure), 'point_measure' (point sets, base/pointSetMatching.py; `faces`
    then holds the template's point weights and `target` is (points, weights)).

    sum_t r dt a_t.K(x_t)a_t  +  (|def|^2 - 2<def, tgt> + |tgt|^2) / sigma_err^2.
    Follows base/surfaceMatching.py:468-565 (objectiveFunDef) and :638-675
    (updateTry: obj0 + deformation + dataTerm), dataTerm :443-466.
    """
    at = _f64(at)
    T = at.shape[0]
    dt = 1.0 / T
    xt = landmarkDirectEvolutionEuler(x0, at, sigma_diff, kernel=kernel_diff, kaff=kaff)
    obj = 0.
    for t in range(T):
        ra = _kv(xt[t], at[t], sigma_diff, -1, kernel_diff, kaff)
        obj += regweight * dt * (at[t] * ra).sum()
        if icost:                              # base/surfaceMatching.py:551-553
            obj += internal_costs(icost, xt[t], faces, ra)[0] * dt
    dterm, obj0 = _data_term(kind, (xt[-1], faces), target, sigma_dist)
    return obj + (obj0 + dterm) / sigma_error ** 2, xt


def matching_gradient(x0, faces, at, target, sigma_diff, sigma_dist, kind='varifold',
                      regweight=1., sigma_error=1., euclidean=True, kernel_diff=GAUSS, kaff=None,
                      icost=None):
    """grad['at'] of the plain surface-matching objective as BFGS sees it.

    px1 = -endPointGradient/sigma_err^2 (base/surfaceMatching.py:702-721,
    :1077), covector by the lddmm adjoint (:783-818), dat_t = 2 r a_t - p_{t+1}
    (:879), K(x_t) dat_t for the Euclidean gradient (:890-891), / T (:1086).
    """
    at = _f64(at)
    T = at.shape[0]
    xt = landmarkDirectEvolutionEuler(x0, at, sigma_diff, kernel=kernel_diff, kaff=kaff)
    px1 = -_data_grad(kind, (xt[-1], faces), target, sigma_dist) / sigma_error ** 2
    if icost:
        dat, pxt = _internal_cost_gradient(xt, at, px1, faces, sigma_diff, regweight, kernel_diff, kaff, icost)
    else:
        dat, xt, pxt = landmarkHamiltonianGradient(x0, at, px1, sigma_diff, regweight, kernel=kernel_diff,
                                                   kaff=kaff)
    if euclidean:
        for t in range(T):
            dat[t] = _kv(xt[t], dat[t], sigma_diff, -1, kernel_diff, kaff)
    return dat / T


# ---------------------------------------------------------------------------
# Geodesic shooting (second order) and its adjoint
# ---------------------------------------------------------------------------

def landmarkEPDiff(T, x0, a0, sigma, KP=-1, kernel=GAUSS):
    """base/pointEvolution.py:320-379 (simple output)."""
    x0, a0 = _f64(x0), _f64(a0)
    dt = 1.0 / T
    at = np.zeros((T,) + x0.shape)
    xt = np.zeros((T + 1,) + x0.shape)
    xt[0] = x0
    at[0] = a0
    for k in range(T):
        z, a = xt[k], at[k]
        xt[k + 1] = z + dt * applyK(z, z, a, sigma, KP, kernel=kernel)
        if k < T - 1:
            at[k + 1] = a - dt * applyDiffKT(z, z, a, a, sigma, KP=KP, kernel=kernel)
    return xt, at


def secondOrderEvolution(x0, a0, sigma, timeStep, KP=-1, kernel=GAUSS, kaff=None):
    """base/pointEvolution.py:385-472 (no spline, no affine control, no point
    set); kaff = the kernel's affine part, which Kernel.applyK and
    Kernel.applyDiffKT add."""
    x0, a0 = _f64(x0), _f64(a0)
    T = int(np.round(1 / timeStep))
    xt = np.zeros((T + 1,) + x0.shape)
    at = np.zeros((T + 1,) + x0.shape)
    xt[0], at[0] = x0, a0
    for k in range(T):
        x, a = xt[k], at[k]
        zx = applyK(x, x, a, sigma, KP, kernel=kernel)
        za = -applyDiffKT(x, x, a, a, sigma, KP=KP, kernel=kernel)
        if kaff is not None:
            zx = zx + affineApplyK(x, x, a, kaff)
            za = za - affineApplyDiffKT(x, a, a, kaff)
        xt[k + 1] = x + timeStep * zx
        at[k + 1] = a + timeStep * za
    return xt, at


def secondOrderHamiltonian(x, a, px, pa, sigma, KP=-1, kernel=GAUSS):
    """base/pointEvolution.py:475-486 (no spline, no affine)."""
    Ht = (px * applyK(x, x, a, sigma, KP, kernel=kernel)).sum()
    Ht += (pa * (-applyDiffKT(x, x, a, a, sigma, KP=KP, kernel=kernel))).sum()
    return Ht


def secondOrderCovector(x0, a0, px1, pa1, sigma, timeStep, KP=-1, kernel=GAUSS, kaff=None):
    """base/pointEvolution.py:489-552 (single target, no affine control; kaff
    = the kernel's affine part, added by the Kernel methods the loop calls)."""
    T = int(np.round(1 / timeStep))
    xt, at = secondOrderEvolution(x0, a0, sigma, timeStep, KP, kernel=kernel, kaff=kaff)
    pxt = np.zeros(xt.shape)
    pat = np.zeros(xt.shape)
    pxt[T], pat[T] = px1, pa1
    for t in range(T):
        px, pa = pxt[T - t], pat[T - t]
        x, a = xt[T - t - 1], at[T - t - 1]
        zpx = (applyDiffKT(x, x, px, a, sigma, sym=True, KP=KP, kernel=kernel)
               - applyDDiffK(x, a, a, pa, sigma, KP=KP, kernel=kernel))
        # Kernel.applyDiffK1and2 -> applykdiff1and2, which forces KP = -1
        zpa = (applyK(x, x, px, sigma, KP, kernel=kernel)
               - applyDiffK(x, pa, a, sigma, KP=-1, option='1and2', kernel=kernel))
        if kaff is not None:
            zpx = zpx + affineApplyDiffKT(x, px, a, kaff, sym=True) - affineApplyDDiffK11and12(x, a, pa, kaff)
            zpa = zpa + affineApplyK(x, x, px, kaff) - affineApplyDiffK1and2(x, pa, a, kaff)
        pxt[T - t - 1] = px + timeStep * zpx
        pat[T - t - 1] = pa + timeStep * zpa
    return pxt, pat, xt, at


def secondOrderGradient(x0, a0, px1, pa1, sigma, timeStep, KP=-1, kernel=GAUSS, kaff=None):
    """da0 = K(x0) a0 - pa_0; base/pointEvolution.py:555-618 (no affine
    control; K with the kernel's affine part kaff)."""
    pxt, pat, xt, at = secondOrderCovector(x0, a0, px1, pa1, sigma, timeStep, KP, kernel=kernel, kaff=kaff)
    Ka = applyK(x0, x0, a0, sigma, KP, kernel=kernel)
    if kaff is not None:
        Ka = Ka + affineApplyK(x0, x0, a0, kaff)
    da0 = Ka - pat[0]
    return da0, xt, at, pxt, pat


def second_order_matching_objective(x0, w0, a0, target, sigma_diff, sigma_dist, timeStep, kind='point_measure',
                                    sigma_error=1., kernel_diff=GAUSS, kaff=None):
    """Second-order (geodesic) point-set matching objective at the initial
    momentum a0: 1/2 a0.K(x0)a0 + (obj0 + dataTerm(x_T)) / sigma_err^2, with
    x_T the endpoint of secondOrderEvolution
    (base/secondOrderPointSetMatching.py:105-140 objectiveFunDef, no affine;
    the data term and obj0 as in base/pointSetMatching.py).  kind
    'point_measure' (w0 = template weights, target = (points, weights)),
    'L2' (target = points of the same shape; |x_T - y|^2 with obj0 = |y|^2),
    or a surface data term 'varifold' / 'current' / 'measure' (w0 = template
    faces, target = (vertices, faces)), the objective of
    base/secondOrderSurfaceMatching.py (SurfaceMatching's data term over the
    second-order control)."""
    a0 = _f64(a0)
    xt, at = secondOrderEvolution(x0, a0, sigma_diff, timeStep, kernel=kernel_diff, kaff=kaff)
    Ka = applyK(x0, x0, a0, sigma_diff, kernel=kernel_diff)
    if kaff is not None:
        Ka = Ka + affineApplyK(x0, x0, a0, kaff)
    obj = 0.5 * (a0 * Ka).sum()
    if kind == 'L2':
        y = _f64(target)
        dterm, obj0 = -2 * (xt[-1] * y).sum() + (xt[-1] ** 2).sum(), (y ** 2).sum()
    else:
        dterm, obj0 = _data_term(kind, (xt[-1], w0), target, sigma_dist)
    return obj + (obj0 + dterm) / sigma_error ** 2, xt


def second_order_matching_gradient(x0, w0, a0, target, sigma_diff, sigma_dist, timeStep, kind='point_measure',
                                   sigma_error=1., kernel_diff=GAUSS, kaff=None):
    """grad['a0'] = da0 of secondOrderGradient with px1 = -dataTerm'(x_T) /
    sigma_err^2 and pa1 = 0 (base/secondOrderPointSetMatching.py:379-460)."""
    a0 = _f64(a0)
    xt, at = secondOrderEvolution(x0, a0, sigma_diff, timeStep, kernel=kernel_diff, kaff=kaff)
    if kind == 'L2':
        g = 2 * (xt[-1] - _f64(target))
    else:
        g = _data_grad(kind, (xt[-1], w0), target, sigma_dist)
    px1 = -g / sigma_error ** 2
    return secondOrderGradient(x0, a0, px1, np.zeros_like(a0), sigma_diff, timeStep, kernel=kernel_diff,
                               kaff=kaff)[0]
