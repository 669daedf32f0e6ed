"""Surface data terms with the reference's signatures (py-lddmm
base/surfaceDistances.py): varifoldNorm0 / varifoldNormDef / varifoldNorm /
varifoldNormGradient (orientation-free varifold, f0(u) = 1 + u^2),
currentNorm0 / currentNormDef / currentNorm / currentNormGradient /
currentMagnitude / currentMagnitudeGradient, measureNorm0 / measureNormDef /
measureNorm / measureNormGradient, the point-set-target forms
measureNormPS0 / measureNormPSDef / measureNormPS / measureNormPSGradient,
and the internal costs of a velocity field normGrad / diffNormGrad,
elasticNorm / diffElasticNorm, normDisplacement / diffNormDisplacement
(lddmm_surface_internal_cost).

The varifold energy and gradient are single fused GPU passes
(lddmm_surface_varifold: face features, one pair kernel over both surfaces,
face -> vertex gather); the energy total is reduced in float64 on the GPU.
The current and measure terms compose applyK / applyDiffKT on face centres
with the GPU gather.  Precision follows KparDist.pk_dtype, as in the reference's pykeops
branch; `cpu` is accepted and ignored (single device path).
"""
import ctypes

import numpy as np
import torch

from . import _backend as B
from . import kernelFunctions_util as ku


def f0_(u):
    return 1 + u ** 2


def df0_(u):
    return 2 * u


def _check_fun(fun):
    if fun is not None and not (fun[0] is f0_ and fun[1] is df0_):
        raise NotImplementedError("lddmm_amd builds the varifold with f0(u) = 1 + u^2 only")


def _kd(KparDist, dtype=None):
    if getattr(KparDist, 'affine', 'none') not in ('none', None):
        raise NotImplementedError("affine terms do not apply to data-term kernels")
    return B.kernel_for(KparDist.name, KparDist.sigma, KparDist.order, KparDist.ms_exponent,
                        dtype or KparDist.pk_dtype)


def varifold_device(kd, s1, s2=None, energy=True, grad=False, v1=None):
    """Device call: returns (energy tensor [1] or None, gradient tensor [nv,3] or None).

    s1, s2 are Surface objects; v1 optionally overrides s1's vertices with a
    device tensor (the deformed template inside a registration)."""
    d1, keep1 = s1.device(vertices=v1)
    dev = B.device()
    e = torch.empty(1, dtype=torch.float64, device=dev) if energy else None
    g = torch.empty((d1.nv, 3), dtype=torch.float64, device=dev) if grad else None
    if s2 is not None:
        d2, keep2 = s2.device()
        p2 = ctypes.byref(d2)
    else:
        p2 = None
    B.check(B.LIB.lddmm_surface_varifold(kd, ctypes.byref(d1), p2, B.ptr(e), B.ptr(g), B.stream_handle()),
            "surface_varifold")
    return e, g


def varifoldNorm0(fv1, KparDist, fun=None, cpu=False, dtype='float64'):
    """|fv1|^2 (base/surfaceDistances.py:286-302)."""
    _check_fun(fun)
    e, _ = varifold_device(_kd(KparDist), fv1)
    return float(e.item())


def varifoldNormDef(fvDef, fv1, KparDist, fun=None, cpu=False, dtype='float64'):
    """|fvDef|^2 - 2 <fvDef, fv1> (base/surfaceDistances.py:328-353)."""
    _check_fun(fun)
    e, _ = varifold_device(_kd(KparDist), fvDef, fv1)
    return float(e.item())


def varifoldNorm(fvDef, fv1, KparDist, fun=None):
    """|fvDef - fv1|^2 (base/surfaceDistances.py:386-390)."""
    return varifoldNormDef(fvDef, fv1, KparDist, fun=fun) + varifoldNorm0(fv1, KparDist, fun=fun)


def varifoldNormGradient(fvDef, fv1, KparDist, with_weights=False, fun=None, cpu=False, dtype='float64'):
    """Gradient of |fvDef - fv1|^2 wrt fvDef's vertices (base/surfaceDistances.py:394-468).
    with_weights also returns the vertex-weight part (:455-467): per vertex
    the sum over its faces of (2/3) fw_f a_f sum_j K f0 (a1w_j - a2w_j terms),
    i.e. (2/3) times the face's row of |S1|^2 minus its row of <S1, S2>
    (applyktensor rows; the reference keeps fw_f in this factor)."""
    _check_fun(fun)
    _, g = varifold_device(_kd(KparDist), fvDef, fv1, energy=False, grad=True)
    if not with_weights:
        return g.cpu().numpy()
    c1, n1, a1, w1, dk = _measure_data(fvDef)
    c2, n2, a2, w2, _ = _measure_data(fv1)
    a1w, a2w = a1 * w1, a2 * w2
    b1, b2 = n1 * torch.sqrt(a1w)[:, None], n2 * torch.sqrt(a2w)[:, None]
    kw = _ap(KparDist)
    r11 = ku.applyktensor(c1, c1, a1w, a1w, b1, b1, **kw)
    r12 = ku.applyktensor(c1, c2, a1w, a2w, b1, b2, **kw)
    return g.cpu().numpy(), _vertex_sums(dk, (2 / 3) * (r11 - r12)).cpu().numpy()


# ------------------------------------------------------------------ currents
def _face_data(fv, vertices=None):
    """(centers [F,3], face_weights*surfel [F,3], face_weights [F]) on the device."""
    d, keep = fv.device(vertices=vertices)
    dev = B.device()
    c = torch.empty((d.nf, 3), dtype=torch.float64, device=dev)
    ws = torch.empty((d.nf, 3), dtype=torch.float64, device=dev)
    w = torch.empty((d.nf,), dtype=torch.float64, device=dev)
    B.check(B.LIB.lddmm_surface_features(ctypes.byref(d), B.ptr(c), B.ptr(ws), B.ptr(w), B.stream_handle()),
            "surface_features")
    return c, ws, w, (d, keep)


def _gather(dk, dz1, z1):
    d, keep = dk
    px = torch.empty((d.nv, 3), dtype=torch.float64, device=dz1.device)
    dz1_, z1_ = dz1.contiguous(), z1.contiguous()
    B.check(B.LIB.lddmm_surface_face_gather(ctypes.byref(d), B.ptr(dz1_), B.ptr(z1_), B.ptr(px),
                                            B.stream_handle()), "surface_face_gather")
    return px


def _vertex_sums(dk, per_face):
    """Per vertex, the sum over its faces of a per-face scalar (the CSR face
    gather with a zero cross-product term: deterministic order)."""
    dz1 = torch.zeros((per_face.shape[0], 3), dtype=torch.float64, device=per_face.device)
    dz1[:, 0] = per_face
    return _gather(dk, dz1, torch.zeros_like(dz1))[:, 0].contiguous()


def _ap(KparDist):
    return dict(name=KparDist.name, scale=KparDist.sigma, order=KparDist.order,
                dtype=KparDist.pk_dtype, KP=KparDist.ms_exponent)


# ------------------------------------------------------------------ measures
def _measure_data(fv, vertices=None):
    """(centers [F,3], unit surfel S/a [F,3], a [F] with a = sqrt(|S|^2 + 1e-10),
    face weights [F], gather handle) on the device; S is the unweighted surfel."""
    d, keep = fv.device(vertices=vertices)
    dev = B.device()
    c = torch.empty((d.nf, 3), dtype=torch.float64, device=dev)
    S = torch.empty((d.nf, 3), dtype=torch.float64, device=dev)
    w = torch.empty((d.nf,), dtype=torch.float64, device=dev)
    B.check(B.LIB.lddmm_surface_features(ctypes.byref(d), None, None, B.ptr(w), B.stream_handle()),
            "surface_features")
    vw = d.vweights
    d.vweights = None                   # unweighted surfel
    B.check(B.LIB.lddmm_surface_features(ctypes.byref(d), B.ptr(c), B.ptr(S), None, B.stream_handle()),
            "surface_features")
    d.vweights = vw
    a = torch.sqrt((S * S).sum(1) + 1e-10)
    return c, S / a[:, None], a, w, (d, keep)


def measureNorm0(fv1, KparDist):
    """|mu_1|^2 with mu = sum_f a_f w_f delta_{c_f} (base/surfaceDistances.py:208-211)."""
    c2, _, a2, w2, _ = _measure_data(fv1)
    cr2 = (a2 * w2)[:, None]
    return float((cr2 * ku.applyK(c2, c2, cr2, **_ap(KparDist))).sum().item())


def measureNormDef_device(fvDef, fv1, KparDist, vertices=None):
    c1, _, a1, w1, _ = _measure_data(fvDef, vertices)
    c2, _, a2, w2, _ = _measure_data(fv1)
    cr1, cr2 = (a1 * w1)[:, None], (a2 * w2)[:, None]
    kw = _ap(KparDist)
    return (cr1 * ku.applyK(c1, c1, cr1, **kw)).sum() - 2 * (cr1 * ku.applyK(c1, c2, cr2, **kw)).sum()


def measureNormDef(fvDef, fv1, KparDist):
    """|mu_def|^2 - 2 <mu_def, mu_1> (base/surfaceDistances.py:215-223)."""
    return float(measureNormDef_device(fvDef, fv1, KparDist).item())


def measureNorm(fvDef, fv1, KparDist):
    """base/surfaceDistances.py:226-227"""
    return measureNormDef(fvDef, fv1, KparDist) + measureNorm0(fv1, KparDist)


def measureNormGradient_device(fvDef, fv1, KparDist, vertices=None, with_weights=False):
    """z1 = K(c1,c1) a1w - K(c1,c2) a2w, w z1 n gathered with the cross-product
    chain rule, dz1 = (2/3)(DiffKT terms) (base/surfaceDistances.py:231-271);
    with_weights also the vertex sums of (2/3) a z1 (:269-276)."""
    c1, n1, a1, w1, dk = _measure_data(fvDef, vertices)
    c2, _, a2, w2, _ = _measure_data(fv1)
    a1w, a2w = (a1 * w1)[:, None], (a2 * w2)[:, None]
    kw = _ap(KparDist)
    z1 = ku.applyK(c1, c1, a1w, **kw) - ku.applyK(c1, c2, a2w, **kw)
    wz1 = w1[:, None] * z1 * n1
    dz1 = (2 / 3) * (ku.applyDiffKT(c1, c1, a1w, a1w, **kw) - ku.applyDiffKT(c1, c2, a1w, a2w, **kw))
    px = _gather(dk, dz1, wz1)
    if not with_weights:
        return px
    return px, _vertex_sums(dk, (2 / 3) * a1 * z1[:, 0])


def measureNormGradient(fvDef, fv1, KparDist, with_weights=False):
    """base/surfaceDistances.py:231-278"""
    if with_weights:
        px, pxw = measureNormGradient_device(fvDef, fv1, KparDist, with_weights=True)
        return px.cpu().numpy(), pxw.cpu().numpy()
    return measureNormGradient_device(fvDef, fv1, KparDist).cpu().numpy()


def _ps(fv1):
    """Points and weights of a point-set target (`points` or `vertices`, `weights`)."""
    pts = getattr(fv1, 'points', None)
    if pts is None:
        pts = fv1.vertices
    return B.to_dev(pts), B.to_dev(np.asarray(fv1.weights, dtype=np.float64))[:, None]


def measureNormPS0(fv1, KparDist):
    """Point-set target |mu_1|^2 (base/surfaceDistances.py:141-144)."""
    p, w = _ps(fv1)
    return float((w * ku.applyK(p, p, w, **_ap(KparDist))).sum().item())


def measureNormPSDef(fvDef, fv1, KparDist):
    """Surface vs point set (base/surfaceDistances.py:148-155)."""
    c1, _, a1, w1, _ = _measure_data(fvDef)
    p, w = _ps(fv1)
    cr1 = (a1 * w1)[:, None]
    kw = _ap(KparDist)
    return float(((cr1 * ku.applyK(c1, c1, cr1, **kw)).sum() - 2 * (cr1 * ku.applyK(c1, p, w, **kw)).sum()).item())


def measureNormPS(fvDef, fv1, KparDist):
    """base/surfaceDistances.py:158-159"""
    return measureNormPSDef(fvDef, fv1, KparDist) + measureNormPS0(fv1, KparDist)


def measureNormPSGradient(fvDef, fv1, KparDist):
    """base/surfaceDistances.py:163-200; the surface side uses the unweighted
    face measure, as the reference does."""
    c1, n1, a1, _, dk = _measure_data(fvDef)
    p, w = _ps(fv1)
    kw = _ap(KparDist)
    a1c = a1[:, None]
    z1 = (ku.applyK(c1, c1, a1c, **kw) - ku.applyK(c1, p, w, **kw)) * n1
    dz1 = (2 / 3) * (ku.applyDiffKT(c1, c1, a1c, a1c, **kw) - ku.applyDiffKT(c1, p, a1c, w, **kw))
    return _gather(dk, dz1, z1).cpu().numpy()


def currentMagnitude(fv1, KparDist):
    """base/surfaceDistances.py:21-25"""
    c2, cr2, _, _ = _face_data(fv1)
    kw = _ap(KparDist)
    return float((cr2 * ku.applyK(c2, c2, cr2, **kw)).sum().item())


def currentNorm0(fv1, KparDist, weight=1.):
    """base/surfaceDistances.py:66-67"""
    return currentMagnitude(fv1, KparDist)


def currentNormDef(fvDef, fv1, KparDist, weight=1.):
    """base/surfaceDistances.py:71-78"""
    c1, cr1, _, _ = _face_data(fvDef)
    c2, cr2, _, _ = _face_data(fv1)
    kw = _ap(KparDist)
    return float(((cr1 * ku.applyK(c1, c1, cr1, **kw)).sum()
                  - 2 * (cr1 * ku.applyK(c1, c2, cr2, **kw)).sum()).item())


def currentNorm(fvDef, fv1, KparDist, weight=1.):
    """base/surfaceDistances.py:82-83"""
    return currentNormDef(fvDef, fv1, KparDist) + currentNorm0(fv1, KparDist)


def _surfel(dk):
    """Unweighted surfels [F,3] of a device surface view."""
    d, keep = dk
    S = torch.empty((d.nf, 3), dtype=torch.float64, device=B.device())
    vw = d.vweights
    d.vweights = None
    B.check(B.LIB.lddmm_surface_features(ctypes.byref(d), None, B.ptr(S), None, B.stream_handle()),
            "surface_features")
    d.vweights = vw
    return S


def currentNormGradient_device(fvDef, fv1, KparDist, vertices=None, with_weights=False):
    """with_weights also returns the vertex sums of (2/3) S.z1, S the unweighted
    surfel (base/surfaceDistances.py:130-137)."""
    c1, wcr1, w1, dk = _face_data(fvDef, vertices)
    kw = _ap(KparDist)
    z1 = ku.applyK(c1, c1, wcr1, **kw)
    if fv1 is not None:
        c2, wcr2, _, _ = _face_data(fv1)
        z1 = z1 - ku.applyK(c1, c2, wcr2, **kw)
        dz1 = (2 / 3) * (ku.applyDiffKT(c1, c1, wcr1, wcr1, **kw) - ku.applyDiffKT(c1, c2, wcr1, wcr2, **kw))
    else:
        dz1 = (2 / 3) * ku.applyDiffKT(c1, c1, wcr1, wcr1, **kw)
    px = _gather(dk, dz1, w1[:, None] * z1)
    if not with_weights:
        return px
    return px, _vertex_sums(dk, (2 / 3) * (_surfel(dk) * z1).sum(1))


def currentNormGradient(fvDef, fv1, KparDist, weight=1., with_weights=False):
    """base/surfaceDistances.py:87-139"""
    if with_weights:
        px, pxw = currentNormGradient_device(fvDef, fv1, KparDist, with_weights=True)
        return px.cpu().numpy(), pxw.cpu().numpy()
    return currentNormGradient_device(fvDef, fv1, KparDist).cpu().numpy()


def currentMagnitudeGradient(fvDef, KparDist, with_weights=False):
    """base/surfaceDistances.py:28-63 (the current gradient with no target)."""
    if with_weights:
        px, pxw = currentNormGradient_device(fvDef, None, KparDist, with_weights=True)
        return px.cpu().numpy(), pxw.cpu().numpy()
    return currentNormGradient_device(fvDef, None, KparDist).cpu().numpy()


# --------------------------------------------------------------------------
# Internal costs of a velocity field phi on fv (base/surfaceDistances.py:631-872)
# --------------------------------------------------------------------------
def internal_cost_device(fv, phi, costs, energy=True, variables='both', vertices=None):
    """sum_k w_k C_k(fv, phi) for costs = [(name, w), ...]; returns (energy
    tensor [1] or None, {'phi': [nv,3], 'x': [nv,3]} device tensors as
    requested by `variables` ('both', 'phi', 'x' or None)."""
    d = B.internal_cost_desc(costs)
    sv, keep = fv.device(vertices=vertices)
    phi_ = B.to_dev(phi).reshape(sv.nv, 3).contiguous()
    dev = phi_.device
    e = torch.empty(1, dtype=torch.float64, device=dev) if energy else None
    want_phi = variables in ('both', 'phi')
    want_x = variables in ('both', 'x')
    gp = torch.empty((sv.nv, 3), dtype=torch.float64, device=dev) if want_phi else None
    gx = torch.empty((sv.nv, 3), dtype=torch.float64, device=dev) if want_x else None
    B.check(B.LIB.lddmm_surface_internal_cost(ctypes.byref(sv), B.ptr(phi_), ctypes.byref(d), B.ptr(e),
                                              B.ptr(gp), B.ptr(gx), B.stream_handle()), "surface_internal_cost")
    grad = {}
    if want_phi:
        grad['phi'] = gp
    if want_x:
        grad['x'] = gx
    return e, grad


def _ic_norm(name, fv, phi):
    e, _ = internal_cost_device(fv, phi, [(name, 1.0)], variables=None)
    return float(e.item())


def _ic_diff(name, fv, phi, variables):
    _, g = internal_cost_device(fv, phi, [(name, 1.0)], energy=False, variables=variables)
    return {k: v.cpu().numpy() for k, v in g.items()}


def normGrad(fv, phi):
    """Discrete H1 (Dirichlet) energy of phi (base/surfaceDistances.py:631-645)."""
    return _ic_norm('h1', fv, phi)


def diffNormGrad(fv, phi, variables='both'):
    """base/surfaceDistances.py:680-720"""
    return _ic_diff('h1', fv, phi, variables)


def elasticNorm(fv, phi):
    """base/surfaceDistances.py:779-801"""
    return _ic_norm('elastic', fv, phi)


def diffElasticNorm(fv, phi, variables='both'):
    """base/surfaceDistances.py:805-856"""
    return _ic_diff('elastic', fv, phi, variables)


def normDisplacement(fv, phi):
    """base/surfaceDistances.py:731-735"""
    return _ic_norm('displacement', fv, phi)


def diffNormDisplacement(fv, phi, variables='both'):
    """base/surfaceDistances.py:737-767"""
    return _ic_diff('displacement', fv, phi, variables)
