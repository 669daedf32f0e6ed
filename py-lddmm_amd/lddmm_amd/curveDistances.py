"""Curve data terms with the reference's signatures (py-lddmm base/curveDistances.py):
varifoldNorm0 / varifoldNormDef / varifoldNorm / varifoldNormGradient
(f(u) = 1 + weight u^2 on unit segment directions, :186-336) and
currentNorm0 / currentNormDef / currentNorm / currentNormGradient (:30-110,
weight=None).

The varifold energy and gradient are single fused GPU passes
(lddmm_curve_varifold: segment features, one pair kernel over both curves,
segment -> vertex gather), with the same pair kernel as the surface
varifold.  The current terms compose applyK / applyDiffKT on segment
centres with the GPU gather.  Precision follows KparDist.pk_dtype.
"""
import ctypes

import torch

from . import _backend as B
from . import kernelFunctions_util as ku
from .surfaceDistances import _ap, _kd


def varifold_device(kd, c1, c2=None, energy=True, grad=False, v1=None, weight=1.):
    """Device call: (energy tensor [1] or None, gradient tensor [nv, dim] or None).
    c1, c2 are Curve objects; v1 optionally overrides c1's vertices with a
    device tensor (the deformed template inside a registration)."""
    d1, keep1 = c1.device(vertices=v1)
    dev = B.device()
    e = torch.empty(1, dtype=torch.float64, device=dev) if energy else None
    g = torch.empty((d1.nv, d1.dim), dtype=torch.float64, device=dev) if grad else None
    if c2 is not None:
        d2, keep2 = c2.device()
        p2 = ctypes.byref(d2)
    else:
        p2 = None
    B.check(B.LIB.lddmm_curve_varifold(kd, ctypes.byref(d1), p2, float(weight), B.ptr(e), B.ptr(g),
                                       B.stream_handle()), "curve_varifold")
    return e, g


def varifoldNorm0(fv1, KparDist=None, weight=1.):
    """|fv1|^2 (base/curveDistances.py:186-200)."""
    e, _ = varifold_device(_kd(KparDist), fv1, weight=weight)
    return float(e.item())


def varifoldNormDef(fvDef, fv1, KparDist=None, weight=1.):
    """|fvDef|^2 - 2 <fvDef, fv1> (base/curveDistances.py:218-242)."""
    e, _ = varifold_device(_kd(KparDist), fvDef, fv1, weight=weight)
    return float(e.item())


def varifoldNorm(fvDef, fv1, KparDist=None, weight=1.):
    """|fvDef - fv1|^2 (base/curveDistances.py:265-267)."""
    return varifoldNormDef(fvDef, fv1, KparDist, weight) + varifoldNorm0(fv1, KparDist, weight)


def varifoldNormGradient(fvDef, fv1, KparDist=None, weight=1., with_weights=False):
    """Gradient of |fvDef - fv1|^2 wrt fvDef's vertices (base/curveDistances.py:280-336)."""
    if with_weights:
        raise NotImplementedError("curve varifold gradient wrt vertex weights is not built")
    _, g = varifold_device(_kd(KparDist), fvDef, fv1, energy=False, grad=True, weight=weight)
    return g.cpu().numpy()


# ------------------------------------------------------------------ currents
def _seg_data(fv, vertices=None):
    """(centers [S,dim], line_weights*linel [S,dim], line_weights [S]) on the device."""
    d, keep = fv.device(vertices=vertices)
    dev = B.device()
    c = torch.empty((d.ns, d.dim), dtype=torch.float64, device=dev)
    ll = torch.empty((d.ns, d.dim), dtype=torch.float64, device=dev)
    w = torch.empty((d.ns,), dtype=torch.float64, device=dev)
    B.check(B.LIB.lddmm_curve_features(ctypes.byref(d), B.ptr(c), B.ptr(ll), B.ptr(w), B.stream_handle()),
            "curve_features")
    return c, ll * w[:, None], w, (d, keep)


def _seg_gather(dk, dz1, wz1):
    d, keep = dk
    px = torch.empty((d.nv, d.dim), dtype=torch.float64, device=dz1.device)
    dz1_, wz1_ = dz1.contiguous(), wz1.contiguous()
    B.check(B.LIB.lddmm_curve_seg_gather(ctypes.byref(d), B.ptr(dz1_), B.ptr(wz1_), B.ptr(px),
                                         B.stream_handle()), "curve_seg_gather")
    return px


def _no_weight(weight):
    if weight:
        raise NotImplementedError("the curve current's magnitude term (weight) is not built")


def currentNorm0(fv1, KparDist=None, weight=None):
    """base/curveDistances.py:30-37"""
    _no_weight(weight)
    c2, cr2, _, _ = _seg_data(fv1)
    return float((cr2 * ku.applyK(c2, c2, cr2, **_ap(KparDist))).sum().item())


def currentNormDef_device(fvDef, fv1, KparDist, vertices=None):
    c1, cr1, _, _ = _seg_data(fvDef, vertices)
    c2, cr2, _, _ = _seg_data(fv1)
    kw = _ap(KparDist)
    return (cr1 * ku.applyK(c1, c1, cr1, **kw)).sum() - 2 * (cr1 * ku.applyK(c1, c2, cr2, **kw)).sum()


def currentNormDef(fvDef, fv1, KparDist=None, weight=None):
    """base/curveDistances.py:41-52"""
    _no_weight(weight)
    return float(currentNormDef_device(fvDef, fv1, KparDist).item())


def currentNorm(fvDef, fv1, KparDist=None, weight=None):
    """base/curveDistances.py:56-57"""
    return currentNormDef(fvDef, fv1, KparDist, weight) + currentNorm0(fv1, KparDist, weight)


def currentNormGradient_device(fvDef, fv1, KparDist, vertices=None):
    """z1 = 2(K(c1,c1) wl1 - K(c1,c2) wl2), dz1 = DiffKT terms, px[f0] += dz1 - w z1,
    px[f1] += dz1 + w z1 (base/curveDistances.py:61-103)."""
    c1, wcr1, w1, dk = _seg_data(fvDef, vertices)
    c2, wcr2, _, _ = _seg_data(fv1)
    kw = _ap(KparDist)
    z1 = 2 * (ku.applyK(c1, c1, wcr1, **kw) - ku.applyK(c1, c2, wcr2, **kw))
    wz1 = w1[:, None] * z1
    dz1 = (ku.applyDiffKT(c1, c1, wcr1, wcr1, **kw) - ku.applyDiffKT(c1, c2, wcr1, wcr2, **kw))
    return _seg_gather(dk, dz1, wz1)


def currentNormGradient(fvDef, fv1, KparDist=None, weight=None, with_weights=False):
    """base/curveDistances.py:61-110"""
    _no_weight(weight)
    if with_weights:
        raise NotImplementedError("curve current gradient wrt vertex weights is not built")
    return currentNormGradient_device(fvDef, fv1, KparDist).cpu().numpy()
