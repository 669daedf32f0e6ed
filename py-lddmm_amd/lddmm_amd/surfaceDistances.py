"""Surface data terms with the reference's signatures (py-lddmm
base/surfaceDistances.py): varifoldNorm0 / varifoldNormDef / varifoldNorm /
varifoldNormGradient (orientation-free varifold, f0(u) = 1 + u^2) and
currentNorm0 / currentNormDef / currentNorm / currentNormGradient /
currentMagnitude / currentMagnitudeGradient.

The varifold energy and gradient are single fused GPU passes
(lddmm_surface_varifold: face features, one pair kernel over both surfaces,
face -> vertex gather); the energy total is reduced in float64 on the GPU.
The current terms compose applyK / applyDiffKT on face centres with the GPU
gather.  Precision follows KparDist.pk_dtype, as in the reference's pykeops
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
    B.kernel_kind(KparDist.name)
    if getattr(KparDist, 'affine', 'none') not in ('none', None):
        raise NotImplementedError("affine terms do not apply to data-term kernels")
    return B.kernel_desc(KparDist.sigma, KparDist.ms_exponent, dtype or KparDist.pk_dtype)


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
    """Gradient of |fvDef - fv1|^2 wrt fvDef's vertices (base/surfaceDistances.py:394-468)."""
    _check_fun(fun)
    if with_weights:
        raise NotImplementedError("varifold gradient wrt vertex weights is not built")
    _, g = varifold_device(_kd(KparDist), fvDef, fv1, energy=False, grad=True)
    return g.cpu().numpy()


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


def _ap(KparDist):
    return dict(name=KparDist.name, scale=KparDist.sigma, order=KparDist.order,
                dtype=KparDist.pk_dtype, KP=KparDist.ms_exponent)


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


def currentNormGradient_device(fvDef, fv1, KparDist, vertices=None):
    c1, wcr1, w1, dk = _face_data(fvDef, vertices)
    c2, wcr2, _, _ = _face_data(fv1)
    kw = _ap(KparDist)
    z1 = ku.applyK(c1, c1, wcr1, **kw) - ku.applyK(c1, c2, wcr2, **kw)
    wz1 = w1[:, None] * z1
    dz1 = (2 / 3) * (ku.applyDiffKT(c1, c1, wcr1, wcr1, **kw) - ku.applyDiffKT(c1, c2, wcr1, wcr2, **kw))
    return _gather(dk, dz1, wz1)


def currentNormGradient(fvDef, fv1, KparDist, weight=1., with_weights=False):
    """base/surfaceDistances.py:87-139"""
    if with_weights:
        raise NotImplementedError("current gradient wrt vertex weights is not built")
    return currentNormGradient_device(fvDef, fv1, KparDist).cpu().numpy()


def currentMagnitudeGradient(fvDef, KparDist, with_weights=False):
    """base/surfaceDistances.py:28-63"""
    if with_weights:
        raise NotImplementedError("current gradient wrt vertex weights is not built")
    c1, wcr1, w1, dk = _face_data(fvDef)
    kw = _ap(KparDist)
    z1 = ku.applyK(c1, c1, wcr1, **kw)
    wz1 = w1[:, None] * z1
    dz1 = (2 / 3) * ku.applyDiffKT(c1, c1, wcr1, wcr1, **kw)
    return _gather(dk, dz1, wz1).cpu().numpy()
