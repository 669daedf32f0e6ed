"""Point-set data terms with the reference's signatures (py-lddmm
base/pointset_distances.py:3-41): L2 (labelled points) and the measure norm
of weighted point sets, on the GPU kernel sums (kernelFunctions_util)."""
import numpy as np
import torch

from . import _backend as B
from . import kernelFunctions_util as ku
from .surfaceDistances import _ap


def L2Norm0(x1):
    return float((np.asarray(x1.vertices) ** 2).sum())


def L2NormDef(xDef, x1):
    return float(-2 * (xDef * x1).sum() + (xDef ** 2).sum())


def L2NormGradient(xDef, x1):
    return 2 * (xDef - x1)


def _xw(fv, vertices=None):
    x = B.to_dev(fv.vertices if vertices is None else vertices)
    return x, B.to_dev(np.asarray(fv.weights, dtype=np.float64))[:, None]


def measureNorm0(fv1, KparDist):
    """base/pointset_distances.py:14-16"""
    x, w = _xw(fv1)
    return float((w * ku.applyK(x, x, w, **_ap(KparDist))).sum().item())


def measureNormDef_device(fvDef, fv1, KparDist, vertices=None):
    xd, wd = _xw(fvDef, vertices)
    x1, w1 = _xw(fv1)
    kw = _ap(KparDist)
    return (wd * ku.applyK(xd, xd, wd, **kw)).sum() - 2 * (wd * ku.applyK(xd, x1, w1, **kw)).sum()


def measureNormDef(fvDef, fv1, KparDist):
    """base/pointset_distances.py:20-25"""
    return float(measureNormDef_device(fvDef, fv1, KparDist).item())


def measureNorm(fvDef, fv1, KparDist):
    """base/pointset_distances.py:29-30"""
    return measureNormDef(fvDef, fv1, KparDist) + measureNorm0(fv1, KparDist)


def measureNormGradient_device(fvDef, fv1, KparDist, vertices=None):
    xd, wd = _xw(fvDef, vertices)
    x1, w1 = _xw(fv1)
    kw = _ap(KparDist)
    return 2 * (ku.applyDiffKT(xd, xd, wd, wd, **kw) - ku.applyDiffKT(xd, x1, wd, w1, **kw))


def measureNormGradient(fvDef, fv1, KparDist):
    """base/pointset_distances.py:34-41"""
    return measureNormGradient_device(fvDef, fv1, KparDist).cpu().numpy()
