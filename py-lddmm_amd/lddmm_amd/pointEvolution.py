"""Landmark evolution and adjoints with the reference's signatures
(py-lddmm base/pointEvolution.py): landmarkDirectEvolutionEuler,
landmarkHamiltonianCovector, landmarkHamiltonianGradient, landmarkEPDiff,
secondOrderEvolution, secondOrderHamiltonian, secondOrderCovector,
secondOrderGradient.

The T-step loops run on the GPU (lddmm_shoot_* / lddmm_epdiff_*): the
trajectory stays in HBM and only the requested outputs come back.  numpy
inputs return numpy arrays; CUDA tensors return CUDA tensors.  A kernel's
affine part (Kernel(affine='affine'|'euclidean')) is applied in first-order
and geodesic (EPDiff) shooting and their adjoints.  The log-Jacobian along the trajectory
(withJacobian) is tracked on the device by the kernel divergence.  Affine
controls, normals, passenger point sets and splines are not built
(NotImplementedError); the registration drivers here do not use them.
"""
import ctypes

import numpy as np
import torch

from . import _backend as B
from . import kernelFunctions_util as ku


def _kd(K):
    """Kernel descriptor, with the kernel's affine part (applied by the
    first-order and geodesic shooting loops)."""
    return B.kernel_of(K, with_affine=True)


def _is_torch(*a):
    return any(isinstance(v, torch.Tensor) for v in a)


def _ret(t, tt):
    return t if tt else t.cpu().numpy()


def _no_affine(affine):
    if affine is not None and len(affine) > 0 and affine[0] is not None and len(affine[0]) > 0:
        raise NotImplementedError("affine controls are not built in lddmm_amd shooting")


def _ic(icost):
    """icost = (surface, [(name, weight), ...]): the template's triangulation
    and the internal-cost list."""
    surf, costs = icost
    sv, keep = surf.device()
    return sv, B.internal_cost_desc(costs), keep


def shoot(x0, at, KparDiff, regweight=1.0, energy=False, icost=None):
    """Device forward flow: returns (xt [T+1,N,d] tensor, energy tensor [1] or None).
    With icost = (surface, costs) the energy includes sum_t dt C(x_t, K(x_t) a_t)."""
    at_ = B.to_dev(at)
    T, n, dim = at_.shape
    xt = torch.empty((T + 1, n, dim), dtype=torch.float64, device=at_.device)
    xt[0] = B.to_dev(x0)
    e = torch.empty(1, dtype=torch.float64, device=at_.device) if energy else None
    if icost is None:
        B.check(B.LIB.lddmm_shoot_forward(_kd(KparDiff), B.ptr(at_), T, n, dim, B.ptr(xt), float(regweight),
                                          B.ptr(e), B.stream_handle()), "shoot_forward")
    else:
        sv, d, keep = _ic(icost)
        if sv.nv != n or dim != 3:
            raise ValueError("internal costs need the template's surface (same vertex count, dim 3)")
        B.check(B.LIB.lddmm_shoot_forward_ic(_kd(KparDiff), ctypes.byref(sv), ctypes.byref(d), B.ptr(at_), T,
                                             B.ptr(xt), float(regweight), B.ptr(e), B.stream_handle()),
                "shoot_forward_ic")
    return xt, e


def shoot_adjoint(xt, at, px1, KparDiff, regweight, euclidean=True, covector=False, icost=None):
    """Device adjoint: returns (grad [T,N,d] (already / T), pxt or None).
    With icost = (surface, costs) the covector carries the costs' extra term."""
    at_ = B.to_dev(at)
    T, n, dim = at_.shape
    xt_ = B.to_dev(xt)
    px1_ = B.to_dev(px1)
    g = torch.empty_like(at_)
    pxt = torch.empty((T + 1, n, dim), dtype=torch.float64, device=at_.device) if covector else None
    if icost is None:
        B.check(B.LIB.lddmm_shoot_adjoint(_kd(KparDiff), B.ptr(at_), B.ptr(xt_), T, n, dim, B.ptr(px1_),
                                          float(regweight), 1 if euclidean else 0, B.ptr(pxt), B.ptr(g),
                                          B.stream_handle()), "shoot_adjoint")
    else:
        sv, d, keep = _ic(icost)
        if sv.nv != n or dim != 3:
            raise ValueError("internal costs need the template's surface (same vertex count, dim 3)")
        B.check(B.LIB.lddmm_shoot_adjoint_ic(_kd(KparDiff), ctypes.byref(sv), ctypes.byref(d), B.ptr(at_),
                                             B.ptr(xt_), T, B.ptr(px1_), float(regweight),
                                             1 if euclidean else 0, B.ptr(pxt), B.ptr(g), B.stream_handle()),
                "shoot_adjoint_ic")
    return g, pxt


def jacobian(xt, at, KparDiff, T=None):
    """Log-Jacobian along a trajectory: J_0 = 0, J_{t+1} = J_t + dt div v_t(x_t)
    with v_t = K(., x_t) a_t and dt = 1/T (base/pointEvolution.py:149-152,
    :366-367, :459-463).  xt [>=T, N, d], at [>=T, N, d] device tensors;
    returns [T+1, N, 1] float64 on the device."""
    T = at.shape[0] if T is None else T
    n = xt.shape[1]
    Jt = torch.zeros((T + 1, n, 1), dtype=torch.float64, device=xt.device)
    dt = 1.0 / T
    for t in range(T):
        Jt[t + 1] = Jt[t] + dt * KparDiff.applyDivergence(xt[t], at[t])
    return Jt


def landmarkDirectEvolutionEuler(x0, at, KparDiff, affine=None, options=None):
    """x_{t+1} = x_t + (1/T) K(x_t) a_t   (base/pointEvolution.py:88-164)."""
    _no_affine(affine)
    opt = {'withJacobian': False, 'withNormals': None, 'withPointSet': None}
    if options:
        opt.update(options)
    if opt['withNormals'] is not None or opt['withPointSet'] is not None:
        raise NotImplementedError("normals / point-set tracking are not built")
    tt = _is_torch(x0, at)
    at_ = B.to_dev(at)
    xt, _ = shoot(x0, at_, KparDiff)
    Jt = _ret(jacobian(xt, at_, KparDiff), tt) if opt['withJacobian'] else None
    return {'xt': _ret(xt, tt), 'Jt': Jt, 'yt': None, 'nt': None}


def landmarkHamiltonianCovector(x0, at, px1, Kpardiff, regweight, affine=None, extraTerm=None):
    """Backward covector p_t (base/pointEvolution.py:167-201); returns (pxt, xt)."""
    _no_affine(affine)
    if extraTerm is not None:
        raise NotImplementedError("extra (internal-cost) terms are not built")
    tt = _is_torch(x0, at, px1)
    xt, _ = shoot(x0, at, Kpardiff)
    _, pxt = shoot_adjoint(xt, at, px1, Kpardiff, regweight, euclidean=False, covector=True)
    return _ret(pxt, tt), _ret(xt, tt)


def landmarkHamiltonianGradient(x0, at, px1, KparDiff, regweight, getCovector=False, affine=None,
                                extraTerm=None):
    """dat_t = 2 r a_t - p_{t+1}  (base/pointEvolution.py:205-236)."""
    _no_affine(affine)
    if extraTerm is not None:
        raise NotImplementedError("extra (internal-cost) terms are not built")
    tt = _is_torch(x0, at, px1)
    at_ = B.to_dev(at)
    xt, _ = shoot(x0, at_, KparDiff)
    _, pxt = shoot_adjoint(xt, at_, px1, KparDiff, regweight, euclidean=False, covector=True)
    dat = 2 * regweight * at_ - pxt[1:]
    if getCovector:
        return _ret(dat, tt), _ret(xt, tt), _ret(pxt, tt)
    return _ret(dat, tt), _ret(xt, tt)


# --------------------------------------------------------------------------
# Second order (geodesic shooting)
# --------------------------------------------------------------------------
def _epdiff(x0, a0, KparDiff, T):
    x0_ = B.to_dev(x0)
    n, dim = x0_.shape
    xt = torch.empty((T + 1, n, dim), dtype=torch.float64, device=x0_.device)
    at = torch.empty_like(xt)
    xt[0] = x0_
    at[0] = B.to_dev(a0)
    B.check(B.LIB.lddmm_epdiff_forward(_kd(KparDiff), T, n, dim, B.ptr(xt), B.ptr(at), B.stream_handle()),
            "epdiff_forward")
    return xt, at


def landmarkEPDiff(T, x0, a0, KparDiff, affine=None, withJacobian=False, withNormals=None, withPointSet=None):
    """Geodesic shooting of (x0, a0) over T steps (base/pointEvolution.py:320-379).

    Returns (xt [T+1], at [T]); at[T-1] is the last momentum used (the
    reference does not update it past T-1).  With withJacobian, returns
    [xt, at, Jt] with Jt [T+1, N, 1] (the reference's list output; its
    update writes the [N, 1] divergence into an [N] slice, here it is
    flattened as in secondOrderEvolution)."""
    _no_affine(affine)
    if withNormals is not None or withPointSet is not None:
        raise NotImplementedError("normals / point-set tracking are not built")
    tt = _is_torch(x0, a0)
    xt, at = _epdiff(x0, a0, KparDiff, T)
    if withJacobian:
        return [_ret(xt, tt), _ret(at[:T], tt), _ret(jacobian(xt, at, KparDiff, T), tt)]
    return _ret(xt, tt), _ret(at[:T], tt)


def geodesicHamiltonian(xt, at, KparDiff):
    """H_t = 1/2 a_t . K(x_t) a_t along a shooting trajectory, in float64.

    The geodesic (EPDiff) flow conserves H up to the O(dt) drift of the Euler
    scheme; the check runs the kernel sums in float64 (and the reductions on
    the device in float64) whatever KparDiff.pk_dtype the trajectory was shot
    with.  xt, at: [T(+1), N, d] arrays (the first min(len) steps are used).
    Returns (H [T] numpy, max_t |H_t - H_0| / |H_0|).  No reference counterpart
    (the reference shoots without checking H)."""
    from .kernelFunctions_util import applyK
    xt_, at_ = B.to_dev(xt), B.to_dev(at)
    T = min(xt_.shape[0], at_.shape[0])
    H = torch.empty(T, dtype=torch.float64, device=xt_.device)
    for t in range(T):
        Ka = applyK(xt_[t], xt_[t], at_[t], KparDiff.name, KparDiff.sigma, KparDiff.order, dtype='float64',
                    KP=KparDiff.ms_exponent)
        H[t] = 0.5 * (at_[t] * Ka).sum()
    h = H.cpu().numpy()
    drift = float(np.max(np.abs(h - h[0])) / max(abs(h[0]), 1e-300)) if T else 0.0
    return h, drift


def secondOrderEvolution(x0, a0, KparDiff, timeStep, withSpline=None, affine=None, options=None):
    """base/pointEvolution.py:385-472 (no spline, no affine)."""
    if withSpline is not None:
        raise NotImplementedError("spline controls are not built")
    _no_affine(affine)
    if options and options.get('withPointSet') is not None:
        raise NotImplementedError("point-set tracking is not built")
    T = int(np.round(1 / timeStep))
    tt = _is_torch(x0, a0)
    xt, at = _epdiff(x0, a0, KparDiff, T)
    Jt = _ret(jacobian(xt, at, KparDiff, T), tt) if options and options.get('withJacobian') else None
    return {'xt': _ret(xt, tt), 'at': _ret(at, tt), 'Jt': Jt, 'yt': None}


def secondOrderHamiltonian(x, a, px, pa, KparDiff, withSpline=None, affine=None):
    """H = <px, K(x)a> - <pa, DiffKT(x, a, a)>  (base/pointEvolution.py:475-486); float64."""
    if withSpline is not None:
        raise NotImplementedError("spline controls are not built")
    _no_affine(affine)
    kw = dict(name=KparDiff.name, scale=KparDiff.sigma, order=KparDiff.order, dtype=KparDiff.pk_dtype,
              KP=KparDiff.ms_exponent)
    x_, a_ = B.to_dev(x), B.to_dev(a)
    if getattr(KparDiff, 'affine', 'none') in ('affine', 'euclidean'):
        # Kernel.applyK / applyDiffKT add the kernel's affine part
        Ka, Da = KparDiff.applyK(x_, a_), KparDiff.applyDiffKT(x_, a_, a_)
    else:
        Ka, Da = ku.applyK(x_, x_, a_, **kw), ku.applyDiffKT(x_, x_, a_, a_, **kw)
    Ht = (B.to_dev(px) * Ka).sum() - (B.to_dev(pa) * Da).sum()
    return float(Ht.item())


def secondOrderCovector(x0, a0, px1, pa1, KparDiff, timeStep, withSpline=None, affine=None, isjump=None,
                        forwardState=None):
    """Backward (px, pa) covectors (base/pointEvolution.py:489-552, single target)."""
    if withSpline is not None or isjump is not None:
        raise NotImplementedError("splines / multiple targets are not built")
    _no_affine(affine)
    if isinstance(px1, (list, tuple)):
        if len(px1) != 1:
            raise NotImplementedError("multiple targets are not built")
        px1, pa1 = px1[0], pa1[0]
    T = int(np.round(1 / timeStep))
    tt = _is_torch(x0, a0, px1, pa1)
    if forwardState is None:
        xt, at = _epdiff(x0, a0, KparDiff, T)
    else:
        xt, at = B.to_dev(forwardState[0]), B.to_dev(forwardState[1])
    pxt, pat, da0 = _epdiff_adjoint(xt, at, px1, pa1, KparDiff, T)
    return _ret(pxt, tt), _ret(pat, tt), _ret(xt, tt), _ret(at, tt)


def _epdiff_adjoint(xt, at, px1, pa1, KparDiff, T):
    n, dim = xt.shape[1], xt.shape[2]
    pxt = torch.empty_like(xt)
    pat = torch.empty_like(xt)
    da0 = torch.empty((n, dim), dtype=torch.float64, device=xt.device)
    # device copies stay referenced until the launch is enqueued (a freed
    # temporary would be handed to the next allocation before the kernel reads it)
    px1_, pa1_ = B.to_dev(px1), B.to_dev(pa1)
    B.check(B.LIB.lddmm_epdiff_adjoint(_kd(KparDiff), T, n, dim, B.ptr(xt), B.ptr(at), B.ptr(px1_),
                                       B.ptr(pa1_), B.ptr(pxt), B.ptr(pat), B.ptr(da0),
                                       B.stream_handle()), "epdiff_adjoint")
    return pxt, pat, da0


def secondOrderGradient(x0, a0, px1, pa1, KparDiff, timeStep, isjump=None, withSpline=None, affine=None,
                        controlWeight=1.0):
    """da0 = K(x0) a0 - pa_0  (base/pointEvolution.py:555-618)."""
    if withSpline is not None or isjump is not None:
        raise NotImplementedError("splines / multiple targets are not built")
    _no_affine(affine)
    T = int(np.round(1 / timeStep))
    tt = _is_torch(x0, a0, px1, pa1)
    xt, at = _epdiff(x0, a0, KparDiff, T)
    pxt, pat, da0 = _epdiff_adjoint(xt, at, px1, pa1, KparDiff, T)
    return {'da0': _ret(da0, tt), 'drhot': None, 'xt': _ret(xt, tt), 'at': _ret(at, tt), 'dA': None,
            'db': None, 'pxt': _ret(pxt, tt), 'pat': _ret(pat, tt)}
