"""Module-level kernel sums (gauss and laplacian) with the reference's signatures.

Drop-in for the dispatch functions of py-lddmm's base/kernelFunctions_util.py
(applyK :443, applyDiffKT :765, applyktensor :1424, applykmat :1510,
applykdiffmat :1631, applyDiffK :646, applyDDiffK :1084, applyDiv :596 and the
applykdiff1/2/1and2, applykdiff11/12/11and12 wrappers :948-955, :1249-1256).
Every call runs on the GPU through the C ABI of liblddmm_amd.so.

Argument conventions follow the reference: `scale` is an array of sigmas,
`KP` the multi-scale exponent, `dtype` selects 'float64' (plain) or
'float32' (compensated FP32) arithmetic.  The `cpu` flag of the reference
selects its numba loops; this backend has one device path and ignores it.
numpy inputs give numpy float64 outputs; CUDA tensors give CUDA float64
tensors (no host round trip).
"""
import numpy as np
import torch

from . import _backend as B


def _is_torch(*arrs):
    return any(isinstance(a, torch.Tensor) for a in arrs)


def _out(t, as_torch):
    return t if as_torch else t.cpu().numpy()


def _pts(a):
    t = B.to_dev(a)
    if t.dim() == 1:
        t = t[:, None]
    return t


def _w(a, n):
    t = B.to_dev(a)
    if t.dim() == 1:
        t = t[:, None]
    width = t.shape[-1] if n == 0 else -1   # reshape(0, -1) is ambiguous
    return t.reshape(n, width).contiguous()


def applyK(y, x, a, name, scale, order, cpu=False, dtype='float64', KP=-1):
    """out[i,:] = sum_s w_s sum_j exp(-|y_i - x_j|^2 / 2 s^2) a[j,:]; shape [N, k]."""
    tt = _is_torch(y, x, a)
    y_, x_ = _pts(y), _pts(x)
    n, m, dim = y_.shape[0], x_.shape[0], y_.shape[1]
    if x_.shape[1] != dim:
        raise ValueError("y and x must have the same dimension")
    a_ = _w(a, m)
    k = a_.shape[1]
    out = torch.empty((n, k), dtype=torch.float64, device=y_.device)
    kd = B.kernel_for(name, scale, order, KP, dtype)
    B.check(B.LIB.lddmm_applyK(kd, B.ptr(y_), n, B.ptr(x_), m, dim, B.ptr(a_), k, B.ptr(out),
                               B.stream_handle()), "applyK")
    return _out(out, tt)


def applyDiffKT(y, x, p, a, name, scale, order, regweight=1., sym=False, lddmm=False, cpu=False,
                dtype='float64', KP=-1):
    """out[i] = sum_s s^(KP-1)/W sum_j -(y_i - x_j)/s K_s(y_i, x_j) c_ij; shape [N, d]."""
    tt = _is_torch(y, x, p, a)
    y_, x_ = _pts(y), _pts(x)
    n, m, dim = y_.shape[0], x_.shape[0], y_.shape[1]
    p_ = _w(p, n)
    a_ = _w(a, m)
    k = a_.shape[1]
    if p_.shape[1] != k:
        raise ValueError("p and a must have the same number of columns")
    mode = B.DIFFKT_LDDMM if lddmm else (B.DIFFKT_SYM if sym else B.DIFFKT_PLAIN)
    if mode != B.DIFFKT_PLAIN and n != m:
        raise ValueError("sym / lddmm modes need as many targets as sources")
    out = torch.empty((n, dim), dtype=torch.float64, device=y_.device)
    kd = B.kernel_for(name, scale, order, KP, dtype)
    B.check(B.LIB.lddmm_applyDiffKT(kd, B.ptr(y_), n, B.ptr(x_), m, dim, B.ptr(p_), B.ptr(a_), k,
                                    mode, float(regweight), B.ptr(out), B.stream_handle()),
            "applyDiffKT")
    return _out(out, tt)


def applyDiv(y, x, a, name, scale, order, cpu=False, dtype='float64', KP=-1):
    """out[i, 0] = sum_s s^(KP-1)/W sum_j -((y_i - x_j)/s).a_j K_s(y_i, x_j): the
    divergence at y_i of the field K(., x) a; shape [N, 1]
    (base/kernelFunctions_util.py:596-620)."""
    tt = _is_torch(y, x, a)
    y_, x_ = _pts(y), _pts(x)
    n, m, dim = y_.shape[0], x_.shape[0], y_.shape[1]
    if x_.shape[1] != dim:
        raise ValueError("y and x must have the same dimension")
    a_ = _w(a, m)
    if a_.shape[1] != dim:
        raise ValueError("a must have one column per coordinate")
    out = torch.empty((n, 1), dtype=torch.float64, device=y_.device)
    kd = B.kernel_for(name, scale, order, KP, dtype)
    B.check(B.LIB.lddmm_applyDiv(kd, B.ptr(y_), n, B.ptr(x_), m, dim, B.ptr(a_), B.ptr(out),
                                 B.stream_handle()), "applyDiv")
    return _out(out, tt)


def applyktensor(y, x, ay, ax, betay, betax, name, scale, order, cpu=False, dtype='float64', KP=-1):
    """out[i] = sum_s w_s sum_j K_s(y_i, x_j)(ay_i ax_j + (betay_i . betax_j)^2); shape [N]."""
    tt = _is_torch(y, x, ay, ax, betay, betax)
    y_, x_ = _pts(y), _pts(x)
    n, m, dim = y_.shape[0], x_.shape[0], y_.shape[1]
    ay_ = B.to_dev(ay).reshape(n).contiguous()
    ax_ = B.to_dev(ax).reshape(m).contiguous()
    by_ = _w(betay, n)
    bx_ = _w(betax, m)
    kb = by_.shape[1]
    out = torch.empty((n,), dtype=torch.float64, device=y_.device)
    kd = B.kernel_for(name, scale, order, KP, dtype)
    B.check(B.LIB.lddmm_applyktensor(kd, B.ptr(y_), n, B.ptr(x_), m, dim, B.ptr(ay_), B.ptr(ax_),
                                     B.ptr(by_), B.ptr(bx_), kb, B.ptr(out), B.stream_handle()),
            "applyktensor")
    return _out(out, tt)


def applykmat(y, x, beta, name, scale, order, KP=-1):
    """out[i,:] = sum_s w_s sum_j K_s(y_i, x_j) beta[i,j,:] (dense beta [N, M, k])."""
    tt = _is_torch(y, x, beta)
    y_, x_ = _pts(y), _pts(x)
    n, m, dim = y_.shape[0], x_.shape[0], y_.shape[1]
    b_ = B.to_dev(beta)
    if b_.dim() == 2:
        b_ = b_[..., None]
    b_ = b_.reshape(n, m, -1).contiguous()
    kb = b_.shape[2]
    out = torch.empty((n, kb), dtype=torch.float64, device=y_.device)
    kd = B.kernel_for(name, scale, order, KP, 'float64')
    B.check(B.LIB.lddmm_applykmat(kd, B.ptr(y_), n, B.ptr(x_), m, dim, B.ptr(b_), kb, B.ptr(out),
                                  B.stream_handle()), "applykmat")
    return _out(out, tt)


def applykdiffmat(y, x, beta, name, scale, order, cpu=False, dtype='float64', KP=-1):
    """out[i] = -sum_s s^(KP-1)/W sum_j K_s(y_i, x_j)(y_i - x_j)/s beta[i,j] (dense [N, M])."""
    tt = _is_torch(y, x, beta)
    y_, x_ = _pts(y), _pts(x)
    n, m, dim = y_.shape[0], x_.shape[0], y_.shape[1]
    b_ = B.to_dev(beta).reshape(n, m).contiguous()
    out = torch.empty((n, dim), dtype=torch.float64, device=y_.device)
    kd = B.kernel_for(name, scale, order, KP, 'float64')
    B.check(B.LIB.lddmm_applykdiffmat(kd, B.ptr(y_), n, B.ptr(x_), m, dim, B.ptr(b_), B.ptr(out),
                                      B.stream_handle()), "applykdiffmat")
    return _out(out, tt)


_DIFFK_OPT = {'1': B.DIFFK_1, '2': B.DIFFK_2, '1and2': B.DIFFK_1AND2}
_DDIFFK_OPT = {'11': B.DDIFFK_11, '12': B.DDIFFK_12, '11and12': B.DDIFFK_11AND12}


def applyDiffK(x, a1, a2, name, scale, order, cpu=False, dtype='float64', KP=-1, option='1'):
    """First-derivative sums on one point set (options '1', '2', '1and2')."""
    if option not in _DIFFK_OPT:
        raise ValueError(f"DiffK: unknown option {option!r}")
    tt = _is_torch(x, a1, a2)
    x_ = _pts(x)
    n, dim = x_.shape
    a1_ = _w(a1, n)
    if a1_.shape[1] != dim:
        raise ValueError("a1 must have the point dimension")
    a2_ = _w(a2, n)
    k = a2_.shape[1]
    out = torch.empty((n, k), dtype=torch.float64, device=x_.device)
    kd = B.kernel_for(name, scale, order, KP, dtype)
    B.check(B.LIB.lddmm_applyDiffK(kd, B.ptr(x_), n, dim, B.ptr(a1_), B.ptr(a2_), k,
                                   _DIFFK_OPT[option], B.ptr(out), B.stream_handle()), "applyDiffK")
    return _out(out, tt)


def applyDDiffK(x, a1, a2, p, name, scale, order, cpu=False, dtype='float64', KP=-1, option='11'):
    """Second-derivative sums on one point set (options '11', '12', '11and12')."""
    if option not in _DDIFFK_OPT:
        raise ValueError(f"DDiffK: unknown option {option!r}")
    tt = _is_torch(x, a1, a2, p)
    x_ = _pts(x)
    n, dim = x_.shape
    a1_ = _w(a1, n)
    a2_ = _w(a2, n)
    if a1_.shape[1] != a2_.shape[1]:
        raise ValueError("a1 and a2 must have the same number of columns")
    p_ = _w(p, n)
    if p_.shape[1] != dim:
        raise ValueError("p must have the point dimension")
    out = torch.empty((n, dim), dtype=torch.float64, device=x_.device)
    kd = B.kernel_for(name, scale, order, KP, dtype)
    B.check(B.LIB.lddmm_applyDDiffK(kd, B.ptr(x_), n, dim, B.ptr(a1_), B.ptr(a2_), a1_.shape[1],
                                    B.ptr(p_), _DDIFFK_OPT[option], B.ptr(out), B.stream_handle()),
            "applyDDiffK")
    return _out(out, tt)


# Reference wrappers (base/kernelFunctions_util.py:948-955, :1249-1256); the
# first-derivative ones always pass KP = -1, as the reference does.
def applykdiff1(x, a1, a2, name, scale, order, KP=-1, dtype='float64'):
    return applyDiffK(x, a1, a2, name, scale, order, KP=-1, option='1', dtype=dtype)


def applykdiff2(x, a1, a2, name, scale, order, KP=-1, dtype='float64'):
    return applyDiffK(x, a1, a2, name, scale, order, KP=-1, option='2', dtype=dtype)


def applykdiff1and2(x, a1, a2, name, scale, order, KP=-1, dtype='float64'):
    return applyDiffK(x, a1, a2, name, scale, order, KP=-1, option='1and2', dtype=dtype)


def applykdiff11(x, a1, a2, p, name, scale, order, KP=-1, dtype='float64'):
    return applyDDiffK(x, a1, a2, p, name, scale, order, KP=KP, option='11', dtype=dtype)


def applykdiff12(x, a1, a2, p, name, scale, order, KP=-1, dtype='float64'):
    return applyDDiffK(x, a1, a2, p, name, scale, order, KP=KP, option='12', dtype=dtype)


def applykdiff11and12(x, a1, a2, p, name, scale, order, KP=-1, dtype='float64'):
    return applyDDiffK(x, a1, a2, p, name, scale, order, KP=KP, option='11and12', dtype=dtype)
