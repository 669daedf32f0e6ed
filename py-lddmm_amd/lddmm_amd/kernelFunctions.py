"""Kernel objects with the reference's API (py-lddmm base/kernelFunctions.py:339-674).

`Kernel(name='gauss', sigma=..., affine='none'|'affine'|'euclidean', w1, w2,
center, ...)` exposes applyK / applyKTensor / applyDiffKT / applyDiffK /
applyDiffK2 / applyDiffK1and2 / applyDiffKmat / applyDDiffK11and12 with the
same signatures.  The Gaussian sums run on the GPU (kernelFunctions_util);
the affine / Euclidean parts are the reference's O(N d^2) corrections,
computed with the array library of the inputs (numpy or torch).
`pk_dtype` ('float64' or 'float32') selects the plain FP64 path or the
compensated FP32 path, as the reference's pykeops flag does.
"""
import logging

import numpy as np
import torch

from . import kernelFunctions_util as ku


def _sum0(a):
    return a.sum(dim=0) if isinstance(a, torch.Tensor) else a.sum(axis=0)


def _like(v, ref):
    """Convert a numpy constant array to the array type of `ref`."""
    if isinstance(ref, torch.Tensor):
        return torch.as_tensor(np.asarray(v), dtype=ref.dtype, device=ref.device)
    return np.asarray(v)


def _copy(x):
    return x.clone() if isinstance(x, torch.Tensor) else np.copy(x)


class KernelSpec:
    """Kernel specification (base/kernelFunctions.py:337-409)."""

    def __init__(self, name='gauss', affine='none', sigma=(1.,), order=3, w1=1.0,
                 w2=1.0, dim=3, weight=1.0, localMaps=None):
        self.name = name
        if np.isscalar(sigma):
            self.sigma = np.array([sigma], dtype=float)
        else:
            self.sigma = np.array(sigma, dtype=float)
        self.order = order
        self.weight = weight
        self.constant_plane = False
        self.affine_basis = []
        self.dim = dim
        self.precomp = []
        self._hold = False
        self._state = False
        self.setAffine(affine, w1, w2)
        self.localMaps = localMaps
        if localMaps:
            raise NotImplementedError("lddmm_amd: local kernel maps are not built")
        self.kff = False
        self.pk_dtype = 'float64'
        self.ms_exponent = -1
        if name == 'gauss' or name == 'gaussian':
            self.order = 10
            self.par = [sigma]
            self.kff = True
        elif name == 'laplacian':
            # orders above 4 fall back to 3, as in the reference
            if self.order > 4:
                self.order = 3
            self.par = [sigma, self.order]
            self.kff = True
        elif name == 'min':
            self.par = []
        elif 'poly' in name:
            self.name = 'poly'
            self.par = [sigma, self.order]
        elif name == 'euclidean':
            self.par = [sigma]
        elif name == 'constant':
            self.par = []
        else:
            self.name = 'none'
            self.par = []

    def setAffine(self, affine, w1=1., w2=1., center=None):
        self.affine = affine
        self.w1 = w1
        self.w2 = w2
        self.center = np.zeros(self.dim) if center is None else np.array(center)
        self.affine_basis = []
        if self.affine == 'euclidean':
            s2 = np.sqrt(2.0)
            if self.dim == 3:
                self.affine_basis.append(np.array([[0, 1, 0], [-1, 0, 0], [0, 0, 0]]) / s2)
                self.affine_basis.append(np.array([[0, 0, 1], [0, 0, 0], [-1, 0, 0]]) / s2)
                self.affine_basis.append(np.array([[0, 0, 0], [0, 0, 1], [0, -1, 0]]) / s2)
            elif self.dim == 2:
                self.affine_basis.append(np.array([[0, 1], [-1, 0]]) / s2)
            else:
                logging.info('Warning: Euclidean kernels only available in dimensions 2 or 3')
                self.affine = 'affine'


class Kernel(KernelSpec):
    """Gaussian kernel operator (base/kernelFunctions.py:415-674)."""

    def hold(self):
        self._state = self._hold
        self._hold = True

    def release(self):
        self._state = False
        self._hold = False

    def reset(self):
        self._hold = self._state

    def getK(self, x, firstVar=None):
        raise NotImplementedError("lddmm_amd: dense kernel matrices (getK) are not built")

    # K(x,x)a or K(y,x)a  (base/kernelFunctions.py:452-506)
    def applyK(self, x, a, firstVar=None, grid=None, matrixWeights=False, cpu=False):
        y = _copy(x) if firstVar is None else firstVar
        if matrixWeights:
            z = ku.applykmat(y, x, a, self.name, self.sigma, self.order, KP=self.ms_exponent)
        else:
            z = ku.applyK(y, x, a, self.name, self.sigma, self.order, cpu=cpu, dtype=self.pk_dtype,
                          KP=self.ms_exponent)
        if self.affine == 'affine':
            xx = x - _like(self.center, x)
            if firstVar is None:
                yy = xx if grid is None else grid - _like(self.center, x)
            else:
                yy = firstVar - _like(self.center, x)
            z = z + self.w1 * (yy @ (xx.T @ a)) + self.w2 * _sum0(a)
        elif self.affine == 'euclidean':
            xx = x - _like(self.center, x)
            z = z + self.w2 * _sum0(a)
            for E in self.affine_basis:
                E_ = _like(E, x)
                xE = xx @ E_.T
                if firstVar is None:
                    if grid is None:
                        z = z + self.w1 * (xE * a).sum() * xE
                    else:
                        z = z + self.w1 * (xE * a).sum() * ((grid - _like(self.center, x)) @ E_.T)
                else:
                    yE = (firstVar - _like(self.center, x)) @ E_.T
                    z = z + self.w1 * (xE * a).sum() * yE
        return z

    # (base/kernelFunctions.py:509-517)
    def applyKTensor(self, x, ay, ax, betay, betax, firstVar=None, grid=None, cpu=False):
        y = _copy(x) if firstVar is None else firstVar
        return ku.applyktensor(y, x, ay, ax, betay, betax, self.name, self.sigma, self.order, cpu=cpu,
                               dtype=self.pk_dtype, KP=self.ms_exponent)

    # A(i) = sum_j D_1[K(x(i), x(j))a2(j)]a1(i)  (base/kernelFunctions.py:520-531)
    def applyDiffK(self, x, a1, a2):
        z = ku.applykdiff1(x, a1, a2, self.name, self.sigma, self.order, KP=self.ms_exponent,
                           dtype=self.pk_dtype)
        if self.affine == 'affine':
            xx = x - _like(self.center, x)
            z = z + self.w1 * (xx * a2).sum() * a1
        elif self.affine == 'euclidean':
            xx = x - _like(self.center, x)
            for E in self.affine_basis:
                E_ = _like(E, x)
                z = z + self.w1 * (xx * (a2 @ E_)).sum() * (a1 @ E_)
        return z

    # A(i) = sum_j D_2[K(x(i), x(j))a2(j)]a1(j)  (base/kernelFunctions.py:534-545)
    def applyDiffK2(self, x, a1, a2):
        z = ku.applykdiff2(x, a1, a2, self.name, self.sigma, self.order, KP=self.ms_exponent,
                           dtype=self.pk_dtype)
        if self.affine == 'affine':
            xx = x - _like(self.center, x)
            z = z + self.w1 * (a1 * a2).sum() * xx
        elif self.affine == 'euclidean':
            xx = x - _like(self.center, x)
            for E in self.affine_basis:
                E_ = _like(E, x)
                z = z + self.w1 * (a1 * (a2 @ E_)).sum() * (xx @ E_.T)
        return z

    # (base/kernelFunctions.py:549-562)
    def applyDiffK1and2(self, x, a1, a2):
        z = ku.applykdiff1and2(x, a1, a2, self.name, self.sigma, self.order, KP=self.ms_exponent,
                               dtype=self.pk_dtype)
        if self.affine == 'affine':
            xx = x - _like(self.center, x)
            z = z + self.w1 * (a1 @ (xx.T @ a2) + (xx @ (a1.T @ a2)))
        elif self.affine == 'euclidean':
            xx = x - _like(self.center, x)
            for E in self.affine_basis:
                E_ = _like(E, x)
                Exx = xx @ E_.T
                Ea1 = a1 @ E_.T
                ETa2 = a2 @ E_
                z = z + self.w1 * ((xx * ETa2).sum() * Ea1 + (a1 * ETa2).sum() * Exx)
        return z

    # (base/kernelFunctions.py:565-570)
    def applyDiffKmat(self, x, beta, firstVar=None):
        y = x if firstVar is None else firstVar
        return ku.applykdiffmat(y, x, beta, self.name, self.sigma, self.order, KP=self.ms_exponent)

    # sum_k sum_j nabla_1[a1(k,i). K(x(i), x(j))a2(k,j)]  (base/kernelFunctions.py:581-626)
    def applyDiffKT(self, x, p0, a, firstVar=None, regweight=1., lddmm=False, sym=False,
                    extra_term=None, cpu=False):
        y = _copy(x) if firstVar is None else firstVar
        p = p0 + extra_term if (lddmm and extra_term is not None) else p0
        zpx = ku.applyDiffKT(y, x, p, a, self.name, self.sigma, self.order, regweight=regweight,
                             lddmm=lddmm, sym=sym, cpu=cpu, dtype=self.pk_dtype, KP=self.ms_exponent)
        if self.affine == 'affine':
            xx = x - _like(self.center, x)
            zpx = zpx + self.w1 * (p @ (a.T @ xx))
            if lddmm:
                zpx = zpx + self.w1 * (a @ (p.T @ xx) - 2 * regweight * a @ (a.T @ xx))
            elif sym:
                zpx = zpx + self.w1 * (a @ (p.T @ xx))
        elif self.affine == 'euclidean':
            xx = x - _like(self.center, x)
            for E in self.affine_basis:
                E_ = _like(E, x)
                ETp = p @ E_
                ETa = a @ E_
                zpx = zpx + self.w1 * (xx * ETa).sum() * ETp
                if lddmm:
                    zpx = zpx + self.w1 * ((xx * ETp).sum() * ETa - 2 * regweight * (xx * ETa).sum() * ETa)
                elif sym:
                    zpx = zpx + self.w1 * (xx * ETp).sum() * ETa
        return zpx

    # divergence of the field K(., x) a at y  (base/kernelFunctions.py:638-663)
    def applyDivergence(self, x, a, firstVar=None):
        y = x if firstVar is None else firstVar
        zpx = ku.applyDiv(y, x, a, self.name, self.sigma, self.order, dtype=self.pk_dtype,
                          KP=self.ms_exponent)
        if self.affine == 'affine':
            # div_y sum_j w1 ((y-c).(x_j-c)) a_j = w1 sum_j (x_j-c).a_j
            xx = x - _like(self.center, x)
            zpx = zpx + self.w1 * (xx * a).sum()
        # 'euclidean': the basis matrices are skew (trace 0), the affine part adds nothing
        return zpx

    def testDiffKT(self, y, x, p, a):
        """Finite-difference check of applyDiffKT (base/kernelFunctions.py:629-635)."""
        k0 = (p * self.applyK(x, a, firstVar=y)).sum()
        dy = np.random.randn(y.shape[0], y.shape[1])
        eps = 1e-8
        k1 = (p * self.applyK(x, a, firstVar=y + eps * dy)).sum()
        grd = self.applyDiffKT(x, p, a, firstVar=y)
        logging.info('test diffKT: {0:.5f}, {1:.5f}'.format((k1 - k0) / eps, (grd * dy).sum()))
        return (k1 - k0) / eps, (grd * dy).sum()

    def solve(self, x, xi, b=None, iterMax=100, sequence='reference'):
        """Solve K(x, x) a = xi by conjugate gradients (base/kernelFunctions.py:676-680;
        `b` is unused there too), following linearcg (base/conjugateGradient.py:351-408):
        zero start; return at once when |xi|^2 < 1e-10; stop when
        alpha * mean|p| < 1e-10, when the quadratic energy a.Ka - 2 xi.a stalls
        (relative change < 1e-20), or after iterMax products.

        sequence 'reference' (default) gives linearcg's iterates as they run:
        its first search direction is the residual array itself, which the
        in-place residual update also moves, so the second direction is
        (1 + beta) times the residual; 'textbook' keeps the two apart
        (conjugate gradients proper; this build's warm start uses it).  The
        reference's per-iteration residual (an extra product used only for
        logging) is not computed.  Without affine part the whole solve runs on
        the device (lddmm_kernel_solve: CG scalars and stop flag in device
        memory, no host round trip per product); with one, the products are
        Kernel.applyK calls and the scalars come back once per product."""
        if sequence not in ('reference', 'textbook'):
            raise ValueError(f"unknown CG sequence {sequence!r}")
        tt = isinstance(x, torch.Tensor) or isinstance(xi, torch.Tensor)
        from . import _backend as B
        x_ = B.to_dev(x)
        rhs = B.to_dev(xi)
        if self.affine == 'none':
            a = B.kernel_solve(self, x_, rhs, iterMax, sequence)
        else:
            a = self._solve_host_loop(x_, rhs, iterMax, sequence)
        return a if tt else a.cpu().numpy()

    def _solve_host_loop(self, x_, rhs, iterMax, sequence):
        """linearcg with Kernel.applyK products (affine part included)."""
        a = torch.zeros_like(rhs)
        z = -2 * rhs                      # K a - 2 xi, tracked with the iterate
        r = rhs.clone()
        ener_old = 0.0
        p = None
        mu = None
        for it in range(iterMax):
            mu_new = (r * r).sum()
            if it == 0:
                if float(mu_new) < 1e-10:
                    break
                p = r.clone()
            else:
                p = r + (mu_new / mu) * p
            mu = mu_new
            q = self.applyK(x_, p)
            alpha = mu / (p * q).sum()
            a = a + alpha * p
            r = r - alpha * q
            z = z + alpha * q
            if it == 0 and sequence == 'reference':
                p = r.clone()             # the reference's direction is its residual array
            chk = torch.stack([alpha * p.abs().mean(), (z * a).sum()]).cpu()
            ener = float(chk[1])
            if float(chk[0]) < 1e-10:
                break
            if abs(ener - ener_old) < 1e-20 * abs(ener_old):
                break
            ener_old = ener
        return a

    # (base/kernelFunctions.py:665-674)
    def applyDDiffK11and12(self, x, n, a, p):
        z = ku.applykdiff11and12(x, n, a, p, self.name, self.sigma, self.order, KP=self.ms_exponent,
                                 dtype=self.pk_dtype)
        if self.affine == 'affine':
            z = z + self.w1 * a @ (a.T @ p)
        elif self.affine == 'euclidean':
            for E in self.affine_basis:
                ETa = a @ _like(E, x)
                z = z + self.w1 * (p * ETa).sum() * ETa
        return z
