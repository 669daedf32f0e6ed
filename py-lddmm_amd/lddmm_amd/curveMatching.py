"""Curve matching (plain LDDMM with a time-dependent momentum control),
restating py-lddmm base/curveMatching.py:191-500 and :755-778 for the
configuration the hot path serves: `CurveMatching(Template, Target,
options).optimizeMatching()` with errorType 'varifold' (f(u) = 1 + u^2,
weight 1) or 'current' (weight None), algorithm 'bfgs', affine None, no
internal cost.  Curves may be 2-D or 3-D.

The control, trajectories and BFGS memory live on the GPU exactly as in
SurfaceMatching (device shooting with the fused energy, device adjoint with
the Euclidean K correction); only the data term and its gradient differ.
"""
import numpy as np

from . import _backend as B
from . import curveDistances as cd
from .curves import Curve
from .surfaceMatching import SurfaceMatching


class CurveMatching(SurfaceMatching):
    unreduced_supported = False

    def getDefaultOptions(self):
        o = super().getDefaultOptions()
        o['errorType'] = 'varifold'
        return o

    def set_template_and_target(self, Template, Target):
        """base/curveMatching.py:138-155 (no file output, no orientation fix)."""
        if Template is None or Target is None:
            raise ValueError("CurveMatching needs a template and a target curve")
        self.fv0 = Curve(curve=Template)
        self.fv1 = Curve(curve=Target)
        self.fvInit = Curve(curve=self.fv0)
        self.dim = self.fv0.vertices.shape[1]
        if self.dim not in (2, 3) or self.fv1.vertices.shape[1] != self.dim:
            raise ValueError("CurveMatching: template and target must be 2-D or 3-D curves of one dimension")

    def set_parameters(self):
        """base/curveMatching.py:112-136 (gradEps = -1; no epsInit scaling)."""
        eps = self.options['epsInit']
        super().set_parameters()
        self.options['epsInit'] = eps
        self.gradEps = -1

    def set_internal_cost(self):
        self.icost = None
        if self.options['internalCost'] is not None:
            raise NotImplementedError("internal costs are built for surfaces only (not curves)")

    def initialize_variables(self):
        super().initialize_variables()
        self.fvDef = Curve(curve=self.fvInit)

    def fun_obj0(self, fv1):
        K = self.options['KparDist']
        if self.options['errorType'] == 'varifold':
            return cd.varifoldNorm0(fv1, K, weight=1.)
        return cd.currentNorm0(fv1, K)

    def dataTerm_device(self, xT):
        K = self.options['KparDist']
        if self.options['errorType'] == 'varifold':
            e, _ = cd.varifold_device(self._dist_kd(), self.fv0, self.fv1, energy=True, grad=False, v1=xT)
            e = e[0]
        else:
            e = cd.currentNormDef_device(self.fv0, self.fv1, K, vertices=xT)
        return e / self.options['sigmaError'] ** 2

    def endPointGradient_device(self, xT):
        K = self.options['KparDist']
        if self.options['errorType'] == 'varifold':
            _, g = cd.varifold_device(self._dist_kd(), self.fv0, self.fv1, energy=False, grad=True, v1=xT)
        else:
            g = cd.currentNormGradient_device(self.fv0, self.fv1, K, vertices=xT)
        return g / self.options['sigmaError'] ** 2

    def optimizeMatching(self):
        """base/curveMatching.py:755-778 (bfgs branch): gradEps = max(gradLB,
        |g| / 10000), epsInit = 1, memory 50."""
        from . import bfgs as bfgs_mod
        grd = self.getGradient(self.gradCoeff)
        [grd2] = self.dotProduct_euclidean(grd, [grd])
        self.gradEps = max(self.options['gradLB'], np.sqrt(grd2) / 10000)
        bfgs_mod.bfgs(self, verb=self.options['verb'], maxIter=self.options['maxIter'],
                      TestGradient=self.options['testGradient'], epsInit=1.,
                      Wolfe=self.options['Wolfe'], lineSearch=self.options['lineSearch'], memory=50)
        self.fvDef.updateVertices(self.state['xt'][-1].cpu().numpy())
        return self.control
