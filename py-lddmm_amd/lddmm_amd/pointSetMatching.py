"""Point-set matching (plain LDDMM with a time-dependent momentum control),
restating py-lddmm base/pointSetMatching.py:58-559 for the configuration the
hot path serves: `PointSetMatching(Template, Target, options).optimizeMatching()`
with errorType 'measure' (weighted point-set measure norm) or 'L2' (labelled
points), algorithm 'bfgs', affine None.  Shooting, adjoint, BFGS memory and
the control live on the device exactly as in SurfaceMatching; only the data
term differs.
"""
import numpy as np
import torch

from . import _backend as B
from . import pointset_distances as psd
from .pointSets import PointSet
from .surfaceMatching import SurfaceMatching


class PointSetMatching(SurfaceMatching):
    unreduced_supported = False

    def getDefaultOptions(self):
        return super().getDefaultOptions()          # errorType 'measure' (base/basicMatching.py:95)

    def setInitialOptions(self, options):
        et = (options or {}).get('errorType', 'measure')
        if et not in ('measure', 'L2'):
            raise NotImplementedError(f"errorType {et!r}: PointSetMatching builds 'measure' and 'L2'")
        opts = dict(options or {})
        opts['errorType'] = 'varifold'              # pass the surface driver's option check
        super().setInitialOptions(opts)
        self.options['errorType'] = et

    def set_template_and_target(self, Template, Target):
        """base/pointSetMatching.py:120-141 (no file output)."""
        if Template is None or Target is None:
            raise ValueError("PointSetMatching needs a template and a target point set")
        self.fv0 = PointSet(data=Template) if isinstance(Template, PointSet) else PointSet(np.asarray(Template))
        self.fv1 = PointSet(data=Target) if isinstance(Target, PointSet) else PointSet(np.asarray(Target))
        self.fvInit = PointSet(data=self.fv0)
        self.dim = self.fv0.vertices.shape[1]
        if self.options['errorType'] == 'L2' and self.fv1.vertices.shape != self.fv0.vertices.shape:
            raise ValueError("errorType 'L2' needs template and target of the same shape")

    def set_parameters(self):
        """base/pointSetMatching.py:156-161 (no epsInit scaling)."""
        eps = self.options['epsInit']
        super().set_parameters()
        self.options['epsInit'] = eps

    def set_internal_cost(self):
        self.icost = None
        if self.options['internalCost'] is not None:
            raise NotImplementedError("internal costs are built for surfaces only (not point sets)")

    def initialize_variables(self):
        super().initialize_variables()
        self.fvDef = PointSet(data=self.fvInit)

    def fun_obj0(self, fv1):
        if self.options['errorType'] == 'measure':
            return psd.measureNorm0(fv1, self.options['KparDist'])
        return psd.L2Norm0(fv1)

    def dataTerm_device(self, xT):
        if self.options['errorType'] == 'measure':
            e = psd.measureNormDef_device(self.fv0, self.fv1, self.options['KparDist'], vertices=xT)
        else:
            x1 = B.to_dev(self.fv1.vertices)
            e = -2 * (xT * x1).sum() + (xT * xT).sum()
        return e / self.options['sigmaError'] ** 2

    def endPointGradient_device(self, xT):
        if self.options['errorType'] == 'measure':
            g = psd.measureNormGradient_device(self.fv0, self.fv1, self.options['KparDist'], vertices=xT)
        else:
            g = 2 * (xT - B.to_dev(self.fv1.vertices))
        return g / self.options['sigmaError'] ** 2

    def optimizeMatching(self):
        """base/pointSetMatching.py:540-559 (bfgs branch): gradEps = max(0.001,
        |g| / 10000), epsInit 1, memory 50."""
        from . import bfgs as bfgs_mod
        grd = self.getGradient(self.gradCoeff)
        [grd2] = self.dotProduct_euclidean(grd, [grd])
        if self.gradEps < 0:
            self.gradEps = max(0.001, np.sqrt(grd2) / 10000)
        bfgs_mod.bfgs(self, verb=self.options['verb'], maxIter=self.options['maxIter'],
                      TestGradient=self.options['testGradient'], epsInit=1.,
                      Wolfe=self.options['Wolfe'], memory=50)
        self.fvDef.updateVertices(self.state['xt'][-1].cpu().numpy())
        return self.control
