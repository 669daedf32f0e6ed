"""Second-order (geodesic) surface matching: the surface drop-in of
py-lddmm base/secondOrderSurfaceMatching.py:41-893.

`SecondOrderSurfaceMatching(Template, Target, options).optimizeMatching()`
registers a template surface to a target by geodesic shooting: the control
is the initial momentum a0 on the template's vertices (option
'initialMomentum' gives a starting value, base/secondOrderSurfaceMatching.py:
58-61 via base/secondOrderPointSetMatching.py:82-88), the trajectory is the
EPDiff geodesic of (x0, a0) over T = 1/timeStep steps, and the objective is
1/2 a0.K(x0)a0 + |S(x_T) - target|^2 / sigmaError^2 with the varifold,
current or measure data term of SurfaceMatching.  The reference assembles
this from SurfaceMatching (surfaces, data term, BFGS driver protocol,
optimizeMatching) and SecondOrderPointSetMatching (solveStateEquation,
objectiveFunDef :188-197, updateTry :281, getGradient :557, dot products);
here GeodesicControl carries the second-order part in front of
SurfaceMatching.

On the device: the geodesic forward and adjoint (lddmm_epdiff_forward /
lddmm_epdiff_adjoint, one call each for the T steps), the fused data-term
energy and gradient (lddmm_surface_varifold or the current / measure kernel
sums), the L-BFGS direction (compact form).  As in the reference, internal
costs are not applied to second-order methods (a given internalCost is
dropped with a warning, base/secondOrderSurfaceMatching.py:44-47), and
affine controls are not built; a kernel with an affine part (affineKernel)
shoots with that part, as the reference's Kernel methods add it.
"""
import logging

from .secondOrderPointSetMatching import GeodesicControl
from .surfaceMatching import SurfaceMatching


class SecondOrderSurfaceMatching(GeodesicControl, SurfaceMatching):

    def set_internal_cost(self):
        """base/secondOrderSurfaceMatching.py:44-47."""
        if self.options['internalCost'] is not None:
            self.options['internalCost'] = None
            logging.info('Warning: Hybrid models not implemented for second-order methods: using basic LDDMM')
        self.icost = None
