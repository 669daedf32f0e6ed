"""Second-order (geodesic) point-set matching: the control is the initial
momentum a0, the trajectory is the geodesic (EPDiff) shot from (x0, a0), and
the objective is 1/2 a0.K(x0)a0 plus the data term at the endpoint.
Restates py-lddmm base/secondOrderPointSetMatching.py:48-460 for the
configuration the hot path serves: `SecondOrderPointSetMatching(Template,
Target, options).optimizeMatching()` with errorType 'measure' or 'L2',
algorithm 'bfgs', affine None (no affine controls), a single target.

The geodesic forward and its adjoint run on the device (lddmm_epdiff_forward
/ lddmm_epdiff_adjoint, T steps each in one call); the control, its BFGS
memory and the trajectory stay on the device as in SurfaceMatching.  The
geodesic control itself is `GeodesicControl`, shared with
SecondOrderSurfaceMatching (the same control over the surface data terms).
"""
import torch

from . import _backend as B
from . import pointEvolution as evol
from .pointSetMatching import PointSetMatching
from .surfaceMatching import Control, State


class GeodesicControl:
    """The second-order control of base/secondOrderPointSetMatching.py placed
    in front of a first-order driver (PointSetMatching or SurfaceMatching)
    in the method order: initial momentum a0 instead of the time-dependent
    momenta, geodesic trajectory, objective 1/2 a0.K(x0)a0 + data term,
    gradient K(x0)a0 - pa_0.  The driver supplies the data term
    (dataTerm_device / endPointGradient_device), the BFGS hooks and the
    device L-BFGS direction."""
    unreduced_supported = False

    def getDefaultOptions(self):
        """base/secondOrderPointSetMatching.py:53-57."""
        options = super().getDefaultOptions()
        options['controlWeight'] = 1.0
        options['initialMomentum'] = None
        return options

    def setInitialOptions(self, options):
        opts = dict(options or {})
        a0 = opts.pop('initialMomentum', None)
        cw = opts.pop('controlWeight', 1.0)
        super().setInitialOptions(opts)
        self.options['initialMomentum'] = a0
        self.options['controlWeight'] = cw

    def initialize_variables(self):
        """base/secondOrderPointSetMatching.py:76-97: a0 = 0 (trajectory at
        rest) or the given initial momentum (trajectory shot from it)."""
        super().initialize_variables()
        self.control = Control()
        self.controlTry = Control()
        a0 = self.options['initialMomentum']
        if a0 is None:
            self.control['a0'] = torch.zeros_like(self.x0_dev)
            self.state = State()
            self.state['xt'] = self.x0_dev.expand(self.Tsize + 1, -1, -1).clone()
            self.state['at'] = torch.zeros_like(self.state['xt'])
        else:
            a0 = B.to_dev(a0)
            if tuple(a0.shape) != tuple(self.x0_dev.shape):
                raise ValueError(f"initialMomentum shape {tuple(a0.shape)} != {tuple(self.x0_dev.shape)}")
            self.control['a0'] = a0.clone()
            _, self.state = self.objectiveFunDef(self.control, withTrajectory=True)
        self.controlTry['a0'] = torch.zeros_like(self.control['a0'])

    def solveStateEquation(self, control=None, init_state=None, kernel=None, options=None):
        """base/secondOrderPointSetMatching.py:100-111 (no affine): the geodesic
        trajectory {'xt', 'at'} of (x0, a0), on the device."""
        control = self.control if control is None else control
        x0 = self.x0_dev if init_state is None else B.to_dev(init_state)
        kernel = self.options['KparDiff'] if kernel is None else kernel
        xt, at = evol._epdiff(x0, control['a0'], kernel, self.Tsize)
        st = State()
        st['xt'] = xt
        st['at'] = at
        return st

    def objectiveFunDef(self, control, var=None, withTrajectory=True, withJacobian=False, display=False):
        """1/2 a0.K(x0)a0 (device float64 scalar) and the trajectory
        (base/secondOrderPointSetMatching.py:113-147)."""
        if var is not None and 'Init' in var:
            raise NotImplementedError("a moving initial shape is not built")
        st = self.solveStateEquation(control=control)
        if withJacobian:
            st['Jt'] = evol.jacobian(st['xt'], st['at'], self.options['KparDiff'], self.Tsize)
        a0 = B.to_dev(control['a0'])
        obj = 0.5 * (a0 * self.options['KparDiff'].applyK(self.x0_dev, a0)).sum()
        return (obj, st) if (withTrajectory or withJacobian) else obj

    def getGradient(self, coeff=1.0, update=None, endpoint_gradient=None):
        """grad['a0'] = (K(x0) a0 - pa_0) / coeff with px1 = -dataTerm'(x_T),
        pa1 = 0 (base/secondOrderPointSetMatching.py:379-460, no affine);
        endpoint_gradient: dataTerm'(x_T) when already computed."""
        if update is None:
            control, st = self.control, self.state
        else:
            lt = self._last_try
            if lt is not None and lt[0] is update[0] and lt[1] == update[1]:
                control, st = lt[2], lt[3]
            else:
                control = Control()
                for k in update[0].keys():
                    if update[0][k] is not None:
                        control[k] = self.control[k] - update[1] * update[0][k]
                st = self.solveStateEquation(control=control)
        px1 = -(self.endPointGradient_device(st['xt'][-1]) if endpoint_gradient is None else endpoint_gradient)
        pa1 = torch.zeros_like(px1)
        _, _, da0 = evol._epdiff_adjoint(st['xt'], st['at'], px1, pa1, self.options['KparDiff'], self.Tsize)
        grd = Control()
        grd['a0'] = da0 / coeff
        return grd

    def randomDir(self):
        d = Control()
        d['a0'] = torch.randn_like(self.control['a0'])
        return d

    def setInitialControl(self, a0):
        """Start the optimisation from the initial momentum a0 [N, d]."""
        a0_ = B.to_dev(a0)
        if tuple(a0_.shape) != tuple(self.control['a0'].shape):
            raise ValueError(f"control shape {tuple(a0_.shape)} != {tuple(self.control['a0'].shape)}")
        self.control['a0'] = a0_.clone()
        self.obj = None
        self.state = self.solveStateEquation()

    def control_numpy(self):
        return self.control['a0'].cpu().numpy()

    def saveOutputs(self):
        """Per-step files of the geodesic (base/secondOrderSurfaceMatching.py:
        819-868): the deformed template with the log-Jacobian fields and the
        velocities v_t = K(x_t) a_t of the geodesic's momenta."""
        if not hasattr(self, 'saveEvolution') or not hasattr(self.fvInit, 'faces'):
            raise NotImplementedError("file output (writeOutput) is built for surface matching only")
        _, st = self.objectiveFunDef(self.control, withTrajectory=True, withJacobian=True)
        K = self.options['KparDiff']
        vel = torch.stack([K.applyK(st['xt'][t], st['at'][t]) for t in range(self.Tsize)])
        self.saveEvolution(self.fvInit, st, fileName=self.options['saveFile'], velocity=vel.cpu().numpy())

    def endOfIteration(self, forceSave=False):
        """base/secondOrderPointSetMatching.py:538-560: count the iteration and
        write the per-step files every saveRate iterations (with the option
        writeOutput); the precision switched to float64 after a restart is
        kept (no return to pk_dtype, unlike the first-order drivers)."""
        self.iter += 1
        if self.options['writeOutput'] and (forceSave or self.iter % self.saveRate == 0):
            self.saveOutputs()


class SecondOrderPointSetMatching(GeodesicControl, PointSetMatching):
    pass
