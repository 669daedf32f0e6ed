"""Weak-Wolfe line search along a descent direction.

The step rules are the reference's (base/linesearch.py:66-163), which the
iterates of the matching drivers depend on: a trial step t is accepted when
it gives sufficient decrease (Armijo, c1) and a slope no steeper than c2
times the initial one; below the sufficient-decrease bound it becomes the
lower end of a bracket, above it the upper end.  With no upper end yet the
step doubles; with one, the next trial is the bracket midpoint after a
sufficient-decrease trial, else the minimiser of the quadratic through
f(0), f'(0), f(t), clipped to [lo + (hi - lo)/10, (lo + hi)/2].  After
`maxiter` slope evaluations it falls back to halving until sufficient
decrease; 100 trials without success return no step.

The optimisable object provides updateTry(dir, t) (objective at the control
minus t*dir) and getGradient(coeff, update=[dir, t]) (gradient there);
directions are subtracted, so the slope along the step is -<grad, dir>.
Decrease tests use gradCoeff-scaled slopes, as the reference does.
"""
import logging

MAX_TRIALS = 100


class _Probe:
    """Objective and slope of t -> f(x - t d), counting evaluations.

    When the object provides updateTryWithSlope, a trial's gradient and slope
    are computed together with its objective (one read-back instead of two);
    the slope is then counted, and its gradient kept, only if the search asks
    for it, so the counts and every decision are those of separate calls."""

    def __init__(self, opt, direction, euclidean, speculate=None):
        self.opt = opt
        self.direction = direction
        self.coeff = getattr(opt, "gradCoeff", 1)
        self._dots = opt.dotProduct_euclidean if euclidean else opt.dotProduct
        self._with_slope = getattr(opt, "updateTryWithSlope", None) if euclidean else None
        self._speculate = speculate if self._with_slope is not None else None
        self._pending = None       # (t, slope, gradient) of the last trial
        self.n_values = 0
        self.n_slopes = 0
        self.last_grad = None

    def slope_of(self, grad):
        return -self._dots(grad, [self.direction])[0]

    def value(self, t):
        self.n_values += 1
        self._pending = None
        if self._with_slope is not None and t != 0:
            if self._speculate is not None:
                # the caller's work for "t accepted" is queued before the read-back
                f, slope, grad = self._with_slope(self.direction, t, self.coeff,
                                                  before_readback=lambda g: self._speculate(t, g))
            else:
                f, slope, grad = self._with_slope(self.direction, t, self.coeff)
            self._pending = (t, slope, grad)
            return f
        return self.opt.updateTry(self.direction, t)

    def slope(self, t):
        self.n_slopes += 1
        if self._pending is not None and self._pending[0] == t:
            _, slope, self.last_grad = self._pending
            self._pending = None
            return slope
        self.last_grad = self.opt.getGradient(self.coeff, update=[self.direction, t])
        return self.slope_of(self.last_grad)


def _next_trial(t, f_t, f0, slope0, lo, hi, decreased):
    if hi < 1e-16:                      # no upper end yet
        return 2 * t
    if decreased:
        return 0.5 * hi + 0.5 * lo
    t_quad = -(t ** 2 * slope0) / (2 * (f_t - f0 - t * slope0))
    return max(0.9 * lo + 0.1 * hi, min(0.5 * hi + 0.5 * lo, t_quad))


def line_search_weak_wolfe(opt, pk, gfk=None, old_fval=None, old_old_fval=None, c1=0.001, c2=0.9,
                           amax=None, maxiter=100, euclidean=True, t_init=1., slope0=None, speculate=None):
    """Returns (t or None, objective evaluations, slope evaluations, f(t),
    f(0), gradient at t or None) for the direction pk (amax and old_old_fval
    are part of the reference's signature and unused).  slope0, when given,
    is the slope at t = 0, -<gfk, pk>, computed by the caller, or a function
    returning it; the function is called after the first trial's objective,
    where the slope is first needed.  speculate(t, g), when given and the
    object has updateTryWithSlope, is called for every trial t with its
    gradient g after the trial's work is queued and before its values are
    read back (the caller may queue what it would do if t were accepted);
    it changes no decision of the search."""
    probe = _Probe(opt, pk, euclidean, speculate)
    if slope0 is None:
        if gfk is None:
            gfk = opt.getGradient(probe.coeff)
        slope0 = probe.slope_of(gfk)
    f0 = probe.value(0) if old_fval is None else old_fval

    lo = hi = 0
    trials = 0
    t = t_init
    f_t = probe.value(t)
    if callable(slope0):
        slope0 = slope0()
    armijo = lambda t, f_t: f_t < f0 + c1 * t * slope0 * probe.coeff
    while trials < MAX_TRIALS and probe.n_slopes < maxiter:
        decreased = armijo(t, f_t)
        if decreased:
            if probe.slope(t) > c2 * slope0 * probe.coeff:
                break                   # both conditions hold
            lo = t
        else:
            hi = t
        t = _next_trial(t, f_t, f0, slope0, lo, hi, decreased)
        f_t = probe.value(t)
        trials += 1

    if probe.n_slopes == maxiter:
        logging.warning('Weak Wolfe condition: maximum gradient computations attained: switching to backtracking')
        probe.last_grad = None
        while trials < MAX_TRIALS:
            if armijo(t, f_t):
                probe.slope(t)
                break
            t /= 2
            f_t = probe.value(t)
            trials += 1

    if trials == MAX_TRIALS:
        logging.warning(f'Weak Wolfe condition: maximum number of iterations attained {f_t:.5f}--{f0:.5f}')
        t = None
    return t, probe.n_values, probe.n_slopes, f_t, f0, probe.last_grad
