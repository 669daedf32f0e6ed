"""Weak-Wolfe line search (py-lddmm base/linesearch.py:66-163).

The optimisable object provides updateTry(dir, t) (objective at control -
t*dir, stored as the trial variable) and getGradient(coeff, update=[dir, t])
(gradient at the trial control); directions are "descent directions to be
subtracted", so phi'(t) = -<grad(t), dir>.  Steps, tests and the quadratic
interpolation follow the reference line by line.
"""
import logging


def line_search_weak_wolfe(opt, pk, gfk=None, old_fval=None, old_old_fval=None, c1=0.001, c2=0.9,
                           amax=None, maxiter=100, euclidean=True, t_init=1.):
    fc = [0]
    gc = [0]
    gval = [None]

    def dot(g, d):
        if euclidean:
            return opt.dotProduct_euclidean(g, [d])[0]
        return opt.dotProduct(g, [d])[0]

    def phi(alpha):
        fc[0] += 1
        return opt.updateTry(pk, alpha)

    def derphi(alpha):
        gc[0] += 1
        gval[0] = opt.getGradient(opt.gradCoeff, update=[pk, alpha])
        return -dot(gval[0], pk)

    if gfk is None:
        gfk = opt.getGradient(opt.gradCoeff)
    derphi0 = -dot(gfk, pk)

    it = 0
    itGrad = 0
    amax = 0
    amin = 0
    if old_fval is None:
        old_fval = phi(0)

    t = t_init
    gradCoeff = getattr(opt, "gradCoeff", 1)
    maxiter2 = 100
    fval = phi(t)
    df = None
    while it < maxiter2 and itGrad < maxiter:
        armijo = False
        if fval < old_fval + c1 * t * derphi0 * gradCoeff:
            df = derphi(t)
            armijo = True
            itGrad += 1
            if df > c2 * derphi0 * gradCoeff:
                break
            else:
                amin = t
        else:
            amax = t
        if amax < 1e-16:
            t *= 2
        else:
            if armijo:
                t = 0.5 * amax + 0.5 * amin
            else:
                tt = -(t ** 2 * derphi0) / (2 * (fval - old_fval - t * derphi0))
                t = max(0.9 * amin + 0.1 * amax, min(0.5 * amax + 0.5 * amin, tt))
        fval = phi(t)
        it += 1

    if itGrad == maxiter:
        logging.warning('Weak Wolfe condition: maximum gradient computations attained: switching to backtracking')
        gval[0] = None
        while it < maxiter2:
            if fval < old_fval + c1 * t * derphi0 * gradCoeff:
                df = derphi(t)
                break
            else:
                t /= 2
            fval = phi(t)
            it += 1

    if it == maxiter2:
        logging.warning(f'Weak Wolfe condition: maximum number of iterations attained {fval:.5f}--{old_fval:.5f}')
        t = None

    return t, fc[0], gc[0], fval, old_fval, gval[0]
