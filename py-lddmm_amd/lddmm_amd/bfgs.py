"""L-BFGS driver (py-lddmm base/bfgs.py:55-312).

Same protocol as the reference: the optimisable object provides
getVariable, objectiveFun, updateTry, acceptVarTry, getGradient and
optionally startOfIteration / endOfIteration / endOfProcedure,
dotProduct_euclidean, addProd / prod / copyDir, gradEps, gradCoeff, epsMax,
burnIn, reset, twoLoop (the two-loop recursion computed by the object, e.g.
with its scalars kept on the device).  Two-loop recursion with memory `memory`, t_init = 1/|q| on
the first (and reset) iterations, weak-Wolfe search (c1 = 1e-4, c2 = 0.9,
at most 10 gradients), stop on |obj - obj_old| < 1e-7 after burn-in or on
|g| < gradEps.  Directions may hold numpy arrays or device tensors; dot
products are reduced to Python floats.
"""
import logging
import time
from copy import deepcopy

import numpy as np

from .linesearch import line_search_weak_wolfe


def _dot_default(x, y):
    return [float((x * yy).sum()) for yy in y]


def _addProd(x, y, a):
    return x + a * y


def _prod(x, a):
    return a * x


def _copyDir(x):
    return x.clone() if hasattr(x, "clone") else np.copy(x)


def bfgs(opt, verb=True, maxIter=1000, TestGradient=False, epsInit=0.01, memory=25, Wolfe=True,
         lineSearch='Weak_Wolfe'):
    for name in ('getVariable', 'objectiveFun', 'updateTry', 'acceptVarTry', 'getGradient'):
        if not hasattr(opt, name):
            logging.error('Error: required functions are not provided')
            return
    if lineSearch != "Weak_Wolfe":
        logging.warning(f'line search {lineSearch!r} is not built: using weak wolfe condition')
    line_search = line_search_weak_wolfe

    if hasattr(opt, 'dotProduct_euclidean'):
        dotProduct = opt.dotProduct_euclidean
    elif hasattr(opt, 'dotProduct'):
        dotProduct = opt.dotProduct
    else:
        dotProduct = _dot_default
    addProd = getattr(opt, 'addProd', _addProd)
    prod = getattr(opt, 'prod', _prod)
    copyDir = getattr(opt, 'copyDir', _copyDir)
    stopCondition = getattr(opt, 'stopCondition', lambda: True)

    if hasattr(opt, 'startOptim'):
        opt.startOptim()

    if hasattr(opt, 'gradEps'):
        gradEps = opt.gradEps
    elif hasattr(opt, 'options') and 'gradEps' in opt.options.keys():
        gradEps = opt.options['gradEps']
    else:
        gradEps = None
    gradCoeff = getattr(opt, 'gradCoeff', 1.0)
    epsMax = getattr(opt, 'epsMax', 1.)
    burnIn = getattr(opt, 'burnIn', 0)
    eps = epsInit
    opt.converged = False

    if hasattr(opt, 'reset') and opt.reset:
        opt.obj = None

    obj = opt.objectiveFun()
    opt.reset = False
    logging.info('iteration 0: obj = {0: .5f}'.format(obj))

    storedGrad = []
    it = 0
    diffVar = None
    grdOld = None
    obj_old = None
    gval = None
    opt.reset = False
    itt0 = time.process_time()
    history = [obj]
    while it < maxIter:
        t0 = time.process_time()
        if hasattr(opt, 'startOfIteration'):
            opt.startOfIteration()
        if opt.reset:
            opt.obj = None
            obj = opt.objectiveFun()
            if verb:
                logging.info(f"recomputed objective {obj:.5f}")
            obj_old = None
            gval = None

        if gval is None:
            grd = opt.getGradient(gradCoeff)
        else:
            grd = deepcopy(gval)

        if TestGradient:
            dirfoo = opt.randomDir()
            epsfoo = 1e-8
            objfoo1 = opt.updateTry(dirfoo, epsfoo, obj - 1e8)
            objfoo2 = opt.updateTry(dirfoo, -epsfoo, obj - 1e8)
            [grdfoo] = dotProduct(grd, [dirfoo])
            logging.info('Test Gradient: %.4f %.4f' % ((objfoo1 - objfoo2) / (2 * epsfoo), -grdfoo * gradCoeff))

        if (not opt.reset) and it > 0:
            storedGrad.append([diffVar, addProd(grd, grdOld, -1)])
            if len(storedGrad) > memory:
                storedGrad.pop(0)
            if hasattr(opt, 'twoLoop'):
                # same recursion with the scalars kept where the vectors live
                q = opt.twoLoop(grd, storedGrad)
            else:
                q = copyDir(grd)
                rho = []
                alpha = []
                for m in reversed(storedGrad):
                    rho.append(1 / dotProduct(m[1], [m[0]])[0])
                    alpha.append(rho[-1] * dotProduct(m[0], [q])[0])
                    q = addProd(q, m[1], -alpha[-1])
                rho.reverse()
                alpha.reverse()
                m = storedGrad[-1]
                c = dotProduct(m[0], [m[1]])[0] / dotProduct(m[1], [m[1]])[0]
                if c < 1e-10:
                    c = 1
                q = prod(q, c)
                for k, m in enumerate(storedGrad):
                    beta = rho[k] * dotProduct(m[1], [q])[0]
                    q = addProd(q, m[0], alpha[k] - beta)
        else:
            storedGrad = []
            q = copyDir(grd)

        grd2 = dotProduct(grd, [grd])[0]
        grdTry = np.sqrt(np.maximum(1e-20, dotProduct(q, [q])[0]))
        dir0 = copyDir(q)
        if it == 0 or opt.reset:
            t_init = 1. / grdTry
        else:
            t_init = 1.

        grdOld = copyDir(grd)

        if it == 0 or it == burnIn:
            if gradEps is None:
                gradEps = max(0.001 * np.sqrt(grd2), 0.0001)
            else:
                gradEps = min(gradEps, 0.0001 * np.sqrt(grd2))
            logging.info(f'Gradient threshold: {gradEps:.6f}')

        if it < burnIn:
            Wolfe = False
            eps = epsInit
        else:
            if Wolfe:
                eps = 1.
            else:
                epsBig = epsMax / grdTry
                if eps > epsBig:
                    eps = epsBig

        eps, fc, gc, phi_star, old_fval, gval = line_search(opt, dir0, gfk=grd, old_fval=obj,
                                                            old_old_fval=obj_old, c1=1e-4, c2=0.9,
                                                            amax=None, t_init=t_init, maxiter=10)
        if eps is not None:
            diffVar = prod(dir0, -eps)
            obj_old = obj
            opt.acceptVarTry()
            obj = phi_star
            newreset = False
        else:
            logging.info('Wolfe search unsuccessful')
            if opt.reset:
                logging.info('Cannot go any further')
                opt.converged = False
                if hasattr(opt, 'endOfProcedure'):
                    opt.endOfProcedure()
                elif hasattr(opt, 'endOfIteration'):
                    opt.endOfIteration()
                break
            newreset = True
            gval = None

        if (obj_old is not None and np.fabs(obj - obj_old) < 1e-7) and stopCondition():
            logging.info(f'iteration {it + 1:d}: obj = {obj:.5f}, eps = {eps:.5f}, gradient: {np.sqrt(grd2):.5f}')
            if it > burnIn or opt.reset:
                logging.info('Stopping Gradient Descent: small variation')
                opt.converged = True
                history.append(obj)
                if hasattr(opt, 'endOfProcedure'):
                    opt.endOfProcedure()
                elif hasattr(opt, 'endOfIteration'):
                    opt.endOfIteration()
                break
            else:
                newreset = True

        tt0 = time.process_time()
        t1 = tt0 - t0
        tt1 = tt0 - itt0
        itt0 = tt0
        if verb | (it == maxIter):
            if eps is None:
                logging.info(f'iteration {it + 1:d}: obj = {obj:.5f}, eps = None, gradient: {np.sqrt(grd2):.5f}, '
                             f'time = {t1:.04f}, {tt1:.04f}')
            else:
                logging.info(f'iteration {it + 1:d}: obj = {obj:.5f}, eps = {eps:.5f}, gradient: '
                             f'{np.sqrt(grd2):.5f}, time = {t1:.04f}, {tt1:.04f}')

        if np.sqrt(grd2) < gradEps and stopCondition():
            logging.info('Stopping Gradient Descent: small gradient')
            opt.converged = True
            history.append(obj)
            if hasattr(opt, 'endOfProcedure'):
                opt.endOfProcedure()
            elif hasattr(opt, 'endOfIteration'):
                opt.endOfIteration()
            break
        if eps is not None:
            eps = np.minimum(100 * eps, epsMax)
        else:
            eps = epsMax

        opt.reset = newreset
        if hasattr(opt, 'endOfIteration'):
            opt.endOfIteration()
        if not opt.reset:
            it += 1
            history.append(obj)
        # optional early stop requested by the optimisable object (no
        # counterpart in the reference loop; absent -> identical behaviour)
        if hasattr(opt, 'stopNow') and opt.stopNow():
            opt.converged = True
            break

    if it == maxIter and hasattr(opt, 'endOfProcedure'):
        opt.endOfProcedure()
    if hasattr(opt, 'endOptim'):
        opt.endOptim()
    opt.bfgs_iterations = it
    opt.bfgs_history = history
    return opt.getVariable()
