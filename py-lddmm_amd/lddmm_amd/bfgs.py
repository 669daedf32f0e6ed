"""Limited-memory BFGS driver for the matching objects.

The protocol of the optimisable object is the reference's (described at
base/bfgs.py:9-35): required getVariable, objectiveFun, updateTry,
acceptVarTry, getGradient; optional startOptim, startOfIteration,
endOfIteration, endOfProcedure, endOptim, stopCondition, dotProduct_euclidean
(or dotProduct), addProd, prod, copyDir, randomDir and the attributes
gradEps (or options['gradEps']), gradCoeff, epsMax, burnIn, reset.  Two
hooks have no counterpart there: twoLoop(grad, pairs), which lets the object
compute the quasi-Newton direction itself (e.g. with its scalars kept on the
device), and stopNow(), an early stop checked after each counted iteration.

The loop is organised as an explicit state (`_Run`) plus one decision
function (`decide`) applied after every line search.  The decision rules are
the reference's (base/bfgs.py:243-300):

* a line search that fails right after a restart ends the run, not converged;
* any other failed line search restarts: the memory is dropped, the objective
  recomputed and the iteration repeated from the steepest-descent direction
  with the first step 1/|q|;
* |f - f_old| < 1e-7 ends the run as converged after the burn-in iterations
  (or right after a restart) and restarts before them;
* |g| < gradEps ends the run as converged;
* otherwise the step is accepted and the iteration counted.

The gradient threshold is fixed at iteration 0 and again at the end of the
burn-in (base/bfgs.py:218-223); the first step of the weak-Wolfe search
(c1 = 1e-4, c2 = 0.9, at most 10 gradients) is 1/|q| at iteration 0 and after
a restart, 1 otherwise (:210-213); the quasi-Newton direction is the
two-loop product with memory `memory` (:181-207).  Directions may hold numpy
arrays or device tensors; dot products come back as Python floats.
"""
import enum
import logging
import time

import numpy as np

from .linesearch import line_search_weak_wolfe

STALL_TOL = 1e-7          # |f - f_old| below which the run has stalled
WOLFE_C1, WOLFE_C2 = 1e-4, 0.9
WOLFE_MAX_GRADIENTS = 10


class Outcome(enum.Enum):
    ACCEPT = "accept"                  # step taken, iteration counted
    RESTART = "restart"                # memory dropped, iteration repeated
    STALLED = "small variation"        # converged: objective no longer moves
    SMALL_GRADIENT = "small gradient"  # converged: |g| below the threshold
    STUCK = "stuck"                    # no step even from steepest descent


def decide(step_found, after_restart, it, burn_in, f, f_old, grad_norm, grad_eps, may_stop):
    """End-of-iteration decision.  `may_stop` is the object's stopCondition;
    it is consulted only for a stopping rule that holds, the stall rule first."""
    if not step_found and after_restart:
        return Outcome.STUCK
    restart = not step_found
    if f_old is not None and abs(f - f_old) < STALL_TOL and may_stop():
        if it > burn_in or after_restart:
            return Outcome.STALLED
        restart = True
    if grad_norm < grad_eps and may_stop():
        return Outcome.SMALL_GRADIENT
    return Outcome.RESTART if restart else Outcome.ACCEPT


def gradient_threshold(current, grad_norm):
    """The |g| stopping threshold: 1e-3 |g| (at least 1e-4) when none is
    given, else the smaller of the given one and 1e-4 |g|."""
    if current is None:
        return max(0.001 * grad_norm, 0.0001)
    return min(current, 0.0001 * grad_norm)


class _VectorOps:
    """Direction arithmetic, from the object's hooks or numpy defaults."""

    def __init__(self, opt):
        self._table = None
        if hasattr(opt, 'dotProduct_euclidean'):
            self._dots = opt.dotProduct_euclidean
            self._table = getattr(opt, 'dotTable', None)
        elif hasattr(opt, 'dotProduct'):
            self._dots = opt.dotProduct
        else:
            self._dots = lambda u, vs: [float((u * v).sum()) for v in vs]
        self.axpy = getattr(opt, 'addProd', lambda u, v, c: u + c * v)        # u + c v
        self.scale = getattr(opt, 'prod', lambda u, c: c * u)
        self.copy = getattr(opt, 'copyDir', lambda u: u.clone() if hasattr(u, 'clone') else np.copy(u))

    def dot(self, u, v):
        return self._dots(u, [v])[0]

    def table(self, probes, targets):
        """A function returning table[i][j] = dot(probes[i], targets[j]).
        With the object's dotTable the products are launched now and read
        back, in one transfer, at the first call (the same values as its dot
        products)."""
        if self._table is not None:
            return _Deferred(self._table(probes, targets, later=True))
        tab = [[self.dot(p, t) for t in targets] for p in probes]
        return lambda: tab


class _Deferred:
    """A value read back at the first call and kept."""

    def __init__(self, read):
        self._read = read
        self._value = None

    def __call__(self):
        if self._read is not None:
            self._value = self._read()
            self._read = None
        return self._value


class _Memory:
    """The last `size` (s, y) pairs, oldest first, as [s, y] lists (twoLoop
    hooks may append their own cached data to an entry)."""

    def __init__(self, size):
        self.size = size
        self.pairs = []

    def push(self, s, y):
        """Appends (s, y); returns the entries dropped from the front."""
        self.pairs.append([s, y])
        n = max(0, len(self.pairs) - self.size)
        dropped = self.pairs[:n]
        del self.pairs[:n]
        return dropped

    def undo_push(self, dropped):
        """Reverts the last push, which dropped `dropped`."""
        self.pairs.pop()
        self.pairs[:0] = dropped

    def clear(self):
        self.pairs = []

    def __len__(self):
        return len(self.pairs)


def quasi_newton_direction(g, pairs, ops):
    """H g for the inverse-Hessian estimate of the stored pairs (two-loop
    product, Nocedal & Wright alg. 7.4), scaled by gamma = s.y / y.y of the
    newest pair (1 when that ratio is below 1e-10)."""
    q = ops.copy(g)
    down = []
    for s, y in reversed([p[:2] for p in pairs]):
        rho = 1 / ops.dot(y, s)
        a = rho * ops.dot(s, q)
        q = ops.axpy(q, y, -a)
        down.append((rho, a))
    s, y = pairs[-1][:2]
    gamma = ops.dot(s, y) / ops.dot(y, y)
    q = ops.scale(q, 1 if gamma < 1e-10 else gamma)
    for (s, y), (rho, a) in zip([p[:2] for p in pairs], reversed(down)):
        b = rho * ops.dot(y, q)
        q = ops.axpy(q, s, a - b)
    return q


def _finish(opt):
    if hasattr(opt, 'endOfProcedure'):
        opt.endOfProcedure()
    elif hasattr(opt, 'endOfIteration'):
        opt.endOfIteration()


def _log_iteration(it, f, step, grad_norm, times=None):
    text = f'iteration {it + 1:d}: obj = {f:.5f}, eps = ' + ('None' if step is None else f'{step:.5f}')
    text += f', gradient: {grad_norm:.5f}'
    if times is not None:
        text += f', time = {times[0]:.04f}, {times[1]:.04f}'
    logging.info(text)


def _fd_gradient_check(opt, ops, grad, f, coeff):
    """Central difference of the objective along a random direction next to
    the gradient's prediction (the reference's TestGradient option)."""
    d = opt.randomDir() if hasattr(opt, 'randomDir') else np.random.normal(size=np.shape(grad))
    h = 1e-8
    fp = opt.updateTry(d, h, f - 1e8)
    fm = opt.updateTry(d, -h, f - 1e8)
    logging.info('Test Gradient: %.4f %.4f' % ((fp - fm) / (2 * h), -ops.dot(grad, d) * coeff))


class _Ahead:
    """The next iteration's quasi-Newton work, queued while the line search
    waits for a trial's values: for trial step t with gradient g, the pair
    (s, y) = (-t d, g - g_old) is pushed, the direction q = twoLoop(g) and
    the dot table of (g, q) launched.  It is used when t is the accepted step
    and g the gradient the search returns; otherwise the push is reverted.
    The values are those the next iteration would compute (the same calls on
    the same inputs), so no decision changes."""

    def __init__(self):
        self.entry = None          # (t, g, s, q, dots, dropped)

    def launch(self, run, opt, ops, direction, t, g):
        self.revert(run)
        s = ops.scale(direction, -t)
        dropped = run.memory.push(s, ops.axpy(g, run.grad_old, -1))
        q = opt.twoLoop(g, run.memory.pairs)
        self.entry = (t, g, s, q, ops.table([g, q], [g, q]), dropped)

    def revert(self, run):
        if self.entry is not None:
            run.memory.undo_push(self.entry[5])
            self.entry = None

    def take(self, run, t, g):
        """(s, q, dots) when the queued work is for step t and gradient g;
        otherwise the queued push is reverted and None returned."""
        e = self.entry
        if e is not None and e[0] == t and e[1] is g:
            self.entry = None
            return e[2], e[3], e[4]
        self.revert(run)
        return None


class _Run:
    """State of one optimisation."""

    def __init__(self, opt, memory):
        self.it = 0
        self.f = None              # objective at the current variable
        self.f_old = None          # objective before the last accepted step
        self.step_taken = None     # s of the last accepted step (-t q)
        self.grad_old = None
        self.grad_next = None      # gradient the line search left at the accepted point
        self.ahead_taken = None    # (q, dots) of the next iteration, queued during the line search
        self.memory = _Memory(memory)
        self.grad_eps = None
        self.history = []
        self.decisions = []        # (iteration, Outcome) after every line search


def bfgs(opt, verb=True, maxIter=1000, TestGradient=False, epsInit=0.01, memory=25, Wolfe=True,
         lineSearch='Weak_Wolfe'):
    """Minimise opt's objective; returns opt.getVariable().  epsInit and
    Wolfe are accepted for the reference's signature: with the weak-Wolfe
    search the first trial step is set as described in the module docstring."""
    required = ('getVariable', 'objectiveFun', 'updateTry', 'acceptVarTry', 'getGradient')
    if not all(hasattr(opt, name) for name in required):
        logging.error('Error: required functions are not provided')
        return None
    if lineSearch != 'Weak_Wolfe':
        logging.warning(f'line search {lineSearch!r} is not built: using weak wolfe condition')
    ops = _VectorOps(opt)
    may_stop = getattr(opt, 'stopCondition', lambda: True)
    coeff = getattr(opt, 'gradCoeff', 1.0)
    burn_in = getattr(opt, 'burnIn', 0)
    if hasattr(opt, 'startOptim'):
        opt.startOptim()

    run = _Run(opt, memory)
    # the next iteration's direction is queued during the line search when the
    # object computes it itself and its fused trial can run work before the
    # read-back
    ahead = _Ahead() if (hasattr(opt, 'twoLoop') and getattr(opt, 'queue_ahead', False)) else None
    if hasattr(opt, 'gradEps'):
        run.grad_eps = opt.gradEps
    elif hasattr(opt, 'options') and 'gradEps' in opt.options:
        run.grad_eps = opt.options['gradEps']
    opt.converged = False
    if getattr(opt, 'reset', False):
        opt.obj = None
    run.f = opt.objectiveFun()
    opt.reset = False
    logging.info('iteration 0: obj = {0: .5f}'.format(run.f))
    run.history.append(run.f)
    clock = time.process_time()

    while run.it < maxIter:
        t_start = time.process_time()
        if hasattr(opt, 'startOfIteration'):
            opt.startOfIteration()
        restarted = opt.reset
        if restarted:
            opt.obj = None
            run.f = opt.objectiveFun()
            if verb:
                logging.info(f"recomputed objective {run.f:.5f}")
            run.f_old = None
            run.grad_next = None

        # directions and gradients are never modified in place (every update
        # makes a new object), so they are shared rather than copied
        grad = opt.getGradient(coeff) if run.grad_next is None else run.grad_next
        if TestGradient:
            _fd_gradient_check(opt, ops, grad, run.f, coeff)

        # direction: quasi-Newton from the stored pairs, steepest descent on
        # the first iteration and after a restart
        if restarted or run.it == 0:
            run.memory.clear()
            q = grad
            dots = None
        elif run.ahead_taken is not None:
            q, dots = run.ahead_taken          # queued during the last line search
        else:
            run.memory.push(run.step_taken, ops.axpy(grad, run.grad_old, -1))
            if hasattr(opt, 'twoLoop'):
                q = opt.twoLoop(grad, run.memory.pairs)
            else:
                q = quasi_newton_direction(grad, run.memory.pairs, ops)
            dots = None
        run.ahead_taken = None
        # |g|^2, |q|^2 and the slope <g, q> along the new direction, launched
        # together; read back when first needed (on most iterations after the
        # line search's first trial, whose own read-back waits for them)
        if dots is None:
            dots = ops.table([grad, q], [grad, q])
        direction = q
        run.grad_old = grad
        if run.it == 0 or run.it == burn_in:
            run.grad_eps = gradient_threshold(run.grad_eps, np.sqrt(dots()[0][0]))
            logging.info(f'Gradient threshold: {run.grad_eps:.6f}')
        t_init = (1. / np.sqrt(np.maximum(1e-20, dots()[1][1]))) if (run.it == 0 or restarted) else 1.

        spec = None
        if ahead is not None:
            spec = lambda t, g, d=direction: ahead.launch(run, opt, ops, d, t, g)
        step, _, _, f_new, _, run.grad_next = line_search_weak_wolfe(
            opt, direction, gfk=grad, old_fval=run.f, old_old_fval=run.f_old, c1=WOLFE_C1, c2=WOLFE_C2,
            amax=None, t_init=t_init, maxiter=WOLFE_MAX_GRADIENTS, slope0=lambda: -dots()[0][1],
            speculate=spec)
        grad_norm = np.sqrt(dots()[0][0])
        if step is not None:
            taken = ahead.take(run, step, run.grad_next) if ahead is not None else None
            if taken is not None:
                run.step_taken, q_next, dots_next = taken
                run.ahead_taken = (q_next, dots_next)
            else:
                run.step_taken = ops.scale(direction, -step)
            run.f_old = run.f
            opt.acceptVarTry()
            run.f = f_new
        else:
            logging.info('Wolfe search unsuccessful')
            run.grad_next = None
            if ahead is not None:
                ahead.revert(run)

        outcome = decide(step is not None, restarted, run.it, burn_in, run.f, run.f_old, grad_norm,
                         run.grad_eps, may_stop)
        run.decisions.append((run.it, outcome))
        if outcome is Outcome.STUCK:
            logging.info('Cannot go any further')
            opt.converged = False
            _finish(opt)
            break
        if outcome is Outcome.STALLED:
            _log_iteration(run.it, run.f, step, grad_norm)
            logging.info('Stopping Gradient Descent: small variation')
            opt.converged = True
            run.history.append(run.f)
            _finish(opt)
            break
        now = time.process_time()
        if verb:
            _log_iteration(run.it, run.f, step, grad_norm, (now - t_start, now - clock))
        clock = now
        if outcome is Outcome.SMALL_GRADIENT:
            logging.info('Stopping Gradient Descent: small gradient')
            opt.converged = True
            run.history.append(run.f)
            _finish(opt)
            break

        opt.reset = outcome is Outcome.RESTART
        if hasattr(opt, 'endOfIteration'):
            opt.endOfIteration()
        if not opt.reset:
            run.it += 1
            run.history.append(run.f)
        if hasattr(opt, 'stopNow') and opt.stopNow():
            opt.converged = True
            break

    if run.it == maxIter and hasattr(opt, 'endOfProcedure'):
        opt.endOfProcedure()
    if hasattr(opt, 'endOptim'):
        opt.endOptim()
    opt.bfgs_iterations = run.it
    opt.bfgs_history = run.history
    opt.bfgs_decisions = run.decisions
    return opt.getVariable()
