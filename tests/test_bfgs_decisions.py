"""The optimiser's decision rules and quasi-Newton direction on the CPU.

Expectations follow the rules of base/bfgs.py:243-300 as stated in
lddmm_amd.bfgs (restart on a failed line search, stop when a failed search
follows a restart, stall test |df| < 1e-7 that restarts inside the burn-in
and stops after it, small-gradient stop, stopCondition veto), and the
two-loop direction is pinned by the dense BFGS inverse-Hessian recursion.
"""
import numpy as np
import pytest

from lddmm_amd import bfgs as B

O = B.Outcome


class Counter:
    def __init__(self, value=True):
        self.calls = 0
        self.value = value

    def __call__(self):
        self.calls += 1
        return self.value


# ------------------------------------------------------------------ decide
@pytest.mark.parametrize("case,expected", [
    # step_found, after_restart, it, burn_in, f, f_old, |g|, gradEps
    ((False, True, 4, 10, 1.0, None, 1.0, 1e-3), O.STUCK),
    ((False, False, 4, 0, 1.0, 2.0, 1.0, 1e-3), O.RESTART),
    ((True, False, 4, 0, 1.0, 2.0, 1.0, 1e-3), O.ACCEPT),
    ((True, False, 11, 10, 1.0, 1.0 + 5e-8, 1.0, 1e-3), O.STALLED),
    ((True, False, 10, 10, 1.0, 1.0 + 5e-8, 1.0, 1e-3), O.RESTART),     # needs it > burn_in
    ((True, True, 3, 10, 1.0, 1.0 + 5e-8, 1.0, 1e-3), O.STALLED),      # right after a restart
    ((True, False, 3, 10, 1.0, 1.0 + 2e-7, 1.0, 1e-3), O.ACCEPT),      # moved by more than 1e-7
    ((True, False, 3, 0, 1.0, 2.0, 1e-4, 1e-3), O.SMALL_GRADIENT),
    ((False, False, 3, 0, 1.0, 2.0, 1e-4, 1e-3), O.SMALL_GRADIENT),    # failed search, small gradient
    ((True, False, 3, 10, 1.0, 1.0, 1e-4, 1e-3), O.SMALL_GRADIENT),    # stall restarts, gradient stops
    ((True, False, 0, 0, 1.0, None, 1.0, 1e-3), O.ACCEPT),             # no previous objective
])
def test_decide_rules(case, expected):
    may = Counter()
    assert B.decide(*case, may) is expected


def test_decide_stop_condition_veto_and_order():
    may = Counter(False)
    assert B.decide(True, False, 20, 0, 1.0, 1.0, 1e-9, 1e-3, may) is O.ACCEPT
    assert may.calls == 2                       # asked for the stall, then for the gradient rule
    may = Counter(True)
    assert B.decide(True, False, 20, 0, 1.0, 1.0, 1e-9, 1e-3, may) is O.STALLED
    assert may.calls == 1
    may = Counter(True)
    assert B.decide(True, False, 20, 0, 1.0, 3.0, 5.0, 1e-3, may) is O.ACCEPT
    assert may.calls == 0                       # no rule held


def test_gradient_threshold():
    assert B.gradient_threshold(None, 50.0) == pytest.approx(0.05)
    assert B.gradient_threshold(None, 0.01) == 1e-4
    assert B.gradient_threshold(1e-2, 50.0) == pytest.approx(5e-3)
    assert B.gradient_threshold(1e-3, 50.0) == 1e-3


# --------------------------------------------------------- direction
class NumpyOps:
    def dot(self, u, v):
        return float(u @ v)

    def axpy(self, u, v, c):
        return u + c * v

    def scale(self, u, c):
        return c * u

    def copy(self, u):
        return u.copy()


@pytest.mark.parametrize("m", [1, 2, 5])
def test_two_loop_equals_dense_bfgs_inverse(m):
    rng = np.random.default_rng(m)
    n = 7
    Q = rng.normal(size=(n, n))
    A = Q @ Q.T + n * np.eye(n)                  # y = A s keeps s.y > 0
    pairs = []
    for _ in range(m):
        s = rng.normal(size=n)
        pairs.append([s, A @ s])
    g = rng.normal(size=n)
    s, y = pairs[-1]
    H = (s @ y) / (y @ y) * np.eye(n)
    for s, y in pairs:
        rho = 1 / (y @ s)
        V = np.eye(n) - rho * np.outer(y, s)
        H = V.T @ H @ V + rho * np.outer(s, s)
    q = B.quasi_newton_direction(g, pairs, NumpyOps())
    assert np.allclose(q, H @ g, rtol=1e-12, atol=1e-12)
    # the secant equation of the newest pair: H y = s
    assert np.allclose(B.quasi_newton_direction(pairs[-1][1], pairs, NumpyOps()), pairs[-1][0], rtol=1e-10)


def test_memory_keeps_newest_pairs():
    mem = B._Memory(3)
    for k in range(5):
        mem.push(k, -k)
    assert [p[0] for p in mem.pairs] == [2, 3, 4]
    mem.clear()
    assert len(mem) == 0


# ------------------------------------------------------------ driver runs
class Quadratic:
    """f(x) = 0.5 x.A x - b.x with directions subtracted (x_try = x - t d);
    trial objectives can be forced to fail (1e500, the NaN guard value)."""

    def __init__(self, A, b, x0, fail=lambda it, k: False):
        self.A, self.b = np.asarray(A, float), np.asarray(b, float)
        self.x = np.array(x0, float)
        self.obj = None
        self.gradCoeff = 1.0
        self.reset = False
        self.fail = fail
        self.iteration = -1
        self.trials = []               # (iteration, t) of every updateTry
        self.events = []
        self.x_starts = []
        self.recomputed = 0
        self.procedure_ends = 0

    def f(self, x):
        return 0.5 * x @ self.A @ x - self.b @ x

    def getVariable(self):
        return self.x

    def objectiveFun(self):
        if self.obj is None:
            self.recomputed += 1
            self.obj = self.f(self.x)
        return self.obj

    def updateTry(self, d, t, objRef=None):
        self.trials.append((self.iteration, t))
        if self.fail(self.iteration, len(self.trials)):
            return 1e500
        self.xTry = self.x - t * d
        self.objTry = self.f(self.xTry)
        return self.objTry

    def acceptVarTry(self):
        self.events.append('accept')
        self.x, self.obj = self.xTry, self.objTry

    def getGradient(self, coeff=1.0, update=None):
        x = self.x if update is None else self.x - update[1] * update[0]
        return (self.A @ x - self.b) / coeff

    def dotProduct_euclidean(self, a, bs):
        return [float(a @ b) for b in bs]

    def startOfIteration(self):
        self.iteration += 1
        self.x_starts.append(self.x.copy())
        self.events.append('restart-start' if self.reset else 'start')

    def endOfProcedure(self):
        self.procedure_ends += 1


def _problem(seed=0, n=12, **kw):
    rng = np.random.default_rng(seed)
    Q = rng.normal(size=(n, n))
    return Quadratic(Q @ Q.T + np.eye(n), rng.normal(size=n), np.zeros(n), **kw)


def test_failed_line_search_restarts_then_recovers():
    P = _problem(fail=lambda it, k: it == 2)      # every trial of the third iteration fails
    B.bfgs(P, maxIter=100, memory=10, verb=False)
    d = P.bfgs_decisions
    assert d[:3] == [(0, O.ACCEPT), (1, O.ACCEPT), (2, O.RESTART)]
    assert d[3][0] == 2                            # the restart repeats iteration 2 (not counted)
    assert d[3][1] is O.ACCEPT
    assert P.events[P.events.index('restart-start') - 1] == 'start'
    assert P.recomputed == 2                       # initial objective + the one after the restart
    assert P.converged and d[-1][1] in (O.STALLED, O.SMALL_GRADIENT)
    objs = P.bfgs_history
    assert all(objs[i + 1] <= objs[i] for i in range(len(objs) - 1))
    assert P.bfgs_iterations == len(objs) - 2      # stop appends the final objective once more


def test_restart_uses_steepest_descent_step():
    P = _problem(seed=1, fail=lambda it, k: it == 1)
    B.bfgs(P, maxIter=4, memory=10, verb=False)
    # iteration index 2 is the repeated one; its first trial is 1/|g| at the point reached after it 0
    P2 = _problem(seed=1)
    B.bfgs(P2, maxIter=1, memory=10, verb=False)
    x_before = P2.x
    g = P2.A @ x_before - P2.b
    first_after = [t for it, t in P.trials if it == 2][0]
    assert first_after == pytest.approx(1.0 / np.linalg.norm(g), rel=1e-12)
    assert [t for it, t in P.trials if it == 1][0] == 1.0          # quasi-Newton unit step first


def test_failure_right_after_restart_stops_unconverged():
    P = _problem(fail=lambda it, k: True)
    x0 = P.x.copy()
    B.bfgs(P, maxIter=20, verb=False)
    assert P.bfgs_decisions == [(0, O.RESTART), (0, O.STUCK)]
    assert P.converged is False
    assert P.procedure_ends == 1
    assert np.array_equal(P.x, x0)
    assert P.bfgs_iterations == 0


def test_stall_inside_burn_in_restarts_then_stops():
    A = 1e-6 * np.diag([2.0, 20.0])
    P = Quadratic(A, np.zeros(2), [1.0, 1.0])
    P.burnIn = 10
    P.gradEps = 1e-30                             # only the stall rule can stop
    B.bfgs(P, maxIter=50, verb=False)
    d = P.bfgs_decisions
    k = d[-1][0]
    assert d[-2:] == [(k, O.RESTART), (k, O.STALLED)]
    assert k <= P.burnIn
    assert all(o is O.ACCEPT for _, o in d[:-2])
    assert P.converged and P.bfgs_iterations == k


def test_small_gradient_stop_and_threshold():
    # a flat quadratic (eigenvalues ~1e-4): the objective still moves by far
    # more than 1e-7 per step when the gradient drops below the threshold
    P = _problem(seed=3)
    P.A *= 1e-4
    P.b *= 10.0
    g0 = np.linalg.norm(P.A @ P.x - P.b)
    P.gradEps = 1e-2
    B.bfgs(P, maxIter=200, verb=False)
    thr = min(1e-2, 1e-4 * g0)
    assert P.bfgs_decisions[-1][1] is O.SMALL_GRADIENT
    assert P.converged
    # the stopping iteration began from a gradient below the threshold, the one before it did not
    gn = [np.linalg.norm(P.A @ x - P.b) for x in P.x_starts]
    assert gn[-1] < thr <= gn[-2]


def test_stop_condition_veto():
    calls = Counter(False)
    P = _problem(seed=4)
    P.stopCondition = calls
    B.bfgs(P, maxIter=40, verb=False)
    outcomes = [o for _, o in P.bfgs_decisions]
    assert calls.calls > 0                        # a stopping rule held and was vetoed
    assert O.STALLED not in outcomes and O.SMALL_GRADIENT not in outcomes
    assert P.converged is False
    # the run ends at maxIter or when even steepest descent finds no step
    assert P.bfgs_iterations == 40 or outcomes[-2:] == [O.RESTART, O.STUCK]
    assert P.procedure_ends == 1


# ------------------------------------------------------- queued-ahead direction
class _Rosenbrock:
    """Chained Rosenbrock objective with the fused-trial protocol of the
    matching drivers (updateTryWithSlope with before_readback, twoLoop):
    enough curvature that the weak-Wolfe search rejects and extends trials."""

    def __init__(self, n=12, queue_ahead=True):
        self.x = np.linspace(-1.2, 1.0, n)
        self.queue_ahead = queue_ahead
        self.trial = None
        self.calls = {"ahead": 0}

    @staticmethod
    def f(x):
        return float(np.sum(100.0 * (x[1:] - x[:-1] ** 2) ** 2 + (1.0 - x[:-1]) ** 2))

    @staticmethod
    def g(x):
        d = np.zeros_like(x)
        r = x[1:] - x[:-1] ** 2
        d[:-1] += -400.0 * x[:-1] * r - 2.0 * (1.0 - x[:-1])
        d[1:] += 200.0 * r
        return d

    def getVariable(self):
        return self.x

    def objectiveFun(self):
        return self.f(self.x)

    def updateTry(self, d, t, objRef=None):
        self.trial = self.x - t * d
        return self.f(self.trial)

    def acceptVarTry(self):
        self.x = self.trial

    def getGradient(self, coeff=1.0, update=None):
        x = self.x if update is None else self.x - update[1] * update[0]
        return self.g(x) / coeff

    def dotProduct_euclidean(self, u, vs):
        return [float(np.dot(u, v)) for v in vs]

    def updateTryWithSlope(self, d, t, coeff=1.0, before_readback=None):
        f = self.updateTry(d, t)
        g = self.getGradient(coeff, update=[d, t])
        if before_readback is not None:
            self.calls["ahead"] += 1
            before_readback(g)
        return f, -float(np.dot(g, d)), g

    def twoLoop(self, g, pairs):
        return B.quasi_newton_direction(g, pairs, B._VectorOps(self))


def test_queued_ahead_direction_keeps_every_decision():
    """bfgs with the next iteration's direction queued during the line search
    (the object's queue_ahead) against the same run without it: the same
    objective history, decisions and final variable, with rejected trials
    and a memory window that slides (memory 3)."""
    runs = []
    for ahead in (False, True):
        obj = _Rosenbrock(queue_ahead=ahead)
        x = B.bfgs(obj, verb=False, maxIter=60, memory=3)
        runs.append((np.array(obj.bfgs_history), obj.bfgs_decisions, np.array(x), obj.calls["ahead"]))
    (h0, d0, x0, a0), (h1, d1, x1, a1) = runs
    assert a0 == 0 and a1 > len(h1)              # queued on every trial, rejected ones included
    np.testing.assert_array_equal(h1, h0)
    assert d1 == d0
    np.testing.assert_array_equal(x1, x0)
    assert any(o is B.Outcome.ACCEPT for _, o in d0) and len(h0) > 20
