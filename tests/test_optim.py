"""Host logic of the optimiser restatement (bfgs + weak-Wolfe), on the CPU.

The optimisable objects here are small numpy problems following the
reference's protocol (base/bfgs.py:9-35).  Checked: convergence on a convex
quadratic and on Rosenbrock, monotone accepted objectives, Armijo on each
accepted step, the reset path (line-search failure -> reset -> stop with
converged False) and the 1e500 NaN guard of updateTry.
"""
import numpy as np
import pytest

from lddmm_amd import bfgs as bfgs_mod
from lddmm_amd.linesearch import line_search_weak_wolfe


class Problem:
    """Minimise f(x) with directions subtracted: x_try = x - t * dir."""

    def __init__(self, f, g, x0, nan_after=None):
        self.f, self.g = f, g
        self.x = np.array(x0, dtype=float)
        self.xTry = None
        self.obj = None
        self.objTry = None
        self.gradCoeff = 1.0
        self.reset = False
        self.accepted = []
        self.nan_after = nan_after
        self.calls = 0

    def getVariable(self):
        return self.x

    def objectiveFun(self):
        if self.obj is None:
            self.obj = self.f(self.x)
        return self.obj

    def updateTry(self, d, t, objRef=None):
        self.calls += 1
        xt = self.x - t * d
        v = self.f(xt)
        if self.nan_after is not None and self.calls > self.nan_after:
            v = np.nan
        if np.isnan(v):
            return 1e500
        if objRef is None or v < objRef:
            self.xTry, self.objTry = xt, v
        return v

    def acceptVarTry(self):
        self.accepted.append((self.obj, self.objTry))
        self.x, self.obj = self.xTry, self.objTry

    def getGradient(self, coeff=1.0, update=None):
        x = self.x if update is None else self.x - update[1] * update[0]
        return self.g(x) / coeff

    def dotProduct_euclidean(self, a, b):
        return [float((a * bb).sum()) for bb in b]


def test_quadratic_converges():
    rng = np.random.default_rng(0)
    Q = rng.normal(size=(20, 20))
    A = Q @ Q.T + np.eye(20)
    b = rng.normal(size=20)
    P = Problem(lambda x: 0.5 * x @ A @ x - b @ x, lambda x: A @ x - b, np.zeros(20))
    bfgs_mod.bfgs(P, maxIter=200, memory=25, verb=False)
    xs = np.linalg.solve(A, b)
    fmin = 0.5 * xs @ A @ xs - b @ xs
    # the reference stops on |obj - obj_old| < 1e-7 (base/bfgs.py:261)
    assert P.obj - fmin < 1e-6
    assert P.converged
    objs = [o for o in P.bfgs_history]
    assert all(objs[i + 1] <= objs[i] + 1e-12 for i in range(len(objs) - 1))


def test_rosenbrock_and_armijo():
    f = lambda x: (1 - x[0]) ** 2 + 100 * (x[1] - x[0] ** 2) ** 2
    g = lambda x: np.array([-2 * (1 - x[0]) - 400 * x[0] * (x[1] - x[0] ** 2), 200 * (x[1] - x[0] ** 2)])
    P = Problem(f, g, [-1.2, 1.0])
    P.burnIn = 0
    bfgs_mod.bfgs(P, maxIter=300, memory=10, verb=False)
    # the absolute 1e-7 variation test stops the run in the flat valley
    assert f(P.x) < 1e-2 < f(np.array([-1.2, 1.0]))
    for before, after in P.accepted:
        assert after <= before


def test_line_search_weak_wolfe_conditions():
    A = np.diag([1.0, 10.0, 100.0])
    P = Problem(lambda x: 0.5 * x @ A @ x, lambda x: A @ x, [1.0, 1.0, 1.0])
    P.objectiveFun()
    g0 = P.getGradient()
    t, fc, gc, fval, f0, gval = line_search_weak_wolfe(P, g0, gfk=g0, old_fval=P.obj, c1=1e-4, c2=0.9,
                                                       t_init=1.0 / np.linalg.norm(g0), maxiter=10)
    assert t is not None
    d0 = -float(g0 @ g0)
    assert fval <= f0 + 1e-4 * t * d0                       # Armijo
    assert -float(gval @ g0) >= 0.9 * d0                    # curvature (weak Wolfe)


def test_nan_objective_guard_and_reset_stop():
    # every trial after the first objective is NaN -> 1e500 -> line search fails,
    # reset, fails again -> stop with converged False
    P = Problem(lambda x: float(x @ x), lambda x: 2 * x, [1.0, 2.0], nan_after=0)
    assert P.updateTry(np.ones(2), 0.1) == 1e500
    P.calls = 0
    bfgs_mod.bfgs(P, maxIter=20, verb=False)
    assert P.converged is False
    assert np.array_equal(P.x, [1.0, 2.0])


def test_mlp_oracle_gradient_fd():
    """The numpy MomentumNet backward against central finite differences."""
    from oracle import mlp
    from lddmm_amd.learned.model import init_params
    rng = np.random.default_rng(0)
    w = init_params(3).astype(np.float64) * 0.5
    # non-zero biases so that every ReLU sees both signs
    lay = mlp.unpack(w)
    for k in lay:
        if k.endswith(".b"):
            lay[k][:] = rng.normal(0, 0.1, lay[k].shape)
    counts = [30, 45]
    X = rng.normal(size=(75, 3))
    T = rng.normal(size=(75, 3))
    L, g = mlp.loss_and_grad(w, w, X, counts, T)
    idx = rng.choice(w.size, 40, replace=False)
    for i in idx:
        e = np.zeros_like(w)
        e[i] = 1e-6
        fd = (mlp.loss_and_grad(w + e, w + e, X, counts, T)[0] - mlp.loss_and_grad(w - e, w - e, X, counts, T)[0]) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd)
