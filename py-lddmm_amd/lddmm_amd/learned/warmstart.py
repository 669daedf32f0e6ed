"""Data generation, training and evaluation of the learned warm start.

Pairs are the plain surface-matching workload (icosphere subdivision 4,
R = 10, seeded Gaussian bumps) with one seed per pair.  Features are the
initial data-term force -grad D(x0) / sigma_err^2 at the template vertices
(no vertex correspondence is assumed); targets are the initial control a_0 =
at[0] that the plain optimiser reaches.  A predicted a0 is turned into a
control with EPDiff (landmarkEPDiff), the zero control is kept whenever it
has the lower objective (the fallback), and BFGS then runs until the
varifold data term matches the plain method's final value within 1e-4.
"""
import logging
import time

import numpy as np
import torch

from .. import _backend as B
from .. import pointEvolution as evol
from .. import synth
from ..kernelFunctions import Kernel
from ..surfaces import Surface
from ..surfaceMatching import SurfaceMatching
from .model import MomentumNet, Segments

DEFAULT = dict(subdiv=4, radius=10.0, nbumps=20, width=3.0, amp_sd=1.0, sigma_def=2.5, sigma_var=1.0,
               timeStep=0.1, maxIter=50, pk_dtype='float64')


def make_matching(seed, cfg=DEFAULT, at=None):
    (v, f), (v2, _) = synth.plain_matching_pair(seed=seed, subdiv=cfg['subdiv'], radius=cfg['radius'],
                                                nbumps=cfg['nbumps'], width=cfg['width'], amp_sd=cfg['amp_sd'])
    opts = {'KparDiff': Kernel(name='gauss', sigma=cfg['sigma_def']),
            'KparDist': Kernel(name='gauss', sigma=cfg['sigma_var']),
            'errorType': 'varifold', 'timeStep': cfg['timeStep'], 'pk_dtype': cfg['pk_dtype'],
            'maxIter': cfg['maxIter'], 'mode': 'quiet'}
    return SurfaceMatching(Template=Surface((f, v)), Target=Surface((f, v2)), options=opts)


def features(sm):
    """Initial data-term force at the template vertices, [N, 3] float64 device tensor."""
    return -sm.endPointGradient_device(sm.x0_dev)


def velocity(sm, a0):
    """v0 = K(x0) a0 (device float64)."""
    return sm.options['KparDiff'].applyK(sm.x0_dev, B.to_dev(a0))


def plain_registration(seed, cfg=DEFAULT):
    """Zero-momentum plain registration: returns dict with features, a0 target,
    final data term, iterations, seconds."""
    sm = make_matching(seed, cfg)
    feats = features(sm)
    t0 = time.perf_counter()
    sm.optimizeMatching()
    torch.cuda.synchronize()
    return {"seed": seed, "features": feats.cpu().numpy(), "a0": sm.control['at'][0].cpu().numpy(),
            "data": sm.varifold_data(), "iterations": int(sm.bfgs_iterations),
            "seconds": time.perf_counter() - t0, "history": list(sm.bfgs_history)}


def train(samples, epochs=300, batch=8, lr=2e-3, seed=0, log_every=50, target='momentum'):
    """Fit MomentumNet on (features -> a0, or -> v0 = K(x0) a0) pairs with the package's HIP kernels."""
    rng = np.random.default_rng(seed)
    key = "a0" if target == 'momentum' else "v0"
    F = np.concatenate([s["features"] for s in samples])
    A = np.concatenate([s[key] for s in samples])
    feat_scale = float(np.sqrt((F ** 2).mean())) or 1.0
    out_scale = float(np.sqrt((A ** 2).mean())) or 1.0
    net = MomentumNet(seed=seed, feat_scale=feat_scale, out_scale=out_scale)
    net.target = target
    dev = net.dev
    Xs = [torch.tensor(s["features"] / feat_scale, dtype=torch.float32, device=dev) for s in samples]
    Ys = [torch.tensor(s[key] / out_scale, dtype=torch.float32, device=dev) for s in samples]
    losses = []
    for ep in range(epochs):
        order = rng.permutation(len(samples))
        tot = 0.0
        for b0 in range(0, len(order), batch):
            idx = order[b0:b0 + batch]
            X = torch.cat([Xs[i] for i in idx])
            Y = torch.cat([Ys[i] for i in idx])
            segs = Segments([Xs[i].shape[0] for i in idx], dev)
            loss = net.loss_and_grad(X, segs, Y)
            net.adam(lr=lr)
            tot += float(loss) * len(idx)
        losses.append(tot / len(order))
        if log_every and (ep % log_every == 0 or ep == epochs - 1):
            logging.info(f"warm-start training epoch {ep}: loss {losses[-1]:.5f}")
    return net, losses


def warm_control(sm, net, cg_iters=20):
    """EPDiff control from the predicted initial momentum.

    A net trained on velocities (net.target == 'velocity') predicts v0 = K(x0) a0,
    which is local in the data force; a0 then comes from a conjugate-gradient
    solve of K(x0) a0 = v0 (Kernel.solve), truncated at cg_iters products."""
    feats = features(sm)
    pred = net.predict(feats, [feats.shape[0]])
    if getattr(net, 'target', 'momentum') == 'velocity':
        a0 = sm.options['KparDiff'].solve(sm.x0_dev, pred, iterMax=cg_iters)
    else:
        a0 = pred
    T = sm.Tsize
    _, at = evol.landmarkEPDiff(T, sm.x0_dev, a0, sm.options['KparDiff'])
    return at, a0


def register(seed, net=None, target_data=None, cfg=DEFAULT, max_iter=None, a0=None, cg_iters=20, at=None):
    """Registration (plain, warm-started by `net`, or warm-started from a given
    initial momentum `a0`) until the data term reaches target_data (1e-4 rel).

    The warm control is kept only when its objective is below the zero
    control's (zero-momentum fallback).  `data0` is the data term at the
    starting control; a start already within tolerance counts 0 iterations."""
    c = dict(cfg)
    if max_iter is not None:
        c['maxIter'] = max_iter
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sm = make_matching(seed, c)
    used_warm = False
    if net is not None or a0 is not None or at is not None:
        if at is not None:
            at = B.to_dev(at)
        elif a0 is None:
            at, _ = warm_control(sm, net, cg_iters=cg_iters)
        else:
            _, at = evol.landmarkEPDiff(sm.Tsize, sm.x0_dev, B.to_dev(a0), sm.options['KparDiff'])
        obj_zero = sm.objectiveFun()
        sm.setInitialControl(at)
        obj_warm = sm.objectiveFun()
        if obj_warm < obj_zero:
            used_warm = True
        else:                                   # fallback: zero momentum
            sm.setInitialControl(torch.zeros_like(at))
    data0 = sm.varifold_data()
    tol = None if target_data is None else target_data * (1 + 1e-4)
    if tol is not None and data0 <= tol:
        iterations = 0
    else:
        if target_data is not None:
            sm.setTargetDataEnergy(target_data, 1e-4)
        sm.optimizeMatching()
        iterations = int(sm.bfgs_iterations)
    torch.cuda.synchronize()
    d = sm.varifold_data()
    reached = tol is None or d <= tol
    return {"seed": seed, "iterations": iterations, "seconds": time.perf_counter() - t0,
            "data": d, "data0": data0, "reached": bool(reached), "used_warm": used_warm}
