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


def train(samples, epochs=300, batch=8, lr=2e-3, seed=0, log_every=50):
    """Fit MomentumNet on (features -> a0) pairs with the package's HIP kernels."""
    rng = np.random.default_rng(seed)
    F = np.concatenate([s["features"] for s in samples])
    A = np.concatenate([s["a0"] for s in samples])
    feat_scale = float(np.sqrt((F ** 2).mean())) or 1.0
    out_scale = float(np.sqrt((A ** 2).mean())) or 1.0
    net = MomentumNet(seed=seed, feat_scale=feat_scale, out_scale=out_scale)
    dev = net.dev
    Xs = [torch.tensor(s["features"] / feat_scale, dtype=torch.float32, device=dev) for s in samples]
    Ys = [torch.tensor(s["a0"] / out_scale, dtype=torch.float32, device=dev) for s in samples]
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


def warm_control(sm, net):
    """EPDiff control from the predicted momentum, or None when the zero control is better."""
    feats = features(sm)
    a0 = net.predict(feats, [feats.shape[0]])
    T = sm.Tsize
    _, at = evol.landmarkEPDiff(T, sm.x0_dev, a0, sm.options['KparDiff'])
    return at, a0


def register(seed, net=None, target_data=None, cfg=DEFAULT, max_iter=None):
    """Registration (plain or warm-started) until the data term reaches target_data (1e-4 rel)."""
    c = dict(cfg)
    if max_iter is not None:
        c['maxIter'] = max_iter
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sm = make_matching(seed, c)
    used_warm = False
    if net is not None:
        at, _ = warm_control(sm, net)
        obj_zero = sm.objectiveFun()
        sm.setInitialControl(at)
        obj_warm = sm.objectiveFun()
        if obj_warm < obj_zero:
            used_warm = True
        else:                                   # fallback: zero momentum
            sm.setInitialControl(torch.zeros_like(at))
    if target_data is not None:
        sm.setTargetDataEnergy(target_data, 1e-4)
    sm.optimizeMatching()
    torch.cuda.synchronize()
    d = sm.varifold_data()
    reached = target_data is None or d <= target_data * (1 + 1e-4)
    return {"seed": seed, "iterations": int(sm.bfgs_iterations), "seconds": time.perf_counter() - t0,
            "data": d, "reached": bool(reached), "used_warm": used_warm}
