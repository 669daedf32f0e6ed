"""Data generation, training and evaluation of the learned warm start.

Pairs are the plain surface-matching workload (icosphere subdivision 4,
R = 10, seeded Gaussian bumps) with one seed per pair; no vertex
correspondence between template and target is assumed.

Feature-only control (no network): the ray-cast normal height h of the
target over each template vertex, taken as the velocity h n and turned into
a time-constant control by a conjugate-gradient solve on K(x0) truncated at
SOLVE_ITERS products (Kernel.solve; the truncation keeps the control smooth).

Default model (features 'normal', target 'vresid'): the network reads per
vertex h and its smoothed versions and predicts, in the vertex's local frame,
the residual that turns h n into the refined velocity: the fixed point of
re-ray-casting from the surface that the current velocity's control reaches
(refined_velocity).  The warm control is then solved from h n plus the
prediction the same way, so a zero prediction reproduces the feature-only
control.  The zero control is kept whenever it has the lower objective (the
fallback), and BFGS then runs until the varifold data term matches the plain
method's final value within 1e-4.

Older targets ('vlocal': the plain optimiser's time-mean velocity in the
local frame, and the world-frame targets) are kept for the experiment
driver; they learn the plain optimiser's balanced solution, which through
the truncated solve is a worse starting point than h n itself.
"""
import functools
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
from . import features as NF
from .model import MomentumNet, Segments

DEFAULT = dict(subdiv=4, radius=10.0, nbumps=20, width=3.0, amp_sd=1.0, sigma_def=2.5, sigma_var=1.0,
               timeStep=0.1, maxIter=50, pk_dtype='float64')


@functools.lru_cache(maxsize=64)
def _pair(seed, subdiv, radius, nbumps, width, amp_sd):
    (v, f), (v2, _) = synth.plain_matching_pair(seed=seed, subdiv=subdiv, radius=radius, nbumps=nbumps, width=width,
                                                amp_sd=amp_sd)
    for arr in (v, f, v2):
        arr.flags.writeable = False
    return v, f, v2


def pair_arrays(seed, cfg=DEFAULT):
    """Template vertices, faces and target vertices of the seeded pair (cached, read-only)."""
    return _pair(int(seed), cfg['subdiv'], float(cfg['radius']), cfg['nbumps'], float(cfg['width']),
                 float(cfg['amp_sd']))


def make_matching(seed, cfg=DEFAULT, at=None):
    v, f, v2 = pair_arrays(seed, cfg)
    opts = {'KparDiff': Kernel(name='gauss', sigma=cfg['sigma_def']),
            'KparDist': Kernel(name='gauss', sigma=cfg['sigma_var']),
            'errorType': 'varifold', 'timeStep': cfg['timeStep'], 'pk_dtype': cfg['pk_dtype'],
            'maxIter': cfg['maxIter'], 'mode': 'quiet'}
    return SurfaceMatching(Template=Surface((f, v)), Target=Surface((f, v2)), options=opts)


FEATURES = ('force', 'meanshift', 'kforce', 'normal', 'normal_lin')
TARGETS = ('vresid', 'abar', 'vbar', 'vlocal', 'momentum', 'velocity')
TARGET_KEY = {"vresid": "vresid", "abar": "abar", "vbar": "vbar", "vlocal": "vlocal", "momentum": "a0",
              "velocity": "v0"}
SOLVE_ITERS = 10      # conjugate-gradient products that turn a velocity into a time-constant control
REFINE_STEPS = 2      # re-ray-casts behind the 'vresid' training target
NET_REFINE_STEPS = 1  # re-ray-casts after the network's prediction in the warm start
MEANSHIFT_SIGMA = 1.5


def features(sm, kind='force'):
    """Per-vertex input channels [N, 3] (float64 device tensor), no correspondence assumed:
    force     - initial data-term force -grad D(x0) / sigma_err^2;
    meanshift - Gaussian-weighted mean of the target's vertices around each
                template vertex minus the vertex (width MEANSHIFT_SIGMA);
    kforce    - the force smoothed by the deformation kernel, K(x0) force;
    normal    - ray-cast normal height of the target over the template and
                its normalised Gaussian averages (learned.features);
    normal_lin - the height h, the part of it that the feature-only
                control's initial velocity misses, h - n.K(x0) a_f
                (linear_miss), and h averaged at width 2.5."""
    if kind in ('normal', 'normal_lin'):
        return height_inputs(sm, kind)
    if kind == 'force':
        return -sm.endPointGradient_device(sm.x0_dev)
    if kind == 'kforce':
        return sm.options['KparDiff'].applyK(sm.x0_dev, -sm.endPointGradient_device(sm.x0_dev))
    if kind == 'meanshift':
        from .. import kernelFunctions_util as ku
        y = B.to_dev(sm.fv1.vertices)
        w = torch.cat([y, torch.ones_like(y[:, :1])], 1)
        s = ku.applyK(sm.x0_dev, y, w, 'gauss', [MEANSHIFT_SIGMA], 10)
        return s[:, :3] / s[:, 3:4].clamp_min(1e-300) - sm.x0_dev
    raise ValueError(f"unknown feature kind {kind!r}")


def frame(sm):
    """Local frames [N, 3, 3] (rows: normal, two tangents) of the template vertices (device)."""
    return B.to_dev(NF.local_frame(synth.vertex_normals(sm.fv0.vertices, sm.fv0.faces)))


def time_constant(sm, a):
    """Control [T, N, 3] that repeats the momentum a at every time step."""
    return a[None].expand(sm.Tsize, -1, -1).contiguous()


def solve_control(sm, v, iters=SOLVE_ITERS):
    """Momentum a with K(x0) a ~ v: conjugate gradients truncated at `iters` products."""
    return sm.options['KparDiff'].solve(sm.x0_dev, v, iterMax=iters, sequence='textbook')


def ray_cast_heights(sm):
    """Normal-height channels [N, 3] and local frames [N, 3, 3] of the pair (device)."""
    return NF.surface_normal_heights(sm.fv0, sm.fv1)


def linear_miss(sm, ch, fr, a=None):
    """h - n.K(x0) a_f [N, 1]: the normal height that the initial velocity of
    the feature-only control a_f (solved from h n) does not reproduce."""
    n = fr[:, 0]
    if a is None:
        a = solve_control(sm, ch[:, :1] * n)
    return ch[:, :1] - (sm.options['KparDiff'].applyK(sm.x0_dev, a) * n).sum(1, keepdim=True)


def height_inputs(sm, kind='normal', heights=None):
    """Network inputs [N, 3] built on the ray-cast heights (features 'normal'
    and 'normal_lin'); `heights` = the pair's (channels, frames) when known."""
    ch, fr = ray_cast_heights(sm) if heights is None else heights
    if kind == 'normal':
        return ch
    if kind == 'normal_lin':
        return torch.cat([ch[:, :1], linear_miss(sm, ch, fr), ch[:, 1:2]], 1)
    raise ValueError(f"feature kind {kind!r} is not built on the ray-cast heights")


def feature_only_control(sm):
    """The warm control without a network: the velocity h n of the ray-cast
    normal heights, solved for a time-constant control.  Returns (at, a)."""
    ch, fr = ray_cast_heights(sm)
    a = solve_control(sm, ch[:, :1] * fr[:, 0])
    return time_constant(sm, a), a


def refined_velocity(sm, steps=REFINE_STEPS, heights=None):
    """Velocity refined by re-ray-casting, started from v = h n (see
    refine_velocity).  `heights` = the pair's (channels, frames) when already
    computed.  Returns v [N, 3] (device)."""
    ch, fr = ray_cast_heights(sm) if heights is None else heights
    return refine_velocity(sm, ch[:, :1] * fr[:, 0], steps)


def refine_velocity(sm, v, steps):
    """`steps` re-ray-casts of the velocity v [N, 3]: each shoots the
    time-constant control solved from v, ray-casts the target along the unit
    vertex normals n_T of the surface it reaches (a missed ray has height 0)
    and adds that height: v <- v + h_T n_T."""
    K = sm.options['KparDiff']
    for _ in range(steps):
        xt, _ = evol.shoot(sm.x0_dev, time_constant(sm, solve_control(sm, v)), K)
        xT = xt[-1].contiguous()
        nT = NF.device_vertex_normals(sm.fv0, xT)
        hT, _ = NF.normal_raycast(xT, nT, *NF.surface_on_device(sm.fv1))
        v = v + hT[:, None] * nT
    return v


def generate_refined(seeds, cfg=DEFAULT, steps=REFINE_STEPS, progress=None):
    """Training samples {feat_normal, feat_normal_lin, vresid} of the pairs `seeds`: the
    residual refined_velocity - h n in the local frames.  No registration is
    run: each pair costs `steps` shoots and ray casts."""
    samples = []
    for k, s in enumerate(seeds):
        sm = make_matching(s, cfg)
        ch, fr = ray_cast_heights(sm)
        r = refined_velocity(sm, steps, heights=(ch, fr)) - ch[:, :1] * fr[:, 0]
        samples.append({"feat_normal": ch.cpu().numpy(),
                        "feat_normal_lin": height_inputs(sm, 'normal_lin', (ch, fr)).cpu().numpy(),
                        "vresid": torch.einsum('nkj,nj->nk', fr, r).cpu().numpy()})
        if progress is not None and (k + 1) % 1000 == 0:
            progress(f"refined targets: {k + 1} / {len(seeds)}")
    return samples


def velocity(sm, a0):
    """v0 = K(x0) a0 (device float64)."""
    return sm.options['KparDiff'].applyK(sm.x0_dev, B.to_dev(a0))


def targets_of(sm, at):
    """Training targets of a finished plain registration with control `at` [T, N, 3]."""
    a0 = at[0]
    abar = at.mean(0)
    vbar = velocity(sm, abar)
    vlocal = torch.einsum('nkj,nj->nk', frame(sm), vbar)
    return {"a0": a0.cpu().numpy(), "abar": abar.cpu().numpy(), "v0": velocity(sm, a0).cpu().numpy(),
            "vbar": vbar.cpu().numpy(), "vlocal": vlocal.cpu().numpy(), "x0": sm.x0_dev.cpu().numpy()}


def generate_batched(seeds, cfg=DEFAULT, batch=256, progress=None):
    """Training samples {feat_normal, vlocal} of the plain registrations of
    `seeds`, computed `batch` pairs at a time by the lockstep population
    registration (population.PopulationMatching: per pair the same BFGS
    iterates as SurfaceMatching, in one persistent launch per kernel sum)."""
    from ..population import PopulationMatching
    samples = []
    for b0 in range(0, len(seeds), batch):
        ss = list(seeds[b0:b0 + batch])
        prs = [synth.plain_matching_pair(seed=s, subdiv=cfg['subdiv'], radius=cfg['radius'], nbumps=cfg['nbumps'],
                                         width=cfg['width'], amp_sd=cfg['amp_sd']) for s in ss]
        Kd = Kernel(name='gauss', sigma=cfg['sigma_def'])
        pm = PopulationMatching([Surface((f, v)) for (v, f), _ in prs], [Surface((f, v2)) for (_, f), (v2, _) in prs],
                                {'KparDiff': Kd, 'KparDist': Kernel(name='gauss', sigma=cfg['sigma_var']),
                                 'timeStep': cfg['timeStep'], 'pk_dtype': cfg['pk_dtype'], 'maxIter': cfg['maxIter']})
        pm.optimizeMatching()
        for p, ((v, f), (v2, _)) in enumerate(prs):
            ch, fr = NF.normal_heights(v, f, v2, f)
            vbar = Kd.applyK(B.to_dev(v), pm.control_of(p).mean(0))
            samples.append({"feat_normal": ch.cpu().numpy(),
                            "vlocal": torch.einsum('nkj,nj->nk', fr, vbar).cpu().numpy()})
        pm.pop.close()
        if progress is not None:
            progress(f"batched plain registrations: {len(samples)} / {len(seeds)}")
    return samples


def plain_registration(seed, cfg=DEFAULT):
    """Zero-momentum plain registration: every feature kind, the targets
    (abar, a0, v0), final data term, iterations, seconds."""
    sm = make_matching(seed, cfg)
    feats = {f"feat_{k}": features(sm, k).cpu().numpy() for k in FEATURES}
    t0 = time.perf_counter()
    sm.optimizeMatching()
    torch.cuda.synchronize()
    out = {"seed": seed, "data": sm.varifold_data(), "iterations": int(sm.bfgs_iterations),
           "seconds": time.perf_counter() - t0, "history": list(sm.bfgs_history)}
    out.update(feats)
    out.update(targets_of(sm, sm.control['at']))
    return out


def train(samples, epochs=300, batch=8, lr=2e-3, seed=0, log_every=50, target='abar', features='force',
          progress=None, loss='mse', sigma_def=None):
    """Fit MomentumNet on (features -> target) pairs with the package's HIP kernels.
    Each sample holds every feature kind under "feat_<kind>" and the targets
    "abar", "a0", "v0" and "vbar" (and the template vertices "x0" for loss='kvel').

    loss 'mse': mean squared error of the output.  loss 'kvel' (momentum
    targets): the error is measured after the deformation kernel,
    |K(x0) (out - target)|^2, i.e. on the velocity field the momentum induces;
    the momentum's high-frequency detail, which barely moves the shape, no
    longer dominates the loss.  Its gradient 2 K(x0)^T K(x0) (out - target) takes
    two kernel sums per pair (float32 compensated)."""
    rng = np.random.default_rng(seed)
    key = TARGET_KEY[target]
    fkey = "feat_" + features
    F = np.concatenate([s[fkey] for s in samples])
    A = np.concatenate([s[key] for s in samples])
    feat_scale = float(np.sqrt((F ** 2).mean())) or 1.0
    out_scale = float(np.sqrt((A ** 2).mean())) or 1.0
    net = MomentumNet(seed=seed, feat_scale=feat_scale, out_scale=out_scale, target=target, features=features,
                      solve_iters=SOLVE_ITERS if target in ('vlocal', 'vresid') else 0)
    dev = net.dev
    Xs = [torch.tensor(s[fkey] / feat_scale, dtype=torch.float32, device=dev) for s in samples]
    Ys = [torch.tensor(s[key] / out_scale, dtype=torch.float32, device=dev) for s in samples]
    losses = []
    if loss == 'kvel':
        from ..kernelFunctions import Kernel
        Kk = Kernel(name='gauss', sigma=sigma_def if sigma_def is not None else DEFAULT['sigma_def'])
        Kk.pk_dtype = 'float32'
        X0 = [B.to_dev(s["x0"]) for s in samples]
        # normalise by the targets' mean squared velocity
        vn = np.mean([float((Kk.applyK(X0[i], Ys[i].double()) ** 2).mean()) for i in range(len(samples))])
        kscale = 1.0 / max(vn, 1e-300)
    for ep in range(epochs):
        order = rng.permutation(len(samples))
        tot = 0.0
        for b0 in range(0, len(order), batch):
            idx = order[b0:b0 + batch]
            X = torch.cat([Xs[i] for i in idx])
            Y = torch.cat([Ys[i] for i in idx])
            counts = [Xs[i].shape[0] for i in idx]
            segs = Segments(counts, dev)
            if loss == 'kvel':
                out = net.forward(X, segs)
                diff = (out - Y).double()
                grads, lsum, o = [], 0.0, 0
                ntot = diff.numel()
                for i, c in zip(idx, counts):
                    d = diff[o:o + c]
                    u = Kk.applyK(X0[i], d)
                    lsum = lsum + (u * u).sum()
                    grads.append(Kk.applyK(X0[i], u))
                    o += c
                lval = lsum * kscale / ntot
                net.backward((torch.cat(grads) * (2.0 * kscale / ntot)).float())
            else:
                lval = net.loss_and_grad(X, segs, Y)
            net.adam(lr=lr)
            tot += float(lval) * len(idx)
        losses.append(tot / len(order))
        if log_every and (ep % log_every == 0 or ep == epochs - 1):
            msg = f"warm-start training epoch {ep}: loss {losses[-1]:.5f}"
            logging.info(msg)
            if progress is not None:
                progress(msg)
    return net, losses


def warm_control(sm, net, cg_iters=None, refine=None):
    """Initial control [T, N, 3] from the network's prediction.

    target 'vresid': the prediction is added, in each vertex's local frame, to
    the ray-cast velocity h n; the sum is re-ray-cast `refine` times
    (refine_velocity, default NET_REFINE_STEPS) and solved for a
    time-constant control with the network's solve_iters products
    (SOLVE_ITERS), as the feature-only control is; 'vlocal': the prediction is K(x0) abar in each vertex's local
    frame (normal, tangents), turned into abar by a conjugate-gradient solve
    truncated at the network's solve_iters products (SOLVE_ITERS) and used as a
    time-constant control; 'abar': the prediction is the time-mean of the plain optimiser's
    control, used as a time-constant control; 'vbar': the prediction is
    K(x0) abar, turned into abar by a conjugate-gradient solve (Kernel.solve,
    truncated at cg_iters products) and used as a time-constant control;
    'momentum': a0, shot with EPDiff; 'velocity': K(x0) a0, solved for a0 the
    same way and shot with EPDiff."""
    target = getattr(net, 'target', 'momentum')
    if cg_iters is None:
        cg_iters = getattr(net, 'solve_iters', None) or 20
    if target == 'vresid':
        ch, fr = ray_cast_heights(sm)
        X = height_inputs(sm, getattr(net, 'features', 'normal'), (ch, fr))
        pred = net.predict(X, [X.shape[0]])
        v = ch[:, :1] * fr[:, 0] + torch.einsum('nkj,nk->nj', fr, pred)
        v = refine_velocity(sm, v, NET_REFINE_STEPS if refine is None else refine)
        a = solve_control(sm, v, cg_iters)
        return time_constant(sm, a), a
    feats = features(sm, getattr(net, 'features', 'force'))
    pred = net.predict(feats, [feats.shape[0]])
    T = sm.Tsize
    if target == 'vlocal':
        v = torch.einsum('nkj,nk->nj', frame(sm), pred)
        a = sm.options['KparDiff'].solve(sm.x0_dev, v, iterMax=cg_iters, sequence='textbook')
        return a[None].expand(T, -1, -1).contiguous(), a
    if target == 'abar':
        return pred[None].expand(T, -1, -1).contiguous(), pred
    if target == 'vbar':
        a = sm.options['KparDiff'].solve(sm.x0_dev, pred, iterMax=cg_iters, sequence='textbook')
        return a[None].expand(T, -1, -1).contiguous(), a
    if target == 'velocity':
        a0 = sm.options['KparDiff'].solve(sm.x0_dev, pred, iterMax=cg_iters, sequence='textbook')
    else:
        a0 = pred
    _, at = evol.landmarkEPDiff(T, sm.x0_dev, a0, sm.options['KparDiff'])
    return at, a0


def register(seed, net=None, target_data=None, cfg=DEFAULT, max_iter=None, a0=None, cg_iters=None, at=None,
             start=None, refine=None):
    """Registration (plain, warm-started by `net`, from a given control `at`,
    from a given initial momentum `a0`, or from the control `start(sm)`
    returns for the new matching) until the data term reaches target_data
    (1e-4 rel).  The timed region covers set-up, the warm control and BFGS.

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
    if net is not None or a0 is not None or at is not None or start is not None:
        if start is not None:
            at = start(sm)
        elif at is not None:
            at = B.to_dev(at)
        elif a0 is None:
            at, _ = warm_control(sm, net, cg_iters=cg_iters, refine=refine)
        else:
            _, at = evol.landmarkEPDiff(sm.Tsize, sm.x0_dev, B.to_dev(a0), sm.options['KparDiff'])
        sm.setInitialControl(at)
        obj_warm = sm.objectiveFun()
        # the zero control leaves x_t = x_0 with no deformation energy; the
        # target's self term obj0 is the one objectiveFun just computed
        obj_zero = sm.obj0 + float(sm.dataTerm_device(sm.x0_dev))
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
            "data": d, "data0": data0, "reached": bool(reached), "used_warm": used_warm,
            "objective": float(sm.objectiveFun())}
