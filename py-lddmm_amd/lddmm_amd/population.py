"""Population registration: many template-target surface pairs at once.

`Population(templates, targets)` concatenates the pairs on the device and
exposes the batched kernels of liblddmm_amd (lddmm_population_*): shooting,
adjoint and varifold data term of every active pair, each Gaussian sum one
persistent launch over a segment table of (pair, tile, chunk) items.

`PopulationMatching(templates, targets, options).optimizeMatching()` runs
the plain SurfaceMatching registration (py-lddmm base/surfaceMatching.py
with algorithm 'bfgs', errorType 'varifold') of every pair in lockstep.  The
per-pair iteration is that of bfgs.bfgs + linesearch.line_search_weak_wolfe
(py-lddmm base/bfgs.py:55-312, base/linesearch.py:66-163) and of
SurfaceMatching's precision switching (startOfIteration / endOfIteration):
each pair keeps its own step sizes, line-search state, L-BFGS memory and
stopping decisions; only the arithmetic is batched.  The L-BFGS two-loop
runs over all pairs with per-pair dot products (lddmm_population_seg_dot);
the line searches are per-pair state machines whose objective and gradient
requests are gathered into one batched evaluation per round.  The reference
has no population driver; per pair, results equal SurfaceMatching's up to
floating-point summation order.
"""
import ctypes
import logging
import time

import numpy as np
import torch

from . import _backend as B
from . import surfaceDistances as sd
from .kernelFunctions import Kernel
from .surfaces import Surface



def closed_and_volumes(surfs):
    """Per surface: closed (no edge belongs to exactly one face, the test of
    Surface.getEdges' boundary flags) and the signed volume sum_f det(v[f]) / 6
    (Surface.surfVolume), for a list of surfaces in one device pass: the
    edges of all surfaces, keyed by global vertex ids, are sorted together.
    Returns (closed [S] bool, volumes [S] float64) on the host."""
    S = len(surfs)
    nv = np.array([s.vertices.shape[0] for s in surfs], dtype=np.int64)
    nf = np.array([s.faces.shape[0] for s in surfs], dtype=np.int64)
    voff = np.concatenate([[0], np.cumsum(nv)])
    foff = np.concatenate([[0], np.cumsum(nf)])
    dev = B.device()
    F = B.upload(np.concatenate([np.asarray(s.faces, dtype=np.int64) + o for s, o in zip(surfs, voff[:-1])]))
    X = B.to_dev(np.concatenate([np.asarray(s.vertices, dtype=np.float64) for s in surfs]))
    a = torch.cat([F[:, 0], F[:, 1], F[:, 2]])
    b = torch.cat([F[:, 1], F[:, 2], F[:, 0]])
    keys, _ = torch.sort(torch.minimum(a, b) * int(voff[-1]) + torch.maximum(a, b))
    new = torch.ones_like(keys, dtype=torch.bool)
    new[1:] = keys[1:] != keys[:-1]
    start = torch.nonzero(new).flatten()
    cnt = torch.diff(torch.cat([start, torch.tensor([keys.numel()], device=dev)]))
    lone = keys[start[cnt == 1]] // int(voff[-1])       # smaller vertex of every one-face edge
    owner = torch.searchsorted(B.upload(voff[1:]), lone, right=True)
    open_ = torch.zeros(S, dtype=torch.bool, device=dev)
    open_[owner] = True
    det = (X[F[:, 0]] * torch.cross(X[F[:, 1]], X[F[:, 2]], dim=1)).sum(1)
    det = det.cpu().numpy()
    vol = np.array([det[foff[i]:foff[i + 1]].sum() / 6 for i in range(S)])
    return ~open_.cpu().numpy(), vol


def to_device_async(a, device):
    """Host numpy array -> device tensor without waiting for the GPU: the
    values go through page-locked memory from torch's caching host allocator
    and are copied on the current stream (a pageable copy would block until
    every queued kernel has finished)."""
    t = torch.from_numpy(np.ascontiguousarray(a)).pin_memory()
    return t.to(device, non_blocking=True)

class Population:
    """Device layout of P template-target surface pairs (vertices [V, 3] of all
    templates, pair p owning rows voff[p]:voff[p+1])."""

    def __init__(self, templates, targets):
        if len(templates) != len(targets) or not templates:
            raise ValueError("Population needs equally many (>= 1) templates and targets")
        self.P = len(templates)
        tv = [np.asarray(s.vertices, dtype=np.float64) for s in templates]
        tf = [np.asarray(s.faces, dtype=np.int64) for s in templates]
        gv = [np.asarray(s.vertices, dtype=np.float64) for s in targets]
        gf = [np.asarray(s.faces, dtype=np.int64) for s in targets]
        self.voff = np.concatenate([[0], np.cumsum([v.shape[0] for v in tv])]).astype(np.int64)
        self.foff = np.concatenate([[0], np.cumsum([f.shape[0] for f in tf])]).astype(np.int64)
        tvoff = np.concatenate([[0], np.cumsum([v.shape[0] for v in gv])]).astype(np.int64)
        self.tfoff = np.concatenate([[0], np.cumsum([f.shape[0] for f in gf])]).astype(np.int64)
        self.V = int(self.voff[-1])
        faces = np.concatenate([f + o for f, o in zip(tf, self.voff[:-1])]).astype(np.int32).reshape(-1, 3)
        tfaces = np.concatenate([f + o for f, o in zip(gf, tvoff[:-1])]).astype(np.int32).reshape(-1, 3)
        ptr = np.zeros(self.V + 1, dtype=np.int32)
        idx = np.zeros(max(3 * faces.shape[0], 1), dtype=np.int32)
        B.check(B.LIB.lddmm_surface_adjacency(faces.ctypes.data, faces.shape[0], self.V, ptr.ctypes.data,
                                              idx.ctypes.data), "surface_adjacency")
        dev = B.device()
        w = np.concatenate([np.asarray(s.weights, dtype=np.float64) for s in templates])
        tw = np.concatenate([np.asarray(s.weights, dtype=np.float64) for s in targets])
        self._keep = {
            "faces": B.upload(faces), "ptr": B.upload(ptr),
            "idx": B.upload(idx), "tvert": B.to_dev(np.concatenate(gv)),
            "tfaces": B.upload(tfaces),
            "w": None if np.all(w == 1) else B.to_dev(w), "tw": None if np.all(tw == 1) else B.to_dev(tw),
        }
        self.x0 = B.to_dev(np.concatenate(tv))
        self.vpair = B.upload(np.repeat(np.arange(self.P), np.diff(self.voff)))
        k = self._keep
        d = B.LddmmPopulationDesc()
        d.npairs = self.P
        d.voff = self.voff.ctypes.data
        d.foff = self.foff.ctypes.data
        d.faces = k["faces"].data_ptr()
        d.vweights = None if k["w"] is None else k["w"].data_ptr()
        d.adj_ptr = k["ptr"].data_ptr()
        d.adj_idx = k["idx"].data_ptr()
        d.tfoff = self.tfoff.ctypes.data
        d.tvertices = k["tvert"].data_ptr()
        d.tfaces = k["tfaces"].data_ptr()
        d.tvweights = None if k["tw"] is None else k["tw"].data_ptr()
        h = ctypes.c_void_p()
        B.check(B.LIB.lddmm_population_create(ctypes.byref(d), ctypes.byref(h)), "population_create")
        self.h = h
        self.active = np.ones(self.P, dtype=bool)

    def close(self):
        if getattr(self, "h", None):
            B.LIB.lddmm_population_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- calls
    def set_active(self, mask=None):
        m = np.ones(self.P, dtype=np.uint8) if mask is None else np.asarray(mask, dtype=np.uint8)
        self.active = m.astype(bool)
        B.check(B.LIB.lddmm_population_set_active(self.h, m.ctypes.data), "population_set_active")

    def set_graphs(self, on):
        B.check(B.LIB.lddmm_population_set_graphs(self.h, 1 if on else 0), "population_set_graphs")

    def shoot(self, kd, at, xt, regweight=1.0, energies=None):
        """xt[1:] of the active pairs from xt[0] and at; energies [P] (device) if given."""
        T = at.shape[0]
        B.check(B.LIB.lddmm_population_shoot_forward(kd, self.h, B.ptr(at), T, B.ptr(xt), float(regweight),
                                                     B.ptr(energies), B.stream_handle()), "population_shoot_forward")

    def adjoint(self, kd, at, xt, px1, grad, regweight=1.0, euclidean=True):
        T = at.shape[0]
        B.check(B.LIB.lddmm_population_shoot_adjoint(kd, self.h, B.ptr(at), B.ptr(xt), T, B.ptr(px1),
                                                     float(regweight), 1 if euclidean else 0, B.ptr(grad),
                                                     B.stream_handle()), "population_shoot_adjoint")

    def varifold(self, kd, vertices, energies=None, grad=None):
        B.check(B.LIB.lddmm_population_varifold(kd, self.h, B.ptr(vertices), B.ptr(energies), B.ptr(grad),
                                                B.stream_handle()), "population_varifold")

    def seg_dot(self, a, b, out=None):
        """[P] device tensor: per-pair sum of a*b over [T, V, 3] (active pairs)."""
        if out is None:
            out = torch.zeros(self.P, dtype=torch.float64, device=a.device)
        T = a.shape[0] if a.dim() == 3 else 1
        B.check(B.LIB.lddmm_population_seg_dot(self.h, B.ptr(a), B.ptr(b), T, B.ptr(out), B.stream_handle()),
                "population_seg_dot")
        return out

    def seg_axpy(self, y, alpha, x):
        """y += alpha[pair] * x on the rows of active pairs (alpha: [P] device float64)."""
        T = y.shape[0] if y.dim() == 3 else 1
        B.check(B.LIB.lddmm_population_seg_axpy(self.h, B.ptr(y), B.ptr(alpha), B.ptr(x), T, B.stream_handle()),
                "population_seg_axpy")

    def tuning(self):
        buf = (ctypes.c_int32 * 64)()
        n = ctypes.c_int32()
        B.check(B.LIB.lddmm_population_tuning(self.h, buf, 16, ctypes.byref(n)), "population_tuning")
        names = {0: "applyK", 1: "applyDiffKT", 2: "varifold_energy", 3: "varifold_gradient", 4: "adjoint_step"}
        return [{"op": names.get(buf[4 * i], buf[4 * i]), "bytes": buf[4 * i + 1], "R": buf[4 * i + 2],
                 "items_per_workgroup": buf[4 * i + 3] / 10} for i in range(min(n.value, 16))]

    def rows(self, mask):
        """Boolean device mask [V] of the rows of the pairs selected by mask [P]."""
        return to_device_async(np.asarray(mask, dtype=bool), self.x0.device)[self.vpair]


# ---------------------------------------------------------------------------
# Lockstep population registration
# ---------------------------------------------------------------------------

def _line_search(t_init, derphi0, old_fval, c1=1e-4, c2=0.9, gradCoeff=1.0, maxiter=10):
    """Weak-Wolfe search of one pair as a generator (linesearch.line_search_weak_wolfe):
    yields ('phi', t) -> objective at t, ('derphi', t) -> -<grad(t), dir>;
    returns (t or None, fval, gval_valid)."""
    it = 0
    itGrad = 0
    amax = 0
    amin = 0
    maxiter2 = 100
    t = t_init
    gval = False
    fval = yield ('phi', t)
    while it < maxiter2 and itGrad < maxiter:
        armijo = False
        if fval < old_fval + c1 * t * derphi0 * gradCoeff:
            df = yield ('derphi', t)
            gval = True
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
        fval = yield ('phi', t)
        it += 1
    if itGrad == maxiter:
        gval = False
        while it < maxiter2:
            if fval < old_fval + c1 * t * derphi0 * gradCoeff:
                yield ('derphi', t)
                gval = True
                break
            else:
                t /= 2
            fval = yield ('phi', t)
            it += 1
    if it == maxiter2:
        t = None
    return t, fval, gval


class PopulationMatching:
    """Plain varifold registration of every template-target pair, in lockstep."""

    def __init__(self, templates, targets, options=None):
        o = {'timeStep': 0.1, 'KparDiff': None, 'KparDist': None, 'sigmaError': 1.0, 'errorType': 'varifold',
             'pk_dtype': 'float32', 'maxIter': 1000, 'regWeight': 1.0, 'burnIn': 10, 'epsInit': 1.,
             'Wolfe': True, 'memory': 50, 'gradTol': -1., 'graphs': True}
        if options:
            o.update(options)
        if o['errorType'] != 'varifold':
            raise NotImplementedError("PopulationMatching builds the varifold data term only")
        self.options = o
        self.fv0 = [Surface(surf=s) for s in templates]
        self.fv1 = [Surface(surf=s) for s in targets]

        # orientation as SurfaceMatching.fix_orientation (base/surfaceMatching.py:284-318):
        # when template and target are both closed and their signed volumes
        # disagree, the target's faces are flipped; all pairs in one device pass
        closed, vol = closed_and_volumes(self.fv0 + self.fv1)
        P_ = len(self.fv0)
        for p in range(P_):
            if closed[p] and closed[P_ + p] and vol[p] * vol[P_ + p] < 0:
                self.fv1[p].flipFaces()
        self.pop = Population(self.fv0, self.fv1)
        self.pop.set_graphs(o['graphs'])
        self.P = self.pop.P
        self.Tsize = int(round(1.0 / o['timeStep']))
        kd_ = o['KparDiff'] if isinstance(o['KparDiff'], Kernel) else Kernel(name='gauss', sigma=6.5)
        kv_ = o['KparDist'] if isinstance(o['KparDist'], Kernel) else Kernel(name='gauss', sigma=2.5)
        self.Kdiff, self.Kdist = kd_, kv_
        self.dtype0 = o['pk_dtype']
        self.dtype = [self.dtype0] * self.P
        self.pkBuffer = np.zeros(self.P, dtype=int)
        dev = B.device()
        V = self.pop.V
        T = self.Tsize
        z = lambda: torch.zeros((T, V, 3), dtype=torch.float64, device=dev)
        self.A = z()                  # control
        self.CT = z()                 # trial control
        self.XT = self.pop.x0.expand(T + 1, -1, -1).clone()
        self.XTT = self.XT.clone()
        self.G = z()                  # gradient at the control
        self.GT = z()                 # gradient at the trial control
        self.GOLD = z()
        self.DIR = z()
        self.DIFF = z()
        self.PX1 = torch.zeros((V, 3), dtype=torch.float64, device=dev)
        self.EN = torch.zeros(self.P, dtype=torch.float64, device=dev)
        self.ED = torch.zeros(self.P, dtype=torch.float64, device=dev)
        # |target|^2 / sigmaError^2 per precision; like SurfaceMatching, a pair
        # uses the value of its precision at the last full objective evaluation.
        # The float64 values of a float32 run are computed when a pair first
        # switches precision (obj0_of).
        self._obj0_by = {}
        self.obj0 = self.obj0_of(self.dtype0).copy()
        self.obj = np.full(self.P, np.nan)
        self.stats = {"evaluations": 0, "gradients": 0, "rounds": 0}

    def obj0_of(self, dt):
        """varifoldNorm0 / sigmaError^2 of every target at precision dt (cached)."""
        if dt not in self._obj0_by:
            kv_ = self.Kdist
            k = Kernel(name=kv_.name, sigma=kv_.sigma, order=kv_.order)
            k.ms_exponent = kv_.ms_exponent
            k.pk_dtype = dt
            kd = sd._kd(k)
            # varifoldNorm0 of every target, read back in one transfer
            e = torch.cat([sd.varifold_device(kd, t)[0] for t in self.fv1])
            self._obj0_by[dt] = e.cpu().numpy() / self.options['sigmaError'] ** 2
        return self._obj0_by[dt]

    # ---------------------------------------------------------- evaluations
    def _kd(self, K, dtype):
        return B.kernel_of(K, dtype)

    def _groups(self, mask):
        """Active masks per precision for the pairs in mask."""
        out = []
        for dt in ('float32', 'float64'):
            m = np.array([mask[p] and self.dtype[p] == dt for p in range(self.P)])
            if m.any():
                out.append((dt, m))
        return out

    def _objective(self, mask, control, state, full=False):
        """Objective of the pairs in mask at `control`; fills `state` rows. -> [P] host (nan elsewhere).
        full=True is SurfaceMatching.objectiveFun with obj None: obj0 is refreshed
        at the pair's current precision; otherwise (updateTry) the stored obj0 is used."""
        out = np.full(self.P, np.nan)
        if full:
            for p in np.nonzero(mask)[0]:
                self.obj0[p] = self.obj0_of(self.dtype[p])[p]
        s2 = self.options['sigmaError'] ** 2
        for dt, m in self._groups(mask):
            self.pop.set_active(m)
            self.pop.shoot(self._kd(self.Kdiff, dt), control, state, self.options['regWeight'], self.EN)
            self.pop.varifold(self._kd(self.Kdist, dt), state[-1], energies=self.ED)
            en = self.EN.cpu().numpy()
            ed = self.ED.cpu().numpy()
            v = self.obj0 + en + ed / s2
            out[m] = v[m]
            self.stats["evaluations"] += int(m.sum())
        out[np.isnan(out) & np.asarray(mask, dtype=bool)] = 1e500
        return out

    def _objective_and_gradient(self, mask, control, state, grad):
        """_objective (updateTry form) of the pairs in mask, with their gradients
        written to `grad` as _gradient would: the varifold gradient pass also
        yields the energies, so a trial costs one varifold pass (the line
        search asks for the slope at the same step whenever the trial passes
        the sufficient-decrease test; SurfaceMatching.updateTryWithSlope does
        the same per pair).  -> [P] host objectives (nan elsewhere)."""
        out = np.full(self.P, np.nan)
        s2 = self.options['sigmaError'] ** 2
        for dt, m in self._groups(mask):
            self.pop.set_active(m)
            kdd = self._kd(self.Kdiff, dt)
            self.pop.shoot(kdd, control, state, self.options['regWeight'], self.EN)
            self.pop.varifold(self._kd(self.Kdist, dt), state[-1], energies=self.ED, grad=self.PX1)
            rows = self.pop.rows(m)
            torch.where(rows[:, None], self.PX1 / -s2, self.PX1, out=self.PX1)
            self.pop.adjoint(kdd, control, state, self.PX1, grad, self.options['regWeight'], euclidean=True)
            en = self.EN.cpu().numpy()
            ed = self.ED.cpu().numpy()
            v = self.obj0 + en + ed / s2
            out[m] = v[m]
            self.stats["evaluations"] += int(m.sum())
            self.stats["gradients"] += int(m.sum())
        out[np.isnan(out) & np.asarray(mask, dtype=bool)] = 1e500
        return out

    def _gradient(self, mask, control, state, grad):
        """Euclidean gradient wrt the control (SurfaceMatching.getGradient) of the pairs in mask."""
        s2 = self.options['sigmaError'] ** 2
        for dt, m in self._groups(mask):
            self.pop.set_active(m)
            self.pop.varifold(self._kd(self.Kdist, dt), state[-1], grad=self.PX1)
            rows = self.pop.rows(m)
            torch.where(rows[:, None], self.PX1 / -s2, self.PX1, out=self.PX1)
            self.pop.adjoint(self._kd(self.Kdiff, dt), control, state, self.PX1, grad,
                             self.options['regWeight'], euclidean=True)
            self.stats["gradients"] += int(m.sum())

    def _dot(self, a, b, mask):
        self.pop.set_active(mask)
        return self.pop.seg_dot(a, b)

    # ------------------------------------------------------------ precision
    def _start_of_iteration(self, p):
        if self.reset[p]:
            self.dtype[p] = 'float64'
            self.pkBuffer[p] = 0

    def _end_of_iteration(self, p):
        if self.reset[p]:
            # the reference recomputes the objective here for logging; the next
            # iteration recomputes it again, so only the precision switch matters
            self.dtype[p] = self.dtype0
            self.pkBuffer[p] = 0
        if self.pkBuffer[p] > 50:
            self.dtype[p] = self.dtype0
            self.pkBuffer[p] = 0
        else:
            self.pkBuffer[p] += 1

    # --------------------------------------------------------------- driver
    def setInitialControl(self, at):
        at = B.to_dev(at)
        if tuple(at.shape) != tuple(self.A.shape):
            raise ValueError(f"control shape {tuple(at.shape)} != {tuple(self.A.shape)}")
        self.A.copy_(at)

    def optimizeMatching(self):
        o = self.options
        P = self.P
        allp = np.ones(P, dtype=bool)
        t0 = time.perf_counter()
        # SurfaceMatching.optimizeMatching: gradient at the initial control, gradEps
        self.obj = self._objective(allp, self.A, self.XT, full=True)
        self._gradient(allp, self.A, self.XT, self.G)
        g2 = self._dot(self.G, self.G, allp).cpu().numpy()
        gradEps = np.maximum(1e-5, np.sqrt(g2) / 10000) if o['gradTol'] < 0 else np.full(P, o['gradTol'])
        maxIter, burnIn, memory = o['maxIter'], o['burnIn'], o['memory']
        # bfgs.bfgs state per pair
        self.reset = np.zeros(P, dtype=bool)
        it = np.zeros(P, dtype=int)
        obj_old = np.full(P, np.nan)
        gval = np.zeros(P, dtype=bool)
        done = np.zeros(P, dtype=bool)
        self.converged = np.zeros(P, dtype=bool)
        self.history = [[self.obj[p]] for p in range(P)]
        slots = []          # L-BFGS memory: dicts S, Y, rho (device [P]), valid (host [P])
        first = True
        while True:
            live = ~done & (it < maxIter)
            if not live.any():
                break
            self.stats["rounds"] += 1
            for p in np.nonzero(live)[0]:
                self._start_of_iteration(p)
            rs = live & self.reset
            if rs.any():
                self.obj[rs] = self._objective(rs, self.A, self.XT, full=True)[rs]
                obj_old[rs] = np.nan
                gval[rs] = False
            # grd = gval or a fresh gradient (the initial one is reused on the first round)
            need = live & ~gval
            if first:
                need &= False
                first = False
            if need.any():
                self._gradient(need, self.A, self.XT, self.G)
            use = live & gval
            if use.any():
                rows = self.pop.rows(use)
                torch.where(rows[None, :, None], self.GT, self.G, out=self.G)
            # L-BFGS two-loop for pairs with memory
            app = live & ~self.reset & (it > 0)
            clear = live & ~app
            for s in slots:
                s["valid"][clear] = False
            if app.any():
                Y = self.G - self.GOLD
                S = self.DIFF.clone()
                sy = self._dot(Y, S, app)
                rho = torch.where(to_device_async(app, sy.device), 1.0 / sy, torch.zeros_like(sy))
                slots.append({"S": S, "Y": Y, "rho": rho, "valid": app.copy()})
                # at most `memory` entries per pair
                cnt = np.zeros(P, dtype=int)
                for s in reversed(slots):
                    cnt += s["valid"]
                    s["valid"][cnt > memory] = False
                slots = [s for s in slots if s["valid"].any()]
            q = self.G.clone()
            if app.any():
                alphas = []
                for s in reversed(slots):
                    m = s["valid"] & app
                    dq = self._dot(s["S"], q, m)
                    alpha = torch.where(to_device_async(m, dq.device), s["rho"] * dq, torch.zeros_like(dq))
                    alphas.append(alpha)
                    self.pop.set_active(m)
                    self.pop.seg_axpy(q, -alpha, s["Y"])
                alphas.reverse()
                last = slots[-1]
                ys = self._dot(last["S"], last["Y"], app)
                yy = self._dot(last["Y"], last["Y"], app)
                c = (ys / yy).cpu().numpy()
                c = np.where(c < 1e-10, 1.0, c)
                cv = to_device_async(np.where(app, c, 1.0), q.device)
                q *= cv[self.pop.vpair][None, :, None]
                for s, alpha in zip(slots, alphas):
                    m = s["valid"] & app
                    dy = self._dot(s["Y"], q, m)
                    beta = s["rho"] * dy
                    coef = torch.where(to_device_async(m, dy.device), alpha - beta, torch.zeros_like(dy))
                    self.pop.set_active(m)
                    self.pop.seg_axpy(q, coef, s["S"])
            grd2 = self._dot(self.G, self.G, live).cpu().numpy()
            qq = self._dot(q, q, live).cpu().numpy()
            grdTry = np.sqrt(np.maximum(1e-20, qq))
            lrows = self.pop.rows(live)
            torch.where(lrows[None, :, None], q, self.DIR, out=self.DIR)
            torch.where(lrows[None, :, None], self.G, self.GOLD, out=self.GOLD)
            t_init = np.where((it == 0) | self.reset, 1.0 / grdTry, 1.0)
            upd = live & ((it == 0) | (it == burnIn))
            gradEps[upd] = np.minimum(gradEps[upd], 0.0001 * np.sqrt(grd2[upd]))
            # line searches
            derphi0 = -self._dot(self.G, self.DIR, live).cpu().numpy()
            gens, reqs = {}, {}
            res = {}
            gt_step = np.full(P, np.nan)      # step at which GT holds each pair's trial gradient
            for p in np.nonzero(live)[0]:
                g = _line_search(t_init[p], derphi0[p], self.obj[p], c1=1e-4, c2=0.9, maxiter=10)
                gens[p] = g
                reqs[p] = next(g)
            while reqs:
                phi = np.zeros(P, dtype=bool)
                der = np.zeros(P, dtype=bool)
                tv = np.zeros(P)
                for p, (kind, t) in reqs.items():
                    tv[p] = t
                    (phi if kind == 'phi' else der)[p] = True
                vals = {}
                if phi.any():
                    rows = self.pop.rows(phi)
                    trow = to_device_async(tv, self.A.device)[self.pop.vpair]
                    torch.where(rows[None, :, None], self.A - trow[None, :, None] * self.DIR, self.CT, out=self.CT)
                    f = self._objective_and_gradient(phi, self.CT, self.XTT, self.GT)
                    gt_step[phi] = tv[phi]
                    for p in np.nonzero(phi)[0]:
                        vals[p] = f[p]
                if der.any():
                    stale = der & ~(gt_step == tv)
                    if stale.any():
                        self._gradient(stale, self.CT, self.XTT, self.GT)
                        gt_step[stale] = tv[stale]
                    d = -self._dot(self.GT, self.DIR, der).cpu().numpy()
                    for p in np.nonzero(der)[0]:
                        vals[p] = d[p]
                nreq = {}
                for p in list(reqs.keys()):
                    try:
                        nreq[p] = gens[p].send(vals[p])
                    except StopIteration as e:
                        res[p] = e.value
                reqs = nreq
            # accept / reset / stop, as bfgs.bfgs
            acc = np.zeros(P, dtype=bool)
            eps = np.zeros(P)
            newreset = np.zeros(P, dtype=bool)
            for p in np.nonzero(live)[0]:
                t, fval, gv = res[p]
                if t is not None:
                    acc[p] = True
                    eps[p] = t
                    obj_old[p] = self.obj[p]
                    self.obj[p] = fval
                    gval[p] = gv
                else:
                    gval[p] = False
                    if self.reset[p]:
                        done[p] = True
                        continue
                    newreset[p] = True
            if acc.any():
                rows = self.pop.rows(acc)
                erow = to_device_async(eps, self.A.device)[self.pop.vpair]
                rows3 = rows[None, :, None]
                torch.where(rows3, -erow[None, :, None] * self.DIR, self.DIFF, out=self.DIFF)
                torch.where(rows3, self.CT, self.A, out=self.A)
                torch.where(rows3, self.XTT, self.XT, out=self.XT)
            for p in np.nonzero(live & ~done)[0]:
                if not np.isnan(obj_old[p]) and abs(self.obj[p] - obj_old[p]) < 1e-7:
                    if it[p] > burnIn or self.reset[p]:
                        self.converged[p] = True
                        self.history[p].append(self.obj[p])
                        done[p] = True
                        continue
                    newreset[p] = True
                if np.sqrt(grd2[p]) < gradEps[p]:
                    self.converged[p] = True
                    self.history[p].append(self.obj[p])
                    done[p] = True
                    continue
                self.reset[p] = newreset[p]
                self._end_of_iteration(p)
                if not self.reset[p]:
                    it[p] += 1
                    self.history[p].append(self.obj[p])
        torch.cuda.synchronize()
        self.iterations = it
        self.seconds = time.perf_counter() - t0
        self.pop.set_active(None)
        return self.A

    def varifold_data(self):
        """Per-pair data term |def - tgt|^2 / sigmaError^2 at the current control (host [P])."""
        allp = np.ones(self.P, dtype=bool)
        out = np.zeros(self.P)
        for dt, m in self._groups(allp):
            self.pop.set_active(m)
            self.pop.varifold(self._kd(self.Kdist, dt), self.XT[-1], energies=self.ED)
            out[m] = (self.obj0 + self.ED.cpu().numpy() / self.options['sigmaError'] ** 2)[m]
        self.pop.set_active(None)
        return out

    def control_of(self, p):
        r0, r1 = int(self.pop.voff[p]), int(self.pop.voff[p + 1])
        return self.A[:, r0:r1]
