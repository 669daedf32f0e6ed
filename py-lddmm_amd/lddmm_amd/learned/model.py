"""MomentumNet: shared per-point MLP 3 -> 64 -> 128 -> 128 (ReLU), max-pool
over the points of each template-target pair (global feature g), per-point
head [h3, g] -> 128 (ReLU) -> 3.  BF16 weights (FP32 master copies for the
optimiser), FP32 activations and accumulation.  Forward, backward and the
Adam step run in this package's HIP kernels (lddmm_nn_* in liblddmm_amd.so);
torch supplies device buffers and the elementwise loss.

Parameter layout (float32, one flat vector): for each layer W [K, N] then
b [N], layers in the order of LAYERS; the head's first layer is split into
W4a (per-point part) and W4b (global part, no bias).
"""
import numpy as np
import torch

from .. import _backend as B

LAYERS = [  # name, K, N, has_bias
    ("l1", 3, 64, True),
    ("l2", 64, 128, True),
    ("l3", 128, 128, True),
    ("l4a", 128, 128, True),
    ("l4b", 128, 128, False),
    ("l5", 128, 3, True),
]


def layout():
    off, out = 0, {}
    for name, K, N, hb in LAYERS:
        out[name + ".W"] = (off, (K, N))
        off += K * N
        if hb:
            out[name + ".b"] = (off, (N,))
            off += N
    return out, off


def bf16_round(a):
    """Round float32 values to the nearest bf16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (r.astype(np.uint32) << 16).view(np.float32)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    lay, total = layout()
    w = np.zeros(total, dtype=np.float32)
    for name, K, N, hb in LAYERS:
        o, shp = lay[name + ".W"]
        fan_in = K if name != "l4b" else 256
        std = np.sqrt(2.0 / fan_in) if name != "l5" else np.sqrt(1.0 / K)
        w[o:o + K * N] = rng.normal(0, std, K * N)
    return w


class Segments:
    """Contiguous point ranges of the pairs in a batch (device int32 arrays)."""

    def __init__(self, counts, device):
        counts = np.asarray(counts, dtype=np.int64)
        off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        self.n = len(counts)
        self.P = int(off[-1])
        self.off = B.upload(off).to(device)
        self.seg = B.upload(np.repeat(np.arange(self.n, dtype=np.int32), counts)).to(device)
        self.all = B.upload(np.array([0, self.P], dtype=np.int32)).to(device)


class MomentumNet:
    def __init__(self, params=None, seed=0, feat_scale=1.0, out_scale=1.0, target='abar', features='force',
                 solve_iters=0):
        # target: 'vlocal' (K(x0) abar in the local vertex frames), 'abar' (time-mean of
        # the control, used as a time-constant control), 'vbar' (K(x0) abar),
        # 'momentum' (a0 = at[0], shot with EPDiff) or 'velocity' (K(x0) a0)
        self.target = target
        self.features = features      # input channels: 'normal', 'force', 'meanshift' or 'kforce'
        self.solve_iters = int(solve_iters)   # CG products of the velocity -> control solve (0: caller's)

        self.dev = B.device()
        self.lay, self.total = layout()
        p = init_params(seed) if params is None else np.asarray(params, dtype=np.float32)
        self.w = torch.from_numpy(p.copy()).to(self.dev)
        self.wbf = torch.empty(self.total, dtype=torch.int16, device=self.dev)
        B.check(B.LIB.lddmm_nn_to_bf16(B.ptr(self.w), B.ptr(self.wbf), self.total, B.stream_handle()), "nn_to_bf16")
        self.g = torch.zeros_like(self.w)
        self.m = torch.zeros_like(self.w)
        self.v = torch.zeros_like(self.w)
        self.step = 0
        self.feat_scale = float(feat_scale)
        self.out_scale = float(out_scale)

    # ---------------------------------------------------------------- views
    def _W(self, name):
        o, _ = self.lay[name + ".W"]
        return self.wbf[o:]

    def _b(self, name):
        key = name + ".b"
        if key not in self.lay:
            return None
        o, _ = self.lay[key]
        return self.w[o:]

    def _gW(self, name):
        o, (K, N) = self.lay[name + ".W"]
        return self.g[o:o + K * N]

    def _gb(self, name):
        o, (N,) = self.lay[name + ".b"]
        return self.g[o:o + N]

    # --------------------------------------------------------------- kernels
    def _lin(self, X, name, relu, segbias=None, seg=None, transW=False, bias=True):
        K, N = self.lay[name + ".W"][1]
        if transW:
            K, N = N, K
        P = X.shape[0]
        Y = torch.empty((P, N), dtype=torch.float32, device=self.dev)
        b = self._b(name) if (bias and not transW) else None
        B.check(B.LIB.lddmm_nn_linear(B.ptr(X), P, K, B.ptr(self._W(name)), 1 if transW else 0, B.ptr(b),
                                      B.ptr(segbias), B.ptr(seg), N, 1 if relu else 0, B.ptr(Y),
                                      B.stream_handle()), "nn_linear")
        return Y

    def _relu_bwd(self, Y, dY):
        dZ = torch.empty_like(dY)
        B.check(B.LIB.lddmm_nn_relu_bwd(B.ptr(Y), B.ptr(dY), B.ptr(dZ), dY.numel(), B.stream_handle()), "relu_bwd")
        return dZ

    def _wgrad(self, X, dZ, out):
        P, K = X.shape
        N = dZ.shape[1]
        B.check(B.LIB.lddmm_nn_wgrad(B.ptr(X), B.ptr(dZ), P, K, N, B.ptr(out), B.stream_handle()), "nn_wgrad")

    def _colsum(self, dZ, off, nseg, out):
        B.check(B.LIB.lddmm_nn_seg_colsum(B.ptr(dZ), B.ptr(off), nseg, dZ.shape[1], B.ptr(out), B.stream_handle()),
                "seg_colsum")

    # --------------------------------------------------------------- passes
    def forward(self, X, segs):
        """X: [P, 3] float32 device features; returns the [P, 3] float32 output (scaled units)."""
        X = X.contiguous()
        h1 = self._lin(X, "l1", True)
        h2 = self._lin(h1, "l2", True)
        h3 = self._lin(h2, "l3", True)
        g = torch.empty((segs.n, 128), dtype=torch.float32, device=self.dev)
        arg = torch.empty((segs.n, 128), dtype=torch.int32, device=self.dev)
        B.check(B.LIB.lddmm_nn_segmax(B.ptr(h3), B.ptr(segs.off), segs.n, 128, B.ptr(g), B.ptr(arg),
                                      B.stream_handle()), "segmax")
        sb = self._lin(g, "l4b", False, bias=False)
        h4 = self._lin(h3, "l4a", True, segbias=sb, seg=segs.seg)
        out = self._lin(h4, "l5", False)
        self._cache = (X, h1, h2, h3, g, arg, h4, segs)
        return out

    def backward(self, dout):
        X, h1, h2, h3, g, arg, h4, segs = self._cache
        dout = dout.contiguous()
        self._wgrad(h4, dout, self._gW("l5"))
        self._colsum(dout, segs.all, 1, self._gb("l5"))
        dh4 = self._lin(dout, "l5", False, transW=True)
        dz4 = self._relu_bwd(h4, dh4)
        self._wgrad(h3, dz4, self._gW("l4a"))
        self._colsum(dz4, segs.all, 1, self._gb("l4a"))
        dsb = torch.empty((segs.n, 128), dtype=torch.float32, device=self.dev)
        self._colsum(dz4, segs.off, segs.n, dsb)
        self._wgrad(g, dsb, self._gW("l4b"))
        dg = self._lin(dsb, "l4b", False, transW=True)
        dh3 = self._lin(dz4, "l4a", False, transW=True)
        B.check(B.LIB.lddmm_nn_segmax_bwd(B.ptr(dg), B.ptr(segs.off), B.ptr(arg), segs.n, 128, B.ptr(dh3),
                                          B.stream_handle()), "segmax_bwd")
        dz3 = self._relu_bwd(h3, dh3)
        self._wgrad(h2, dz3, self._gW("l3"))
        self._colsum(dz3, segs.all, 1, self._gb("l3"))
        dh2 = self._lin(dz3, "l3", False, transW=True)
        dz2 = self._relu_bwd(h2, dh2)
        self._wgrad(h1, dz2, self._gW("l2"))
        self._colsum(dz2, segs.all, 1, self._gb("l2"))
        dh1 = self._lin(dz2, "l2", False, transW=True)
        dz1 = self._relu_bwd(h1, dh1)
        self._wgrad(X, dz1, self._gW("l1"))
        self._colsum(dz1, segs.all, 1, self._gb("l1"))
        return self.g

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        self.step += 1
        B.check(B.LIB.lddmm_nn_adam(B.ptr(self.w), B.ptr(self.wbf), B.ptr(self.g), B.ptr(self.m), B.ptr(self.v),
                                    self.total, lr, b1, b2, eps, self.step, wd, B.stream_handle()), "nn_adam")

    def loss_and_grad(self, X, segs, target):
        out = self.forward(X, segs)
        diff = out - target
        n = diff.numel()
        loss = (diff * diff).sum() / n
        self.backward(2.0 * diff / n)
        return loss

    # ---------------------------------------------------------------- use
    def predict(self, feats, counts):
        """feats: [P, 3] (float64/32, device or numpy) in physical units -> a0 [P, 3] float64."""
        X = B.to_dev(feats, torch.float32) / self.feat_scale
        segs = Segments(counts, self.dev)
        out = self.forward(X, segs)
        return out.double() * self.out_scale

    def params_numpy(self):
        return self.w.cpu().numpy().copy()

    def save(self, path):
        np.savez(path, params=self.params_numpy(), feat_scale=self.feat_scale, out_scale=self.out_scale,
                 target=np.array(self.target), features=np.array(self.features),
                 solve_iters=np.array(self.solve_iters))

    @classmethod
    def load(cls, path):
        d = np.load(path)
        target = str(d["target"]) if "target" in d.files else 'momentum'
        features = str(d["features"]) if "features" in d.files else 'force'
        solve_iters = int(d["solve_iters"]) if "solve_iters" in d.files else 0
        return cls(params=d["params"], feat_scale=float(d["feat_scale"]), out_scale=float(d["out_scale"]),
                   target=target, features=features, solve_iters=solve_iters)
