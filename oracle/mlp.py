"""Float64 numpy forward / backward of MomentumNet (test infrastructure only).

Mirrors lddmm_amd/learned/model.py layer by layer with the BF16-rounded
weights the GPU uses, so that the GPU gradients can be checked against it
(and the numpy gradient itself is checked by finite differences in tests).
"""
import numpy as np

LAYERS = [("l1", 3, 64, True), ("l2", 64, 128, True), ("l3", 128, 128, True),
          ("l4a", 128, 128, True), ("l4b", 128, 128, False), ("l5", 128, 3, True)]


def unpack(w):
    out, off = {}, 0
    for name, K, N, hb in LAYERS:
        out[name + ".W"] = w[off:off + K * N].reshape(K, N)
        off += K * N
        if hb:
            out[name + ".b"] = w[off:off + N]
            off += N
    return out


def forward(wW, wb, X, counts):
    """wW: params with W entries as used in products (bf16-rounded), wb: params for biases.
    Returns (out, cache)."""
    P = unpack(np.asarray(wW, dtype=np.float64))
    Q = unpack(np.asarray(wb, dtype=np.float64))
    relu = lambda z: np.maximum(z, 0)
    h1 = relu(X @ P["l1.W"] + Q["l1.b"])
    h2 = relu(h1 @ P["l2.W"] + Q["l2.b"])
    h3 = relu(h2 @ P["l3.W"] + Q["l3.b"])
    off = np.concatenate([[0], np.cumsum(counts)])
    g = np.stack([h3[off[s]:off[s + 1]].max(0) for s in range(len(counts))])
    arg = np.stack([off[s] + h3[off[s]:off[s + 1]].argmax(0) for s in range(len(counts))])
    seg = np.repeat(np.arange(len(counts)), counts)
    sb = g @ P["l4b.W"]
    h4 = relu(h3 @ P["l4a.W"] + Q["l4a.b"] + sb[seg])
    out = h4 @ P["l5.W"] + Q["l5.b"]
    return out, (X, h1, h2, h3, g, arg, seg, h4, P, off)


def backward(cache, dout):
    X, h1, h2, h3, g, arg, seg, h4, P, off = cache
    G = {}
    G["l5.W"] = h4.T @ dout
    G["l5.b"] = dout.sum(0)
    dz4 = (dout @ P["l5.W"].T) * (h4 > 0)
    G["l4a.W"] = h3.T @ dz4
    G["l4a.b"] = dz4.sum(0)
    dsb = np.stack([dz4[off[s]:off[s + 1]].sum(0) for s in range(len(off) - 1)])
    G["l4b.W"] = g.T @ dsb
    dg = dsb @ P["l4b.W"].T
    dh3 = dz4 @ P["l4a.W"].T
    for s in range(len(off) - 1):
        dh3[arg[s], np.arange(h3.shape[1])] += dg[s]
    dz3 = dh3 * (h3 > 0)
    G["l3.W"] = h2.T @ dz3
    G["l3.b"] = dz3.sum(0)
    dz2 = (dz3 @ P["l3.W"].T) * (h2 > 0)
    G["l2.W"] = h1.T @ dz2
    G["l2.b"] = dz2.sum(0)
    dz1 = (dz2 @ P["l2.W"].T) * (h1 > 0)
    G["l1.W"] = X.T @ dz1
    G["l1.b"] = dz1.sum(0)
    flat = []
    for name, K, N, hb in LAYERS:
        flat.append(G[name + ".W"].ravel())
        if hb:
            flat.append(G[name + ".b"])
    return np.concatenate(flat)


def loss_and_grad(wW, wb, X, counts, target):
    out, cache = forward(wW, wb, X, counts)
    d = out - target
    n = d.size
    return (d * d).sum() / n, backward(cache, 2 * d / n)
