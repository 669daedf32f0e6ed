"""GPU checks of the learned warm start: the HIP forward / backward of
MomentumNet against the float64 numpy network (same BF16-rounded weights),
and one Adam step against its closed form."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lddmm_amd.learned import model
    return model, torch


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


@pytest.mark.parametrize("counts", [[1], [37, 64, 130], [2562, 2562]])
def test_forward_backward_match_oracle(net_mod, counts):
    model, torch = net_mod
    from oracle import mlp
    rng = np.random.default_rng(sum(counts))
    w = model.init_params(5)
    lay, _ = model.layout()
    for k, (o, shp) in lay.items():
        if k.endswith(".b"):
            w[o:o + shp[0]] = rng.normal(0, 0.1, shp[0])
    net = model.MomentumNet(params=w)
    P = sum(counts)
    X = rng.normal(size=(P, 3)).astype(np.float32)
    T = rng.normal(size=(P, 3)).astype(np.float32)
    segs = model.Segments(counts, net.dev)
    Xd, Td = torch.tensor(X, device="cuda"), torch.tensor(T, device="cuda")
    loss = net.loss_and_grad(Xd, segs, Td)
    out = net.forward(Xd, segs).cpu().numpy()
    g = net.g.cpu().numpy()
    wbf = model.bf16_round(w)                     # weights as the kernels read them
    out_r, _ = mlp.forward(wbf, w, X.astype(np.float64), counts)
    L_r, g_r = mlp.loss_and_grad(wbf, w, X.astype(np.float64), counts, T.astype(np.float64))
    assert rel(out, out_r) < 1e-5
    assert abs(float(loss) - L_r) / L_r < 1e-5
    assert rel(g, g_r) < 1e-4


@pytest.mark.parametrize("N", [3, 128, 300])
def test_segment_reductions_ragged(net_mod, N):
    """Per-segment column sums and max / first argmax over ragged segments
    (an empty one, one shorter than the chunk count, long ones), against numpy."""
    model, torch = net_mod
    from lddmm_amd import _backend as B
    rng = np.random.default_rng(N)
    counts = [0, 5, 3000, 20000, 1, 77]
    P = sum(counts)
    H = rng.normal(size=(P, N)).astype(np.float32)
    H[3000:3010] = H[3005]                       # ties: the first row must win
    segs = model.Segments(counts, "cuda")
    Hd = torch.tensor(H, device="cuda")
    out = torch.empty((len(counts), N), dtype=torch.float32, device="cuda")
    g = torch.empty_like(out)
    arg = torch.empty((len(counts), N), dtype=torch.int32, device="cuda")
    B.check(B.LIB.lddmm_nn_seg_colsum(B.ptr(Hd), B.ptr(segs.off), len(counts), N, B.ptr(out), B.stream_handle()))
    B.check(B.LIB.lddmm_nn_segmax(B.ptr(Hd), B.ptr(segs.off), len(counts), N, B.ptr(g), B.ptr(arg),
                                  B.stream_handle()))
    off = np.concatenate([[0], np.cumsum(counts)])
    for s in range(len(counts)):
        blk = H[off[s]:off[s + 1]].astype(np.float64)
        if counts[s] == 0:
            np.testing.assert_array_equal(out[s].cpu().numpy(), 0)
            np.testing.assert_array_equal(g[s].cpu().numpy(), 0)
            np.testing.assert_array_equal(arg[s].cpu().numpy(), off[s])
            continue
        np.testing.assert_allclose(out[s].cpu().numpy(), blk.sum(0), rtol=1e-4, atol=1e-3)
        np.testing.assert_array_equal(g[s].cpu().numpy(), blk.max(0).astype(np.float32))
        np.testing.assert_array_equal(arg[s].cpu().numpy(), off[s] + blk.argmax(0))


def test_adam_step(net_mod):
    model, torch = net_mod
    net = model.MomentumNet(seed=1)
    w0 = net.params_numpy()
    g = np.random.default_rng(2).normal(size=w0.size).astype(np.float32)
    net.g.copy_(torch.tensor(g, device="cuda"))
    net.adam(lr=1e-2)
    w1 = net.params_numpy()
    # first step: m = 0.1 g, v = 0.001 g^2 -> update = lr * g / (|g| + eps) after bias correction
    expect = w0 - 1e-2 * g / (np.abs(g) + 1e-8)
    assert rel(w1, expect) < 1e-5
    bf = net.wbf.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    np.testing.assert_array_equal(bf.view(np.float32), model.bf16_round(w1))


def test_trained_network_warm_start_gain(net_mod):
    """The committed network on held-out pairs of the plain workload.  It reads
    the ray-cast normal height h, the part of h n that the feature-only
    control's initial velocity misses and h averaged at width 2.5, and
    predicts the residual of the re-ray-cast (refined) velocity over h n in
    the local vertex frames; trained on 20,000 pairs.  Each warm-started
    registration reaches the plain run's final varifold energy (1e-4
    relative) from the warm control.  In total over six seeds:
    * the network's velocity solved directly takes no more BFGS iterations
      than the feature-only control (h n solved the same way, no network);
    * the committed pipeline (the network's velocity re-ray-cast once) takes
      no more than h n re-ray-cast twice without the network;
    * both take at least 3x fewer than the zero-momentum plain runs.
    Measured on 60 held-out pairs: network only 4.63, feature-only 6.05,
    plain 49.85 iterations; on the six seeds below 28 vs 34 vs 299."""
    model, torch = net_mod
    import os
    from lddmm_amd.learned import warmstart as W
    from lddmm_amd.learned.experiment import plain_with_trace, iterations_to
    path = os.path.join(os.path.dirname(W.__file__), "momentum_net.npz")
    net = model.MomentumNet.load(path)
    assert net.target == 'vresid' and net.features == 'normal_lin' and net.solve_iters == W.SOLVE_ITERS
    assert W.NET_REFINE_STEPS == 1
    its = {k: [] for k in ("plain", "network", "network_only", "feature_only", "refined2")}
    refined2 = lambda sm: W.time_constant(sm, W.solve_control(sm, W.refined_velocity(sm, 2)))
    for seed in range(900000, 900006):
        sm, _, _ = plain_with_trace(seed, W.DEFAULT)
        E = sm.varifold_data()
        it_plain = iterations_to(sm.data_trace, E)
        runs = {"network": W.register(seed, net=net, target_data=E),
                "network_only": W.register(seed, net=net, target_data=E, refine=0),
                "feature_only": W.register(seed, start=lambda sm: W.feature_only_control(sm)[0], target_data=E),
                "refined2": W.register(seed, start=refined2, target_data=E)}
        print(f"[warmstart] seed {seed}: plain {it_plain} it, "
              + ", ".join(f"{k} {r['iterations']} it" for k, r in runs.items()), flush=True)
        for k, r in runs.items():
            assert r["reached"], k
            assert r["iterations"] < it_plain, k
            its[k].append(r["iterations"])
        assert runs["network"]["used_warm"] and runs["network_only"]["used_warm"]
        its["plain"].append(it_plain)
    tot = {k: sum(v) for k, v in its.items()}
    assert tot["network_only"] <= tot["feature_only"]
    assert tot["network"] <= tot["refined2"]
    assert tot["plain"] >= 3 * tot["network_only"]
    assert tot["plain"] >= 3 * tot["network"]
