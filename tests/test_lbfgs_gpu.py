"""The device L-BFGS direction of SurfaceMatching.twoLoop against the reference
recursion of base/bfgs.py:188-207 evaluated in float64 on the host: the
compact form (lddmm_lbfgs_multidot + lddmm_lbfgs_compact) for memories of 1
to 64 pairs, including a sliding window whose cached Gram columns are longer
than the window, and the two-loop recursion (lddmm_lbfgs_two_loop) beyond 64
pairs; determinism (two calls give the same bits)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def reference_two_loop(g, stored):
    dot = lambda a, b: float((a * b).sum())
    q = g.copy()
    rho, alpha = [], []
    for s, y in reversed(stored):
        rho.append(1 / dot(y, s))
        alpha.append(rho[-1] * dot(s, q))
        q = q - alpha[-1] * y
    rho.reverse()
    alpha.reverse()
    s, y = stored[-1]
    c = dot(s, y) / dot(y, y)
    c = 1 if c < 1e-10 else c
    q = c * q
    for k, (s, y) in enumerate(stored):
        beta = rho[k] * dot(y, q)
        q = q + (alpha[k] - beta) * s
    return q


@pytest.mark.parametrize("m,shape", [(1, (3, 17, 3)), (2, (10, 2562, 3)), (7, (4, 1000, 2)), (50, (10, 2562, 3)),
                                     (64, (2, 700, 3)), (70, (2, 700, 3))])
def test_device_two_loop_equals_reference(m, shape):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lddmm_amd.surfaceMatching import SurfaceMatching, Control
    rng = np.random.default_rng(m)
    stored_h, stored_d = [], []
    for _ in range(m):
        s = rng.normal(size=shape)
        y = s * rng.uniform(0.5, 2.0) + 0.1 * rng.normal(size=shape)
        stored_h.append((s, y))
        cs, cy = Control(), Control()
        cs['at'], cy['at'] = torch.tensor(s, device='cuda'), torch.tensor(y, device='cuda')
        stored_d.append([cs, cy])
    g = rng.normal(size=shape)
    cg = Control()
    cg['at'] = torch.tensor(g, device='cuda')
    q1 = SurfaceMatching.twoLoop(None, cg, stored_d)['at']
    q2 = SurfaceMatching.twoLoop(None, cg, stored_d)['at']
    assert torch.equal(q1, q2)
    qr = reference_two_loop(g, stored_h)
    assert np.max(np.abs(q1.cpu().numpy() - qr)) <= 1e-11 * np.max(np.abs(qr))


def test_sliding_window_reuses_cached_columns():
    """bfgs.bfgs appends one pair per iteration and drops the oldest beyond the
    memory: every direction of a run with memory 5 over 12 pairs matches the
    reference recursion on the current window, and a fresh copy of the
    window (no caches) gives the same bits as the cached one."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lddmm_amd.surfaceMatching import SurfaceMatching, Control
    rng = np.random.default_rng(12)
    shape = (5, 800, 3)
    win_h, win_d = [], []
    for it in range(12):
        s = rng.normal(size=shape)
        y = s * rng.uniform(0.5, 2.0) + 0.2 * rng.normal(size=shape)
        cs, cy = Control(), Control()
        cs['at'], cy['at'] = torch.tensor(s, device='cuda'), torch.tensor(y, device='cuda')
        win_h.append((s, y))
        win_d.append([cs, cy])
        if len(win_h) > 5:
            win_h.pop(0)
            win_d.pop(0)
        g = rng.normal(size=shape)
        cg = Control()
        cg['at'] = torch.tensor(g, device='cuda')
        q = SurfaceMatching.twoLoop(None, cg, win_d)['at']
        qr = reference_two_loop(g, win_h)
        assert np.max(np.abs(q.cpu().numpy() - qr)) <= 1e-11 * np.max(np.abs(qr)), it
        fresh = [[e[0], e[1]] for e in win_d]
        q2 = SurfaceMatching.twoLoop(None, cg, fresh)['at']
        assert torch.equal(q, q2), it
