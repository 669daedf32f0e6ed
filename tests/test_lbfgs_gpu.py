"""The device L-BFGS two-loop (lddmm_lbfgs_two_loop through
SurfaceMatching.twoLoop) against the reference recursion of base/bfgs.py:188-207
evaluated in float64 on the host, for memories of 1 to 50 pairs, and its
determinism (two calls give the same bits)."""
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


@pytest.mark.parametrize("m,shape", [(1, (3, 17, 3)), (2, (10, 2562, 3)), (7, (4, 1000, 2)), (50, (10, 2562, 3))])
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
