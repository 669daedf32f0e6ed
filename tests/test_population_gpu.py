"""Population (batched) kernels and lockstep registration on the GPU.

Each batched result is compared, pair by pair, with the single-pair entry
points (themselves pinned to the oracle): FP64 <= 1e-11 relative, FP32
compensated <= 2e-5.  Rows of inactive pairs must stay untouched; graph
replays must equal eager runs bit for bit, replay after replay.  The
lockstep registration must follow each pair's own SurfaceMatching run.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = {'float64': 1e-11, 'float32': 2e-5}


@pytest.fixture(scope="module")
def env():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lddmm_amd import synth, pointEvolution as evol, surfaceDistances as sd, _backend as B
    from lddmm_amd.population import Population, PopulationMatching
    from lddmm_amd.surfaces import Surface
    from lddmm_amd.kernelFunctions import Kernel
    return dict(torch=torch, synth=synth, evol=evol, sd=sd, B=B, Population=Population,
                PopulationMatching=PopulationMatching, Surface=Surface, Kernel=Kernel)


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def pairs(env, seed=3, P=5):
    return env["synth"].population_pairs(P, seed=seed, nmin=60, nmax=1500)


def build(env, prs):
    S = env["Surface"]
    tm = [S((f, v)) for (v, f), _ in prs]
    tg = [S((f2, v2)) for _, (v2, f2) in prs]
    return tm, tg, env["Population"](tm, tg)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("sig", [[3.0], [2.0, 5.0]])
def test_population_forward_adjoint_match_single(env, dtype, sig):
    torch, evol, B = env["torch"], env["evol"], env["B"]
    prs = pairs(env)
    tm, tg, pop = build(env, prs)
    K = env["Kernel"](name='gauss', sigma=sig)
    K.pk_dtype = dtype
    kd = B.kernel_desc(K.sigma, K.ms_exponent, dtype)
    T = 6
    rng = np.random.default_rng(1)
    at = torch.tensor(rng.normal(size=(T, pop.V, 3)) * 0.3, device='cuda')
    xt = pop.x0.expand(T + 1, -1, -1).clone()
    en = torch.zeros(pop.P, dtype=torch.float64, device='cuda')
    pop.shoot(kd, at, xt, 0.7, en)
    px1 = torch.tensor(rng.normal(size=(pop.V, 3)), device='cuda')
    g = torch.zeros_like(at)
    pop.adjoint(kd, at, xt, px1, g, 0.7, euclidean=True)
    gn = torch.zeros_like(at)
    pop.adjoint(kd, at, xt, px1, gn, 0.7, euclidean=False)
    for p in range(pop.P):
        r0, r1 = int(pop.voff[p]), int(pop.voff[p + 1])
        xs, es = evol.shoot(pop.x0[r0:r1], at[:, r0:r1], K, regweight=0.7, energy=True)
        assert rel(xt[:, r0:r1].cpu(), xs.cpu()) < TOL[dtype]
        assert abs(en[p].item() - es.item()) <= TOL[dtype] * abs(es.item())
        gs, _ = evol.shoot_adjoint(xs, at[:, r0:r1], px1[r0:r1], K, 0.7, euclidean=True)
        assert rel(g[:, r0:r1].cpu(), gs.cpu()) < 5 * TOL[dtype]
        gs2, _ = evol.shoot_adjoint(xs, at[:, r0:r1], px1[r0:r1], K, 0.7, euclidean=False)
        assert rel(gn[:, r0:r1].cpu(), gs2.cpu()) < 5 * TOL[dtype]
    pop.close()


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_population_varifold_matches_single(env, dtype):
    torch, sd, B = env["torch"], env["sd"], env["B"]
    prs = pairs(env, seed=5, P=6)
    tm, tg, pop = build(env, prs)
    kd = B.kernel_desc([1.5], -1, dtype)
    x = pop.x0 + 0.2 * torch.randn_like(pop.x0)
    e = torch.zeros(pop.P, dtype=torch.float64, device='cuda')
    pop.varifold(kd, x, energies=e)
    e2 = torch.zeros_like(e)
    g = torch.zeros_like(x)
    pop.varifold(kd, x, energies=e2, grad=g)
    for p in range(pop.P):
        r0, r1 = int(pop.voff[p]), int(pop.voff[p + 1])
        es, gs = sd.varifold_device(kd, tm[p], tg[p], energy=True, grad=True, v1=x[r0:r1])
        assert abs(e[p].item() - es.item()) <= TOL[dtype] * abs(es.item()) + 1e-12
        assert abs(e2[p].item() - es.item()) <= TOL[dtype] * abs(es.item()) + 1e-12
        assert rel(g[r0:r1].cpu(), gs.cpu()) < TOL[dtype]
    pop.close()


def test_population_inactive_rows_untouched_and_seg_ops(env):
    torch, B = env["torch"], env["B"]
    prs = pairs(env, seed=7, P=5)
    tm, tg, pop = build(env, prs)
    kd = B.kernel_desc([3.0], -1, 'float64')
    T = 4
    at = torch.randn((T, pop.V, 3), dtype=torch.float64, device='cuda') * 0.2
    full = pop.x0.expand(T + 1, -1, -1).clone()
    pop.shoot(kd, at, full, 1.0)
    act = np.array([1, 0, 1, 0, 1], dtype=bool)
    pop.set_active(act)
    xt = torch.full_like(full, 7.0)
    xt[0] = pop.x0
    pop.shoot(kd, at, xt, 1.0)
    rows = pop.rows(act)
    assert torch.equal(xt[1:, ~rows], torch.full_like(xt[1:, ~rows], 7.0))
    assert rel(xt[:, rows].cpu(), full[:, rows].cpu()) < 1e-14
    # per-pair dot / axpy against torch on the rows
    a = torch.randn_like(at)
    b = torch.randn_like(at)
    d = pop.seg_dot(a, b).cpu().numpy()
    for p in range(pop.P):
        r0, r1 = int(pop.voff[p]), int(pop.voff[p + 1])
        want = float((a[:, r0:r1] * b[:, r0:r1]).sum()) if act[p] else 0.0
        assert abs(d[p] - want) <= 1e-12 * max(1.0, abs(want))
    y = a.clone()
    alpha = torch.arange(pop.P, dtype=torch.float64, device='cuda') + 0.5
    pop.seg_axpy(y, alpha, b)
    want = a + torch.where(rows[None, :, None], alpha[pop.vpair][None, :, None] * b, torch.zeros_like(b))
    assert torch.allclose(y, want, rtol=0, atol=1e-14)
    pop.close()


def test_population_graph_replays_equal_eager(env):
    """Captured shooting loops replayed several times, with eager work and
    synchronisations in between, equal the eager results exactly."""
    torch, B = env["torch"], env["B"]
    prs = pairs(env, seed=9, P=4)
    tm, tg, pop = build(env, prs)
    kd = B.kernel_desc([2.5], -1, 'float32')
    T = 5
    at = torch.randn((T, pop.V, 3), dtype=torch.float64, device='cuda') * 0.3
    px1 = torch.randn((pop.V, 3), dtype=torch.float64, device='cuda')
    pop.set_graphs(False)
    xe = pop.x0.expand(T + 1, -1, -1).clone()
    ee = torch.zeros(pop.P, dtype=torch.float64, device='cuda')
    pop.shoot(kd, at, xe, 1.0, ee)
    ge = torch.zeros_like(at)
    pop.adjoint(kd, at, xe, px1, ge, 1.0)
    pop.set_graphs(True)
    xg = pop.x0.expand(T + 1, -1, -1).clone()
    eg = torch.zeros_like(ee)
    gg = torch.zeros_like(at)
    for rep in range(4):
        xg[1:] = 0.0
        eg.zero_()
        gg.zero_()
        pop.shoot(kd, at, xg, 1.0, eg)
        pop.adjoint(kd, at, xg, px1, gg, 1.0)
        torch.cuda.synchronize()
        assert torch.equal(xg, xe), rep
        assert torch.equal(eg, ee), rep
        assert torch.equal(gg, ge), rep
        _ = (at * 2.0).sum().item()          # eager work between replays
    # a different active set replays the same graph on other pairs (its own
    # item table): equal to the eager run of that set, other rows untouched
    act = np.array([0, 1, 1, 0], dtype=bool)
    pop.set_active(act)
    rows = pop.rows(act)
    pop.set_graphs(False)
    xs = torch.full_like(xe, 3.0)
    xs[0] = pop.x0
    pop.shoot(kd, at, xs, 1.0, eg)
    pop.set_graphs(True)
    xg2 = torch.empty_like(xe)
    for rep in range(3):
        xg2.fill_(3.0)
        xg2[0] = pop.x0
        pop.shoot(kd, at, xg2, 1.0, eg)
        torch.cuda.synchronize()
        assert torch.equal(xg2, xs), rep
    assert rel(xs[:, rows].cpu(), xe[:, rows].cpu()) < 2e-6
    assert torch.equal(xs[1:, ~rows], torch.full_like(xs[1:, ~rows], 3.0))
    assert len(pop.tuning()) >= 2
    pop.close()


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_lockstep_registration_follows_surface_matching(env, dtype):
    from lddmm_amd.surfaceMatching import SurfaceMatching
    Kernel, Surface = env["Kernel"], env["Surface"]
    prs = pairs(env, seed=11, P=3)
    opts = dict(KparDiff=Kernel(name='gauss', sigma=3.0), KparDist=Kernel(name='gauss', sigma=1.0),
                errorType='varifold', timeStep=0.1, pk_dtype=dtype, maxIter=8, mode='quiet')
    pm = env["PopulationMatching"]([Surface((f, v)) for (v, f), _ in prs], [Surface((f2, v2)) for _, (v2, f2) in prs],
                                   {k: v for k, v in opts.items() if k != 'mode'})
    pm.optimizeMatching()
    tol = 1e-9 if dtype == 'float64' else 2e-5
    for p, ((v, f), (v2, f2)) in enumerate(prs):
        o = dict(opts)
        o['KparDiff'] = Kernel(name='gauss', sigma=3.0)
        o['KparDist'] = Kernel(name='gauss', sigma=1.0)
        sm = SurfaceMatching(Template=Surface((f, v)), Target=Surface((f2, v2)), options=o)
        sm.optimizeMatching()
        h1, h2 = np.array(pm.history[p]), np.array(sm.bfgs_history)
        assert len(h1) == len(h2), (p, len(h1), len(h2))
        assert np.max(np.abs(h1 - h2) / np.abs(h2)) < tol, (p, h1, h2)
        assert int(pm.iterations[p]) == int(sm.bfgs_iterations)


def test_batched_orientation_check_matches_surface_methods(env):
    """closed_and_volumes (all surfaces' edges sorted together on the device)
    against Surface.getEdges' boundary flags and Surface.surfVolume: closed
    spheres, flipped ones (negative volume), a mesh with a face removed (open)
    and an ellipsoid hull."""
    from lddmm_amd import synth
    from lddmm_amd.population import closed_and_volumes
    Surface = env["Surface"]
    v1, f1 = synth.icosphere(2, 3.0)
    v2, f2 = synth.icosphere(3, 1.5)
    surfs = [Surface((f1, v1)), Surface((f2[:, [0, 2, 1]], v2 + 4.0)), Surface((f1[1:], v1)),
             Surface((f2, v2 * np.array([2.0, 1.0, 0.5])))]
    (vh, fh), _ = synth.population_pairs(1, seed=3)[0]
    surfs.append(Surface((fh, vh)))
    closed, vol = closed_and_volumes(surfs)
    for s, c, v in zip(surfs, closed, vol):
        s.getEdges()
        assert bool(c) == (s.bdry.max() == 0)
        assert abs(v - s.surfVolume()) <= 1e-9 * abs(s.surfVolume())
    assert list(closed) == [True, True, False, True, True]
    assert vol[1] < 0 < vol[0]
