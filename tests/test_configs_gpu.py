"""The workload configurations at their full sizes against the FP64 oracle.

Every GPU result here is a sum over all sources; the oracle recomputes a
sample of its rows exactly (each sampled row is still a sum over every
source), so the checks do not depend on the problem size:

* kernel-sum microbenchmark, varifold norm: two seeded sphere hulls of
  100,000 triangles each, spatial sigma 1 -- 400 sampled face rows of each
  varifold block (applyktensor) against the oracle over all 100k sources, and
  the fused varifold energy against the sum of the pinned rows;
* large single registration: icosphere 6 (40,962 vertices, 81,920 faces,
  R = 50) to a 100-bump target, sigma_def 10, varifold sigma 2, T = 20 --
  shooting steps, adjoint (lddmm covector) steps, the Euclidean gradient and
  the varifold gradient on sampled rows, FP64 and compensated FP32;
* batched population: 256 pairs of 506-19,741 vertices, sigma_def 3,
  varifold sigma 1, T = 10 -- for the smallest, median and largest pair, the
  batched forward, adjoint and varifold gradient on sampled rows, energies
  where the oracle finishes in seconds.

Tolerances: FP64 1e-11 relative (norm-wise over the sampled rows), FP32
compensated 2e-5.
"""
import time

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {'float64': 1e-11, 'float32': 2e-5}


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def say(msg):
    print(f"[configs] {msg}", flush=True)


@pytest.fixture(scope="module")
def env():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lddmm_amd import synth, pointEvolution as evol, surfaceDistances as sd, _backend as B
    from lddmm_amd import kernelFunctions_util as ku
    from lddmm_amd.surfaces import Surface
    from lddmm_amd.kernelFunctions import Kernel
    from lddmm_amd.population import Population
    return dict(torch=torch, synth=synth, evol=evol, sd=sd, B=B, ku=ku, Surface=Surface, Kernel=Kernel,
                Population=Population)


def varifold_feats(v, f):
    c, S, fw = O.surface_geometry(v, f)
    a = np.sqrt((S ** 2).sum(1) + 1e-10)
    n = S / a[:, None]
    return c, a * fw, n * np.sqrt((a * fw)[:, None])


# ------------------------------------------------ config 2: varifold norm
@pytest.fixture(scope="module")
def cfg2(env):
    """The two 100k-triangle hulls, their face features and the oracle's rows
    of the three varifold blocks (computed once for both precisions)."""
    synth = env["synth"]
    v1, f1 = synth.sphere_hull(50_002, 50.0, seed=11)
    v2, f2 = synth.sphere_hull(50_002, 50.0, seed=12)
    v2 = synth.bump_deform(v2, f2, 30, 8.0, 2.0, seed=13)
    assert f1.shape[0] == f2.shape[0] == 100_000
    F1, F2 = varifold_feats(v1, f1), varifold_feats(v2, f2)
    rows = np.random.default_rng(2).choice(100_000, size=400, replace=False)
    blocks = {"S1.S1": (F1, F1), "S1.S2": (F1, F2), "S2.S2": (F2, F2)}
    refs = {}
    for name, ((cy, ay, by), (cx, ax, bx)) in blocks.items():
        say(f"varifold 100k oracle rows, block {name}")
        refs[name] = O.applyktensor(cy[rows], cx, ay[rows], ax, by[rows], bx, [1.0])
    return (v1, f1), (v2, f2), blocks, rows, refs


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_varifold_norm_100k_faces(env, cfg2, dtype):
    ku, sd, B = env["ku"], env["sd"], env["B"]
    (v1, f1), (v2, f2), blocks, rows, refs = cfg2
    sums = {}
    for name, ((cy, ay, by), (cx, ax, bx)) in blocks.items():
        say(f"varifold 100k {dtype} block {name}")
        got = ku.applyktensor(cy, cx, ay, ax, by, bx, 'gauss', [1.0], 10, dtype=dtype)
        assert rel(got[rows], refs[name]) < TOL[dtype], name
        sums[name] = got.sum()
    # the fused data-term kernels against the sums of the pinned rows
    S1, S2 = env["Surface"]((f1, v1)), env["Surface"]((f2, v2))
    kd = B.kernel_desc([1.0], -1, dtype)
    eD, _ = sd.varifold_device(kd, S1, S2)
    e0, _ = sd.varifold_device(kd, S2)
    want = sums["S1.S1"] - 2 * sums["S1.S2"] + sums["S2.S2"]
    got = float(eD.item()) + float(e0.item())
    assert abs(got - want) / abs(want) < (1e-11 if dtype == 'float64' else 1e-6)


# ---------------------------------------- config 3: large registration
@pytest.fixture(scope="module")
def cfg3(env):
    synth = env["synth"]
    v, f = synth.icosphere(6, 50.0)
    v2 = synth.bump_deform(v, f, 100, 8.0, 2.0, seed=11)
    rng = np.random.default_rng(31)
    at = rng.normal(size=(20, v.shape[0], 3)) * 0.05
    return v, f, v2, at


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_large_registration_rows(env, cfg3, dtype):
    torch, evol, sd, B = env["torch"], env["evol"], env["sd"], env["B"]
    v, f, v2, at = cfg3
    T, N = at.shape[0], v.shape[0]
    assert (N, f.shape[0], T) == (40962, 81920, 20)
    dt = 1.0 / T
    K = env["Kernel"](name='gauss', sigma=10.0)
    K.pk_dtype = dtype
    say(f"cfg3 {dtype}: shoot")
    xt_d, _ = evol.shoot(v, at, K)
    xt = xt_d.cpu().numpy()
    rng = np.random.default_rng(7)
    rows = rng.choice(N, size=256, replace=False)
    for t in (0, 10, T - 1):
        inc = (xt[t + 1, rows] - xt[t, rows]) / dt
        assert rel(inc, O.applyK(xt[t, rows], xt[t], at[t], [10.0])) < TOL[dtype], t
    # varifold gradient at the endpoint (sigma 2), used as the adjoint's start
    S, Tg = env["Surface"]((f, v)), env["Surface"]((f, v2))
    kdist = B.kernel_desc([2.0], -1, dtype)
    say(f"cfg3 {dtype}: varifold gradient")
    _, g1 = sd.varifold_device(kdist, S, Tg, energy=False, grad=True, v1=xt_d[-1].contiguous())
    g1h = g1.cpu().numpy()
    vrows = rows[:32]
    refg = O.varifoldNormGradient((xt[-1], f), (v2, f), [2.0], vertex_rows=vrows)
    assert rel(g1h[vrows], refg) < TOL[dtype]
    px1 = -g1h
    say(f"cfg3 {dtype}: adjoint")
    grad_d, pxt_d = evol.shoot_adjoint(xt_d, at, px1, K, 1.0, euclidean=False, covector=True)
    pxt, grad = pxt_d.cpu().numpy(), grad_d.cpu().numpy()
    np.testing.assert_allclose(pxt[T], px1, rtol=0, atol=0)
    for t in (0, 10, T - 1):
        inc = (pxt[t, rows] - pxt[t + 1, rows]) / dt
        ref = O.applyDiffKT(xt[t], xt[t], pxt[t + 1], at[t], [10.0], regweight=1.0, lddmm=True, rows=rows)
        assert rel(inc, ref) < TOL[dtype], t
        np.testing.assert_allclose(grad[t], (2 * at[t] - pxt[t + 1]) / T, rtol=1e-12, atol=1e-14)
    geu_d, _ = evol.shoot_adjoint(xt_d, at, px1, K, 1.0, euclidean=True)
    geu = geu_d.cpu().numpy()
    for t in (0, T - 1):
        ref = O.applyK(xt[t, rows], xt[t], 2 * at[t] - pxt[t + 1], [10.0]) / T
        assert rel(geu[t, rows], ref) < TOL[dtype], t


# ---------------------------------------------- config 4: population
@pytest.fixture(scope="module")
def cfg4(env):
    t0 = time.perf_counter()
    prs = env["synth"].population_pairs(256, seed=4)
    S = env["Surface"]
    tm = [S((f, v)) for (v, f), _ in prs]
    tg = [S((f2, v2)) for _, (v2, f2) in prs]
    pop = env["Population"](tm, tg)
    n = np.diff(pop.voff)
    order = np.argsort(n, kind='stable')
    pick = {"smallest": int(order[0]), "median": int(order[len(order) // 2]), "largest": int(order[-1])}
    say(f"cfg4 built in {time.perf_counter() - t0:.1f} s: V = {pop.V}, sizes "
        f"{[int(n[p]) for p in pick.values()]}")
    return prs, pop, pick


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_population_full_size_rows(env, cfg4, dtype):
    torch, B = env["torch"], env["B"]
    prs, pop, pick = cfg4
    assert pop.P == 256 and 500 <= np.diff(pop.voff).min() and np.diff(pop.voff).max() <= 20000
    T, dt = 10, 0.1
    kd = B.kernel_desc([3.0], -1, dtype)
    kdist = B.kernel_desc([1.0], -1, dtype)
    rng = np.random.default_rng(41)
    at_h = rng.normal(size=(T, pop.V, 3)) * 0.05
    at = torch.tensor(at_h, device='cuda')
    xt_d = pop.x0.expand(T + 1, -1, -1).clone()
    en = torch.zeros(pop.P, dtype=torch.float64, device='cuda')
    say(f"cfg4 {dtype}: forward")
    pop.shoot(kd, at, xt_d, 1.0, en)
    ven = torch.zeros(pop.P, dtype=torch.float64, device='cuda')
    vgrad = torch.zeros((pop.V, 3), dtype=torch.float64, device='cuda')
    say(f"cfg4 {dtype}: varifold")
    pop.varifold(kdist, xt_d[-1].contiguous(), ven, vgrad)
    px1 = -vgrad
    grad_d = torch.zeros_like(at)
    say(f"cfg4 {dtype}: adjoint")
    pop.adjoint(kd, at, xt_d, px1, grad_d, 1.0, euclidean=False)
    geu_d = torch.zeros_like(at)
    pop.adjoint(kd, at, xt_d, px1, geu_d, 1.0, euclidean=True)
    torch.cuda.synchronize()
    xt, en, ven = xt_d.cpu().numpy(), en.cpu().numpy(), ven.cpu().numpy()
    vg, grad, geu, p1 = vgrad.cpu().numpy(), grad_d.cpu().numpy(), geu_d.cpu().numpy(), px1.cpu().numpy()
    for name, p in pick.items():
        r0, r1 = int(pop.voff[p]), int(pop.voff[p + 1])
        n = r1 - r0
        (v, f), (v2, f2) = prs[p]
        rows = rng.choice(n, size=min(n, 256), replace=False)
        X = xt[:, r0:r1]
        A = at_h[:, r0:r1]
        say(f"cfg4 {dtype}: {name} pair {p} ({n} vertices)")
        for t in (0, 5, T - 1):
            inc = (X[t + 1, rows] - X[t, rows]) / dt
            assert rel(inc, O.applyK(X[t, rows], X[t], A[t], [3.0])) < TOL[dtype], (name, t)
        # adjoint: non-Euclidean gradient_t = (2 a_t - p_{t+1}) / T gives p_1 .. p_T
        Pn = np.empty((T + 1, n, 3))
        Pn[T] = p1[r0:r1]
        for t in range(T):
            Pn[t + 1] = 2 * A[t] - T * grad[t, r0:r1]
        np.testing.assert_allclose(Pn[T], p1[r0:r1], rtol=1e-12,
                                   atol=1e-12 * (np.abs(p1[r0:r1]).max() + np.abs(A).max()))
        for t in (1, 5, T - 1):
            inc = (Pn[t, rows] - Pn[t + 1, rows]) / dt
            ref = O.applyDiffKT(X[t], X[t], Pn[t + 1], A[t], [3.0], regweight=1.0, lddmm=True, rows=rows)
            assert rel(inc, ref) < max(TOL[dtype], 1e-10), (name, t)
        for t in (0, T - 1):
            ref = O.applyK(X[t, rows], X[t], 2 * A[t] - Pn[t + 1], [3.0]) / T
            assert rel(geu[t, r0:r1][rows], ref) < TOL[dtype], (name, t)
        vrows = rows[:48]
        refg = O.varifoldNormGradient((X[T], f), (v2, f2), [1.0], vertex_rows=vrows)
        assert rel(vg[r0:r1][vrows], refg) < TOL[dtype], name
        if n <= 5000:
            e_ref = sum(dt * (A[t] * O.applyK(X[t], X[t], A[t], [3.0])).sum() for t in range(T))
            assert abs(en[p] - e_ref) / abs(e_ref) < TOL[dtype], name
            d_ref = O.varifoldNormDef((X[T], f), (v2, f2), [1.0])
            assert abs(ven[p] - d_ref) / abs(d_ref) < (1e-10 if dtype == 'float64' else 1e-5), name
