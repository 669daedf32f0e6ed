#!/usr/bin/env python3
"""Kernel-sum benchmark of the py-lddmm hot path on one MI355X.

Workload (the reference's kernel-sum microbenchmark): applyK and applyDiffKT
(plain) with N = M = 50,000 points drawn uniform in [0, 50]^3 (seeded), weights
a, p ~ N(0, 1) of dimension 3, Gaussian sigma in {2.5, 5, 10}, each sigma
single-scale.  One step = applyK + applyDiffKT for the three sigmas = 6 sums of
N*M = 2.5e9 pair evaluations, in the compensated FP32 form.  Inputs are device
resident before the timed region; calls go through the C ABI on the current
stream.

Prints one JSON line (the driver's contract) with `roofline` (dominant kernel:
the applyDiffKT pair kernel, timed by HIP events around each launch on its own
stream during the timed region) and `cpu_baseline` (the FP64 OpenMP
restatement of the reference's numba loops on a bounded row sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "py-lddmm_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "Gaussian pair-evals/s; s per surface registration at matched energy vs plain"
SIGMAS = (2.5, 5.0, 10.0)
# algorithmic work per pair (FMA = 2 FLOP) and the MI355X published FP32 peak
FLOP_PER_PAIR = {"applyK": 15, "applyDiffKT": 22}
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_EXP_PER_S = 256 * 4 * 16 / 2 * 2.4e9   # v_exp_f32 issues at half the FMA rate (profiles/gfx950_rates.md)


def measured_traffic(n):
    """HBM bytes per launch of the dominant kernel (applyDiffKT at N = M = 50k),
    from the rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarised, with their
    commands, in profiles/k_pair_kernel_sums_50k.md; None for other sizes."""
    if n != 50_000:
        return None
    try:
        with open(os.path.join(ROOT, "profiles", "k_pair_kernel_sums_50k.md")) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch"):
                    return int(line.split("=")[1])
    except (OSError, ValueError):
        pass
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(y, x, a, p, target_s=12.0):
    """FP64 OpenMP restatement timed on a row sample of the same workload."""
    import numpy as np
    from oracle import cpu
    n = y.shape[0]
    rows = 256
    t0 = time.perf_counter()
    cpu.applyK(y[:rows], x, a, [SIGMAS[0]])
    cpu.applyDiffKT(y[:rows], x, p[:rows], a, [SIGMAS[0]])
    dt = time.perf_counter() - t0
    rate = 2 * rows * x.shape[0] / dt
    rows = int(min(n, max(rows, target_s * rate / (2 * x.shape[0] * len(SIGMAS)))))
    total_pairs = 0
    t0 = time.perf_counter()
    for s in SIGMAS:
        cpu.applyK(y[:rows], x, a, [s])
        cpu.applyDiffKT(y[:rows], x, p[:rows], a, [s])
        total_pairs += 2 * rows * x.shape[0]
    dt = time.perf_counter() - t0
    return {
        "value": total_pairs / dt, "unit": "pair-evals/s", "cores": cpu.threads(),
        "kind": "port",
        "sample": (f"first {rows} of {n} targets x all {x.shape[0]} sources, applyK + applyDiffKT "
                   f"(plain) for sigma in {list(SIGMAS)}, float64, {total_pairs:.3g} pairs in {dt:.1f} s"),
    }


def varifold_microbench(B, torch, synth, np):
    """Varifold norm between two seeded random triangulated surfaces of 100,000
    triangles each (50,002 points on a radius-50 sphere -> convex hull; the
    second with another seed and radial bumps), spatial sigma = 1: three blocks
    of 1e10 face pairs (|S1|^2, -2<S1,S2>, |S2|^2)."""
    from lddmm_amd.surfaces import Surface
    from lddmm_amd import surfaceDistances as sd
    v1, f1 = synth.sphere_hull(50_002, 50.0, seed=11)
    v2, f2 = synth.sphere_hull(50_002, 50.0, seed=12)
    v2 = synth.bump_deform(v2, f2, 30, 8.0, 2.0, seed=13)
    S1, S2 = Surface((f1, v1)), Surface((f2, v2))
    res = {"faces": [int(f1.shape[0]), int(f2.shape[0])], "sigma": 1.0}
    vals = {}
    for pr in ("float32", "float64"):
        kd = B.kernel_desc([1.0], -1, pr)
        sd.varifold_device(kd, S1, S2)          # warm-up (features, adjacency, workspace)
        sd.varifold_device(kd, S2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eD, _ = sd.varifold_device(kd, S1, S2)
        e0, _ = sd.varifold_device(kd, S2)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        vals[pr] = float(eD.item()) + float(e0.item())
        pairs = f1.shape[0] * (f1.shape[0] + f2.shape[0]) + f2.shape[0] ** 2
        res[pr] = {"seconds": dt, "face_pairs_per_s": pairs / dt, "varifold_norm": vals[pr]}
        log(f"varifold 100k ({pr}): {dt * 1e3:.1f} ms, {pairs / dt:.3g} face-pairs/s, norm {vals[pr]:.6g}")
    res["fp32_rel_err_vs_fp64"] = abs(vals["float32"] - vals["float64"]) / abs(vals["float64"])
    # A/B: the float32 energy with the cross terms on the matrix cores
    kd = B.kernel_desc([1.0], -1, "float32")
    res["mfma_ab"] = {}
    for engine in ("mfma_cos", "mfma_cos_r2"):
        B.set_varifold_engine(engine)
        try:
            sd.varifold_device(kd, S1, S2)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            eD, _ = sd.varifold_device(kd, S1, S2)
            e0, _ = sd.varifold_device(kd, S2)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
        finally:
            B.set_varifold_engine("valu")
        val = float(eD.item()) + float(e0.item())
        res["mfma_ab"][engine] = {"seconds": dt, "face_pairs_per_s": pairs / dt,
                                  "speedup_vs_valu": res["float32"]["seconds"] / dt,
                                  "rel_err_vs_fp64": abs(val - vals["float64"]) / abs(vals["float64"])}
        log(f"varifold 100k ({engine}): {dt * 1e3:.1f} ms, {pairs / dt:.3g} face-pairs/s, "
            f"rel err vs fp64 {res['mfma_ab'][engine]['rel_err_vs_fp64']:.2e}")
    return res


def registration_cfg1(torch):
    """Plain surface matching (icosphere 4, R=10, 20 bumps width 3 / amp N(0,1),
    sigma_def 2.5, varifold sigma 1, T = 10, zero initial momentum, BFGS 50
    iterations, float64 and float32): wall seconds per registration (matching
    set-up included), after one untimed registration of the same pair."""
    import logging
    from lddmm_amd import synth
    from lddmm_amd.surfaces import Surface
    from lddmm_amd.kernelFunctions import Kernel
    from lddmm_amd.surfaceMatching import SurfaceMatching
    logging.getLogger().setLevel(logging.WARNING)
    (v, f), (v2, _) = synth.plain_matching_pair(seed=0)
    out = {}
    for pr in ("float64", "float32"):
        opts = {'KparDiff': Kernel(name='gauss', sigma=2.5), 'KparDist': Kernel(name='gauss', sigma=1.0),
                'errorType': 'varifold', 'timeStep': 0.1, 'pk_dtype': pr, 'maxIter': 50, 'mode': 'quiet'}
        # one untimed registration first: code-object loading and launch plans are one-time costs
        SurfaceMatching(Template=Surface((f, v)), Target=Surface((f, v2)), options=opts).optimizeMatching()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sm = SurfaceMatching(Template=Surface((f, v)), Target=Surface((f, v2)), options=opts)
        sm.optimizeMatching()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out[pr] = {"seconds": dt, "iterations": int(sm.bfgs_iterations), "final_objective": sm.obj,
                   "final_data_term": sm.varifold_data(), "initial_objective": sm.bfgs_history[0]}
        log(f"registration cfg1 ({pr}): {dt * 1e3:.1f} ms, {sm.bfgs_iterations} iterations, "
            f"objective {sm.bfgs_history[0]:.4f} -> {sm.obj:.4f}")
    return out


def population_registration(torch, npairs=256, iters=10, sample=16):
    """Batched population registration (PopulationMatching, lockstep L-BFGS,
    one persistent launch per kernel sum) of the population workload (256
    pairs, 506-19,741 vertices, sigma_def 3, varifold sigma 1, T = 10, FP32),
    `iters` BFGS iterations per pair, against the same registrations run one
    pair at a time by SurfaceMatching on every (256 / sample)-th pair (pair
    sizes are drawn independently, so the sample's time times 256 / sample
    estimates the whole population's)."""
    import logging
    import numpy as np
    from lddmm_amd import synth
    from lddmm_amd.population import PopulationMatching
    from lddmm_amd.surfaces import Surface
    from lddmm_amd.kernelFunctions import Kernel
    from lddmm_amd.surfaceMatching import SurfaceMatching
    logging.getLogger().setLevel(logging.WARNING)
    prs = synth.population_pairs(npairs, seed=4)

    def opts():
        return {'KparDiff': Kernel(name='gauss', sigma=3.0), 'KparDist': Kernel(name='gauss', sigma=1.0),
                'timeStep': 0.1, 'pk_dtype': 'float32', 'maxIter': iters}
    tm = [Surface((f, v)) for (v, f), _ in prs]
    tg = [Surface((f2, v2)) for _, (v2, f2) in prs]
    # first run: the launch-shape autotuner times its candidates on each op's
    # first call; the second run of the same population shape reuses them
    t_runs = []
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pm = PopulationMatching(tm, tg, opts())
        pm.optimizeMatching()
        torch.cuda.synchronize()
        t_runs.append(time.perf_counter() - t0)
        pm.pop.close()
    t_batch = t_runs[1]
    idx = list(range(0, npairs, npairs // sample))
    o = dict(opts(), errorType='varifold', mode='quiet')
    t_seq = 0.0
    for i in idx:
        (v, f), (v2, f2) = prs[i]
        o['KparDiff'] = Kernel(name='gauss', sigma=3.0)
        o['KparDist'] = Kernel(name='gauss', sigma=1.0)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        SurfaceMatching(Template=Surface((f, v)), Target=Surface((f2, v2)), options=o).optimizeMatching()
        torch.cuda.synchronize()
        t_seq += time.perf_counter() - t1
    t_seq_all = t_seq * npairs / len(idx)
    out = {"pairs": npairs, "bfgs_iterations": iters, "batched_seconds": t_batch,
           "batched_seconds_first_run_with_tuning": t_runs[0],
           "batched_seconds_per_pair": t_batch / npairs, "sequential_sample_pairs": len(idx),
           "sequential_seconds_estimated": t_seq_all, "batched_speedup": t_seq_all / t_batch}
    log(f"population registration: {npairs} pairs x {iters} iterations batched {t_batch:.2f} s "
        f"(first run, tuning included, {t_runs[0]:.2f} s), one at a time ~{t_seq_all:.2f} s "
        f"(sample of {len(idx)}), {t_seq_all / t_batch:.2f}x")
    return out


def second_order_registration(torch):
    """Geodesic point-set matching (SecondOrderPointSetMatching): the plain
    workload's template and target vertices (2,562 points each) as point sets,
    measure data term (sigma 1), Gaussian deformation kernel sigma 2.5, T = 10,
    50 BFGS iterations on the initial momentum; wall seconds per registration
    after one untimed run, float64 and float32."""
    import logging
    from lddmm_amd import synth
    from lddmm_amd.pointSets import PointSet
    from lddmm_amd.kernelFunctions import Kernel
    from lddmm_amd.secondOrderPointSetMatching import SecondOrderPointSetMatching
    logging.getLogger().setLevel(logging.WARNING)
    (v, f), (v2, _) = synth.plain_matching_pair(seed=0)
    out = {}
    for pr in ("float64", "float32"):
        def make():
            opts = {'KparDiff': Kernel(name='gauss', sigma=2.5), 'KparDist': Kernel(name='gauss', sigma=1.0),
                    'errorType': 'measure', 'timeStep': 0.1, 'pk_dtype': pr, 'maxIter': 50, 'mode': 'quiet'}
            return SecondOrderPointSetMatching(Template=PointSet(v), Target=PointSet(v2), options=opts)
        make().optimizeMatching()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m = make()
        m.optimizeMatching()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out[pr] = {"seconds": dt, "iterations": int(m.bfgs_iterations), "initial_objective": m.bfgs_history[0],
                   "final_objective": m.bfgs_history[-1]}
        log(f"second-order registration ({pr}): {dt:.3f} s, {m.bfgs_iterations} iterations, "
            f"objective {m.bfgs_history[0]:.4f} -> {m.bfgs_history[-1]:.4f}")
    return out


def warmstart_cfg5(torch, pairs=8):
    """Learned warm start on held-out pairs of the plain workload (seeds
    900000..): the zero-momentum plain registration (float64, 50 BFGS
    iterations) sets the target data term E*; each warm-started registration
    runs until its varifold data term is within 1e-4 of E*.  Side by side on
    the same pairs:
      network      - the committed pipeline: the network (ray-cast heights ->
                     residual of the refined velocity over h n), then one
                     re-ray-cast from the surface that velocity's control
                     reaches, then the truncated CG control;
      network_only - the network's velocity solved directly (no re-ray-cast);
      feature_only - h n solved the same way, no network;
      refined1 / refined2 - h n refined by one / two re-ray-casts (one extra
                     shoot and ray cast each), no network.
    Wall seconds cover matching set-up, features, network, solve and BFGS
    (the synthetic pair is generated before the timed region)."""
    import logging
    import os
    import numpy as np
    from lddmm_amd.learned import warmstart as W
    from lddmm_amd.learned.model import MomentumNet
    logging.getLogger().setLevel(logging.WARNING)
    net = MomentumNet.load(os.path.join(os.path.dirname(W.__file__), "momentum_net.npz"))

    def feature_only(sm):
        return W.feature_only_control(sm)[0]

    def refined(steps):
        return lambda sm: W.time_constant(sm, W.solve_control(sm, W.refined_velocity(sm, steps)))

    def network_only(sm):
        return W.warm_control(sm, net, refine=0)[0]

    starts = {"network": None, "network_only": network_only, "feature_only": feature_only,
              "refined1": refined(1), "refined2": refined(2)}
    # untimed warm-up of every start on a pair outside the measured seeds (first kernel loads)
    W.register(899999, net=net, max_iter=1)
    for name, make in starts.items():
        if make is not None:
            W.register(899999, start=make, max_iter=1)
    rows = []
    for k in range(pairs):
        seed = 900000 + k
        W.pair_arrays(seed)                   # generate the synthetic pair outside the timed region
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sm = W.make_matching(seed)
        sm.setTargetDataEnergy(None)
        sm.optimizeMatching()
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t0
        E = sm.varifold_data()
        obj = float(sm.objectiveFun())
        it_plain = next((i + 1 for i, d in enumerate(sm.data_trace) if d <= E * (1 + 1e-4)), len(sm.data_trace))
        row = {"plain": (it_plain, t_plain * it_plain / max(len(sm.data_trace), 1), True, 1.0)}
        for name, make in starts.items():
            r = W.register(seed, net=net, target_data=E) if make is None else W.register(seed, start=make,
                                                                                          target_data=E)
            row[name] = (r["iterations"], r["seconds"], r["reached"], r["objective"] / obj)
        rows.append(row)
    out = {"pairs": pairs}
    for name in ["plain"] + list(starts):
        it = np.array([r[name][0] for r in rows], dtype=float)
        out[name] = {"mean_iterations": float(it.mean()), "mean_seconds_to_E": float(np.mean([r[name][1] for r in rows])),
                     "all_reached": bool(all(r[name][2] for r in rows)),
                     "mean_final_objective_ratio": float(np.mean([r[name][3] for r in rows]))}
    it_p = sum(r["plain"][0] for r in rows)
    for name in starts:
        out[name]["iteration_ratio_vs_plain"] = float(it_p / max(sum(r[name][0] for r in rows), 1))
    log("warm start cfg5: " + ", ".join(f"{n} {out[n]['mean_iterations']:.2f} it {out[n]['mean_seconds_to_E']:.3f} s"
                                         for n in ["plain"] + list(starts)))
    return out


def precision_cfg3(torch):
    """Large single registration with the precision form: icosphere subdivision 6
    (40,962 vertices, 81,920 faces, R = 50), target with 100 bumps (width 8,
    amplitude N(0, 2^2)), sigma_def 10, varifold sigma 2, T = 20.  The control is
    the one after 5 FP64 BFGS iterations; at that control the FP32 compensated
    objective and gradient are compared with FP64, and one objective + gradient
    evaluation is timed in each precision."""
    import logging
    import numpy as np
    from lddmm_amd import synth
    from lddmm_amd.surfaces import Surface
    from lddmm_amd.kernelFunctions import Kernel
    from lddmm_amd.surfaceMatching import SurfaceMatching
    logging.getLogger().setLevel(logging.WARNING)
    v, f = synth.icosphere(6, 50.0)
    v2 = synth.bump_deform(v, f, 100, 8.0, 2.0, seed=11)

    def make(pr, iters):
        o = {'KparDiff': Kernel(name='gauss', sigma=10.0), 'KparDist': Kernel(name='gauss', sigma=2.0),
             'errorType': 'varifold', 'timeStep': 0.05, 'pk_dtype': pr, 'maxIter': iters, 'mode': 'quiet'}
        return SurfaceMatching(Template=Surface((f, v)), Target=Surface((f, v2)), options=o)

    sm64 = make('float64', 5)
    t0 = time.perf_counter()
    sm64.optimizeMatching()
    torch.cuda.synchronize()
    t_reg5 = time.perf_counter() - t0
    at = sm64.control['at'].clone()
    res = {"vertices": int(v.shape[0]), "faces": int(f.shape[0]), "T": 20,
           "fp64_5_iterations_s": t_reg5}
    vals = {}
    for pr in ("float64", "float32"):
        sm = make(pr, 1)
        sm.setInitialControl(at)
        sm.obj = None
        sm.objectiveFun()
        sm.getGradient()
        torch.cuda.synchronize()
        t = time.perf_counter()
        sm.obj = None
        obj = sm.objectiveFun()
        g = sm.getGradient()['at']
        torch.cuda.synchronize()
        res[f"{pr}_objective_plus_gradient_s"] = time.perf_counter() - t
        vals[pr] = (obj, g, sm.varifold_data())
    o64, g64, d64 = vals["float64"]
    o32, g32, d32 = vals["float32"]
    res["objective_fp64"] = o64
    res["objective_rel_err_fp32"] = abs(o32 - o64) / abs(o64)
    res["data_term_rel_err_fp32"] = abs(d32 - d64) / abs(d64)
    res["gradient_rel_err_fp32"] = float((g32 - g64).norm() / g64.norm())
    log(f"precision cfg3: objective rel err {res['objective_rel_err_fp32']:.2e}, gradient rel err "
        f"{res['gradient_rel_err_fp32']:.2e}; objective+gradient FP64 "
        f"{res['float64_objective_plus_gradient_s'] * 1e3:.0f} ms, FP32 "
        f"{res['float32_objective_plus_gradient_s'] * 1e3:.0f} ms")
    return res


def population_cfg4(torch, npairs=256, T=10):
    """Batched population workload: 256 template-target pairs, vertex counts
    log-uniform in [500, 20000] (Fibonacci sphere / ellipsoid hulls with bumps),
    sigma_def 3, varifold sigma 1, T = 10, FP32.  Times one batched forward
    shoot, one Euclidean adjoint and one varifold gradient of all pairs (one
    persistent launch per Gaussian sum, shooting loops replayed from hipGraphs),
    and the same forward shoot as 256 single-pair calls."""
    import numpy as np
    from lddmm_amd import synth, _backend as B, pointEvolution as evol
    from lddmm_amd.population import Population
    from lddmm_amd.surfaces import Surface
    from lddmm_amd.kernelFunctions import Kernel
    t0 = time.perf_counter()
    prs = synth.population_pairs(npairs, seed=4)
    tm = [Surface((f, v)) for (v, f), _ in prs]
    tg = [Surface((f2, v2)) for _, (v2, f2) in prs]
    pop = Population(tm, tg)
    gen_s = time.perf_counter() - t0
    n = np.diff(pop.voff).astype(np.float64)
    nf = np.diff(pop.foff).astype(np.float64)
    ntf = np.diff(pop.tfoff).astype(np.float64)
    kd = B.kernel_desc([3.0], -1, "float32")
    kv = B.kernel_desc([1.0], -1, "float32")
    rng = np.random.default_rng(0)
    at = torch.tensor(rng.normal(size=(T, pop.V, 3)) * 0.05, device="cuda")
    xt = pop.x0.expand(T + 1, -1, -1).clone()
    en = torch.zeros(pop.P, dtype=torch.float64, device="cuda")
    px1 = torch.zeros((pop.V, 3), dtype=torch.float64, device="cuda")
    g = torch.zeros_like(at)
    e = torch.zeros_like(en)

    def timed(fn, reps=3):
        fn()                                  # tuning / capture
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps

    fwd = timed(lambda: pop.shoot(kd, at, xt, 1.0, en))
    pop.varifold(kv, xt[-1], energies=e, grad=px1)
    adj = timed(lambda: pop.adjoint(kd, at, xt, px1, g, 1.0, euclidean=True))
    vf = timed(lambda: pop.varifold(kv, xt[-1], energies=e, grad=px1))
    # the same forward shoot, one pair per call
    K = Kernel(name="gauss", sigma=3.0)
    K.pk_dtype = "float32"
    offs = [(int(pop.voff[p]), int(pop.voff[p + 1])) for p in range(pop.P)]
    ats = [at[:, r0:r1].contiguous() for r0, r1 in offs]
    x0s = [pop.x0[r0:r1].contiguous() for r0, r1 in offs]
    for p in range(min(4, pop.P)):
        evol.shoot(x0s[p], ats[p], K, energy=True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for p in range(pop.P):
        evol.shoot(x0s[p], ats[p], K, energy=True)
    torch.cuda.synchronize()
    seq = time.perf_counter() - t
    pk = float(T * (n ** 2).sum())
    out = {
        "pairs": int(pop.P), "vertices": int(pop.V), "faces": int(pop.foff[-1]),
        "vertex_count_range": [int(n.min()), int(n.max())], "generation_s": gen_s,
        "pair_evals_per_forward": pk,
        "forward_ms": fwd * 1e3, "forward_pair_evals_per_s": pk / fwd,
        "adjoint_ms": adj * 1e3, "adjoint_pair_evals_per_s": 2 * pk / adj,
        "varifold_gradient_ms": vf * 1e3,
        "varifold_face_pairs_per_s": float((nf * (nf + ntf)).sum()) / vf,
        "sequential_single_pair_forward_ms": seq * 1e3, "batched_speedup_forward": seq / fwd,
        "tuning": pop.tuning(),
    }
    log(f"population cfg4: {pop.P} pairs, V={pop.V}: forward {fwd * 1e3:.1f} ms "
        f"({pk / fwd:.3g} pair-evals/s), adjoint {adj * 1e3:.1f} ms, varifold grad {vf * 1e3:.1f} ms; "
        f"256 single-pair forwards {seq * 1e3:.1f} ms")
    pop.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=50_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 reference pass")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the varifold and registration measurements reported under `extra`")
    ap.add_argument("--probe", action="store_true",
                    help="only launch the pair kernels (applyK and plain applyDiffKT, sigma 2.5, FP32 then "
                         "FP64) 3 times each and a 256 MiB device copy 3 times, for rocprofv3 counter runs")
    ap.add_argument("--registration", action="store_true",
                    help="only run the plain registration (icosphere 4, float64 and float32) three times, "
                         "for rocprofv3 kernel traces of the registration loop")
    ap.add_argument("--second-order", action="store_true",
                    help="only run the geodesic point-set registration (2,562 points, float64 and float32: one "
                         "untimed and one timed run each), for rocprofv3 kernel traces")
    ap.add_argument("--population", action="store_true",
                    help="only run the batched population workload (256 pairs: forward, adjoint, varifold "
                         "gradient, and the forward as 256 single-pair calls), for rocprofv3 kernel traces")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the kernel-sum path shards by target rows with no "
              "communication; multi-GPU runs are not built.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from lddmm_amd import _backend as B
    from lddmm_amd import synth

    if not torch.cuda.is_available():
        log("no GPU visible")
        sys.exit(1)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.registration:
        for rep in range(3):
            registration_cfg1(torch)
        return
    if args.second_order:
        print(json.dumps(second_order_registration(torch)), flush=True)
        return
    if args.population:
        print(json.dumps(population_cfg4(torch)), flush=True)
        return
    N = M = args.n
    y_h = synth.uniform_cloud(N, seed=1)
    x_h = synth.uniform_cloud(M, seed=2)
    rng = np.random.default_rng(3)
    a_h = rng.normal(size=(M, 3))
    p_h = rng.normal(size=(N, 3))
    y, x, a, p = (torch.tensor(v, device=dev) for v in (y_h, x_h, a_h, p_h))
    outK = torch.empty((N, 3), dtype=torch.float64, device=dev)
    outD = torch.empty((N, 3), dtype=torch.float64, device=dev)
    st = B.stream_handle()
    kds = {(s, pr): B.kernel_desc([s], -1, pr) for s in SIGMAS for pr in ("float32", "float64")}

    def applyK(s, pr, out):
        B.check(B.LIB.lddmm_applyK(kds[(s, pr)], B.ptr(y), N, B.ptr(x), M, 3, B.ptr(a), 3,
                                   B.ptr(out), st), "applyK")

    def applyDiffKT(s, pr, out):
        B.check(B.LIB.lddmm_applyDiffKT(kds[(s, pr)], B.ptr(y), N, B.ptr(x), M, 3, B.ptr(p),
                                        B.ptr(a), 3, B.DIFFKT_PLAIN, 1.0, B.ptr(out), st),
                "applyDiffKT")

    def step(pr="float32"):
        for s in SIGMAS:
            applyK(s, pr, outK)
            applyDiffKT(s, pr, outD)

    if args.probe:
        for pr in ("float32", "float64"):
            for _ in range(3):
                applyK(SIGMAS[0], pr, outK)
            for _ in range(3):
                applyDiffKT(SIGMAS[0], pr, outD)
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device=dev)     # 256 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("probe done: 3 launches each of applyK and applyDiffKT in FP32 then FP64 (N=M=%d), "
            "3 copies of 256 MiB" % N)
        return

    log(f"warmup {args.warmup} steps, N=M={N}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    B.profile_enable(True)
    B.profile_read()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_each = B.profile_read()
    B.profile_enable(False)
    elapsed = t1 - t0
    pairs_per_step = 2 * len(SIGMAS) * N * M
    value = args.steps * pairs_per_step / elapsed
    log(f"timed {args.steps} steps: {elapsed * 1e3 / args.steps:.2f} ms/step, {value:.4g} pair-evals/s")

    # launch order per step: K(s0), D(s0), K(s1), D(s1), K(s2), D(s2)
    per_step = 2 * len(SIGMAS)
    kt = [t for i, t in enumerate(ms_each) if (i % per_step) % 2 == 0]
    dt = [t for i, t in enumerate(ms_each) if (i % per_step) % 2 == 1]
    avgK = sum(kt) / max(len(kt), 1)
    avgD = sum(dt) / max(len(dt), 1)
    achievedD = N * M * FLOP_PER_PAIR["applyDiffKT"] / (avgD * 1e-3) / 1e12
    achievedK = N * M * FLOP_PER_PAIR["applyK"] / (avgK * 1e-3) / 1e12
    log(f"pair kernels: applyK {avgK:.3f} ms ({achievedK:.1f} TFLOP/s), "
        f"applyDiffKT {avgD:.3f} ms ({achievedD:.1f} TFLOP/s), {len(ms_each)} launches")

    extra = {
        "pair_kernel_ms": {"applyK": round(avgK, 4), "applyDiffKT": round(avgD, 4)},
        "pair_evals_per_s_kernel": {"applyK": N * M / (avgK * 1e-3), "applyDiffKT": N * M / (avgD * 1e-3)},
        "exp_per_s": N * M / (avgK * 1e-3),
        "exp_frac_of_peak": N * M / (avgK * 1e-3) / PEAK_EXP_PER_S,
    }

    # A/B: the same step with MFMA accumulation (applyK, plain applyDiffKT)
    B.set_pair_engine('mfma')
    step()
    torch.cuda.synchronize()
    B.profile_enable(True)
    B.profile_read()
    t0 = time.perf_counter()
    step()
    torch.cuda.synchronize()
    t_mfma = time.perf_counter() - t0
    ms_m = B.profile_read()
    B.profile_enable(False)
    B.set_pair_engine('valu')
    extra["mfma_engine"] = {"ms_per_step": t_mfma * 1e3,
                            "pair_kernel_ms": {"applyK": round(sum(ms_m[0::2]) / max(len(ms_m[0::2]), 1), 4),
                                               "applyDiffKT": round(sum(ms_m[1::2]) / max(len(ms_m[1::2]), 1), 4)}}
    log(f"mfma engine: {t_mfma * 1e3:.2f} ms/step, kernels {extra['mfma_engine']['pair_kernel_ms']}")

    if not args.no_fp64:
        # FP64 plain path: rate and the FP32 error against it (norm-wise relative)
        errs = {}
        ref = torch.empty_like(outK)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s in SIGMAS:
            applyK(s, "float64", ref)
            applyK(s, "float32", outK)
            torch.cuda.synchronize()
            errs[f"applyK_sigma{s}"] = float((outK - ref).norm() / ref.norm())
            applyDiffKT(s, "float64", ref)
            applyDiffKT(s, "float32", outD)
            torch.cuda.synchronize()
            errs[f"applyDiffKT_sigma{s}"] = float((outD - ref).norm() / ref.norm())
        B.profile_enable(True)
        B.profile_read()
        t0 = time.perf_counter()
        applyK(SIGMAS[0], "float64", ref)
        applyDiffKT(SIGMAS[0], "float64", ref)
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t0
        ms64 = B.profile_read()
        B.profile_enable(False)
        extra["fp64_pair_evals_per_s"] = 2 * N * M / t64
        extra["fp64_pair_kernel_ms"] = [round(v, 3) for v in ms64]
        if len(ms64) == 2:
            tf64 = {op: N * M * FLOP_PER_PAIR[op] / (ms * 1e-3) / 1e12
                    for op, ms in zip(("applyK", "applyDiffKT"), ms64)}
            extra["fp64_tflops"] = {op: round(v, 2) for op, v in tf64.items()}
            extra["fp64_frac_of_peak"] = {op: round(v / PEAK_FP64_TFLOPS, 3) for op, v in tf64.items()}
        extra["fp32_rel_err_vs_fp64"] = errs
        log(f"fp64: {2 * N * M / t64:.4g} pair-evals/s; max fp32 rel err {max(errs.values()):.2e}")

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded uniform points in [0,50]^3, N(0,1) weights)",
        "config": {
            "workload": ("kernel-sum microbenchmark: applyK + applyDiffKT (plain), N=M=50,000 uniform "
                         "in [0,50]^3, sigma in {2.5,5,10} single-scale, compensated FP32"),
            "n": N, "m": M, "sigmas": list(SIGMAS), "pairs_per_step": pairs_per_step,
        },
        "roofline": {
            "bound": "valu",
            "engine": "FP32 VALU + v_exp_f32 (compute-bound; packed FP32 disabled)",
            "kernel": "k_pair<float, OpDiffKT<3,false,false>, 2> (applyDiffKT plain)",
            "achieved": achievedD, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achievedD / PEAK_FP32_TFLOPS, "traffic": measured_traffic(N),
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
            "flop_per_pair": FLOP_PER_PAIR["applyDiffKT"], "pairs_per_launch": N * M,
            "avg_launch_ms": avgD,
        },
        "extra": extra,
    }
    if not args.no_extras:
        extra["varifold_100k"] = varifold_microbench(B, torch, synth, np)
        extra["registration_plain_cfg1"] = registration_cfg1(torch)
        extra["population_cfg4"] = population_cfg4(torch)
        extra["precision_cfg3"] = precision_cfg3(torch)
        extra["learned_warmstart_cfg5"] = warmstart_cfg5(torch)
        extra["second_order_registration"] = second_order_registration(torch)
        extra["population_registration"] = population_registration(torch)

    if not args.no_cpu:
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline(y_h, x_h, a_h, p_h)
        log(f"cpu baseline {result['cpu_baseline']['value']:.4g} pair-evals/s "
            f"on {result['cpu_baseline']['cores']} threads")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
