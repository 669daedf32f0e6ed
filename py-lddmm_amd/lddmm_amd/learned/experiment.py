"""End-to-end learned warm-start experiment (run on the GPU):

    python -m lddmm_amd.learned.experiment --batched 256 --train 20000 --epochs 30 --batch 32 \
        --test 1000 --out momentum_net.npz --report report.json
    python -m lddmm_amd.learned.experiment --net momentum_net.npz --test 1000   # evaluate only

1. plain zero-momentum registrations of seeded deformed spheres (training
   seeds 10000.., held-out seeds 900000..) record the features and targets;
   with --batched they run that many pairs at a time through the lockstep
   population registration (features 'normal', target 'vlocal');
2. MomentumNet is trained on them with the package's HIP kernels, one
   network per feature:target variant (default: ray-cast normal heights ->
   time-mean velocity in the local vertex frames, control by a truncated
   conjugate-gradient solve);
3. on held-out pairs, the plain run's final varifold data term E* is the
   target; iterations / seconds to reach E* (1e-4 relative) are compared for
   plain (from its own per-iteration trace), warm-started registrations and
   the same pipeline without the network (raw ray-cast heights).
Progress lines go to stderr; the report is JSON.
"""
import argparse
import json
import logging
import sys
import time

import numpy as np
import torch

from . import warmstart as W


def log(msg):
    print(f"[warmstart] {msg}", file=sys.stderr, flush=True)


def plain_with_trace(seed, cfg, with_features=True):
    sm = W.make_matching(seed, cfg)
    feats = {f"feat_{k}": W.features(sm, k).cpu().numpy() for k in W.FEATURES} if with_features else {}
    sm.setTargetDataEnergy(None)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sm.optimizeMatching()
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    return sm, feats, sec


def iterations_to(trace, target, rtol=1e-4):
    for i, d in enumerate(trace):
        if d <= target * (1 + rtol):
            return i + 1
    return len(trace)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--test", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default="momentum_net.npz")
    ap.add_argument("--report", default=None)
    ap.add_argument("--time-budget", type=float, default=600.0, help="seconds for data generation")
    ap.add_argument("--net", default=None,
                    help="evaluate this saved network (no data generation or training)")
    ap.add_argument("--feature-only-pairs", type=int, default=None,
                    help="held-out pairs that also run the feature-only baseline (default: all)")
    ap.add_argument("--refined", action="store_true",
                    help="train on refined-velocity residuals (target 'vresid', no registrations run)")
    ap.add_argument("--refine-pairs", type=int, default=0,
                    help="held-out pairs that also run the re-ray-cast velocity without a network "
                         "(one and REFINE_STEPS refinements)")
    ap.add_argument("--batched", type=int, default=0,
                    help="generate the training registrations this many pairs at a time with the lockstep "
                         "population registration (features 'normal', target 'vlocal' only)")
    ap.add_argument("--net-refine", type=int, default=None,
                    help="re-ray-casts after the network's prediction (target 'vresid'; default "
                         "NET_REFINE_STEPS)")
    ap.add_argument("--cg-iters", type=int, default=None,
                    help="CG products of the velocity -> control solve (default: the network's own)")
    ap.add_argument("--oracle-check", type=int, default=5,
                    help="held-out pairs also warm-started from the plain run's own a0")
    ap.add_argument("--variants", default="normal:vlocal",
                    help="comma-separated feature:target[:loss] (features normal/force/meanshift/kforce; "
                         "targets vlocal/abar/vbar/momentum/velocity; loss mse (default) or kvel)")
    a = ap.parse_args(argv)
    logging.getLogger().setLevel(logging.WARNING)
    cfg = dict(W.DEFAULT)
    rep = {"config": cfg, "train_requested": a.train, "test_requested": a.test}

    t0 = time.perf_counter()
    samples = []
    if a.net:
        from .model import MomentumNet
        a.train = 0
        a.variants = ""
    if a.refined:
        if ":vresid" not in a.variants:
            a.variants = "normal:vresid"
        if a.train:
            samples = W.generate_refined([10000 + k for k in range(a.train)], cfg, progress=log)
        a.train = 0
    elif a.batched and a.train:
        samples = W.generate_batched([10000 + k for k in range(a.train)], cfg, batch=a.batched, progress=log)
    for k in range(a.train if not (a.batched or a.refined) else 0):
        seed = 10000 + k
        sm, feats, sec = plain_with_trace(seed, cfg)
        smp = dict(feats)
        smp.update(W.targets_of(sm, sm.control['at']))
        samples.append(smp)
        if k % 100 == 0:
            log(f"train registration {k}: {sec:.2f} s, data {sm.varifold_data():.4f}")
        if time.perf_counter() - t0 > a.time_budget:
            log(f"data-generation budget reached after {k + 1} registrations")
            break
    rep["train_pairs"] = len(samples)
    rep["generation_seconds"] = time.perf_counter() - t0

    nets = {}
    if a.net:
        net = MomentumNet.load(a.net)
        nets[f"{net.features}:{net.target}"] = net
        a.variants = f"{net.features}:{net.target}"
    for variant in (a.variants.split(",") if not a.net else []):
        parts = variant.split(":")
        feat_kind, target = parts[0], parts[1]
        loss_kind = parts[2] if len(parts) > 2 else 'mse'
        t1 = time.perf_counter()
        net, losses = W.train(samples, epochs=a.epochs, batch=a.batch, lr=a.lr, log_every=10, target=target,
                              features=feat_kind, progress=lambda m: log(f"[{variant}] {m}"), loss=loss_kind)
        target = variant
        torch.cuda.synchronize()
        rep[f"train_seconds_{target}"] = time.perf_counter() - t1
        rep[f"loss_first_last_{target}"] = [losses[0], losses[-1]]
        log(f"[{target}] trained on {len(samples)} pairs in {rep[f'train_seconds_{target}']:.1f} s, "
            f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
        net.save(a.out.replace(".npz", "_" + variant.replace(":", "_") + ".npz"))
        nets[target] = net

    rows = []
    for k in range(a.test):
        seed = 900000 + k
        sm, _, sec_plain = plain_with_trace(seed, cfg, with_features=False)
        E = sm.varifold_data()
        obj_plain = float(sm.objectiveFun())
        it_plain = iterations_to(sm.data_trace, E)
        sec_plain_to = sec_plain * it_plain / max(len(sm.data_trace), 1)
        row = {"seed": seed, "E_plain": E, "it_plain": it_plain, "it_plain_total": len(sm.data_trace),
               "sec_plain_total": sec_plain, "sec_plain_to_E": sec_plain_to}
        if k < 3:
            row["plain_trace"] = list(sm.data_trace)
        for target, net in nets.items():
            r = W.register(seed, net=net, target_data=E, cfg=cfg, cg_iters=a.cg_iters, refine=a.net_refine)
            row.update({f"it_warm_{target}": r["iterations"], f"sec_warm_{target}": r["seconds"],
                        f"E_warm_{target}": r["data"], f"E0_warm_{target}": r["data0"],
                        f"obj_ratio_{target}": r["objective"] / obj_plain,
                        f"reached_{target}": r["reached"], f"used_warm_{target}": r["used_warm"]})
            if net.target == 'vresid':
                smr = W.make_matching(seed, cfg)
                ch, fr = W.ray_cast_heights(smr)
                res = W.refined_velocity(smr, heights=(ch, fr)) - ch[:, :1] * fr[:, 0]
                tg = torch.einsum('nkj,nj->nk', fr, res).cpu().numpy()
                fx = W.height_inputs(smr, net.features, (ch, fr))
            else:
                tg = W.targets_of(sm, sm.control['at'])[W.TARGET_KEY[net.target]]
                fx = W.features(W.make_matching(seed, cfg), net.features)
            pred = net.predict(fx, [fx.shape[0]]).double().cpu().numpy()
            row[f"pred_rel_err_{target}"] = float(np.linalg.norm(pred - tg) / np.linalg.norm(tg))
        if "normal" in a.variants and (a.feature_only_pairs is None or k < a.feature_only_pairs):
            # the same pipeline without the network: the ray-cast normal height
            # itself as the velocity, solved with the same truncated CG
            r = W.register(seed, start=lambda s_: W.feature_only_control(s_)[0], target_data=E, cfg=cfg)
            row.update({"it_feature_only": r["iterations"], "obj_ratio_feature_only": r["objective"] / obj_plain,
                        "sec_feature_only": r["seconds"]})
        if k < a.refine_pairs:
            # the re-ray-cast velocity itself (no network), one and REFINE_STEPS refinements
            for steps in sorted({1, W.REFINE_STEPS}):
                r = W.register(seed, start=lambda s_: W.time_constant(s_, W.solve_control(s_, W.refined_velocity(
                    s_, steps))), target_data=E, cfg=cfg)
                row[f"it_refined{steps}"] = r["iterations"]
        if k < a.oracle_check:
            # upper bounds: the plain run's own time-mean control (perfect predictor),
            # and its velocity K abar turned back into abar by truncated CG, exact
            # and with 10% / 30% relative white noise added to the velocity
            abar = sm.control['at'].mean(0)
            T_ = sm.Tsize
            r = W.register(seed, at=abar[None].expand(T_, -1, -1).contiguous(), target_data=E, cfg=cfg)
            row.update({"it_warm_true_abar": r["iterations"], "E0_true_abar": r["data0"]})
            # perfect prediction through the default pipeline: K abar -> truncated CG
            a_v = sm.options['KparDiff'].solve(sm.x0_dev, W.velocity(sm, abar), iterMax=W.SOLVE_ITERS, sequence='textbook')
            r = W.register(seed, at=a_v[None].expand(T_, -1, -1).contiguous(), target_data=E, cfg=cfg)
            row["it_true_vbar_solve"] = r["iterations"]
            # tolerance of the time-mean control to a smooth error: abar + e with
            # e = K(x0) white noise, scaled so that |K e| = lvl |K abar|
            K_ = sm.options['KparDiff']
            vb = W.velocity(sm, abar)
            gen = torch.Generator(device=vb.device).manual_seed(seed)
            for lvl in (0.02, 0.05, 0.1, 0.2):
                noise = torch.randn(vb.shape, generator=gen, device=vb.device, dtype=vb.dtype)
                e = K_.applyK(sm.x0_dev, noise)
                e = e * (lvl * vb.norm() / K_.applyK(sm.x0_dev, e).norm())
                ah = abar + e
                r = W.register(seed, at=ah[None].expand(T_, -1, -1).contiguous(), target_data=E, cfg=cfg)
                row[f"it_abar_smooth_err{lvl}"] = r["iterations"]
        rows.append(row)
        if k % 10 == 0 or k < a.oracle_check:
            log(f"test {k}: plain {it_plain} it ({sec_plain:.2f} s) -> warm "
                + ", ".join(f"{t}: {row[f'it_warm_{t}']} it ({row[f'sec_warm_{t}']:.2f} s)" for t in nets))
    itp = np.array([r["it_plain"] for r in rows], dtype=float)
    rep["test"] = {"pairs": len(rows), "mean_it_plain": float(itp.mean()),
                   "mean_sec_plain_to_E": float(np.mean([r["sec_plain_to_E"] for r in rows]))}
    for target in nets:
        itw = np.array([r[f"it_warm_{target}"] for r in rows], dtype=float)
        rep["test"][target] = {
            "mean_it_warm": float(itw.mean()),
            "iteration_ratio": float(itp.mean() / max(itw.mean(), 1e-9)),
            "mean_sec_warm": float(np.mean([r[f"sec_warm_{target}"] for r in rows])),
            "reached_fraction": float(np.mean([r[f"reached_{target}"] for r in rows])),
            "warm_used_fraction": float(np.mean([r[f"used_warm_{target}"] for r in rows])),
            "mean_E0_over_E": float(np.mean([r[f"E0_warm_{target}"] / r["E_plain"] for r in rows])),
            "mean_objective_ratio": float(np.mean([r[f"obj_ratio_{target}"] for r in rows])),
        }
        rep["test"][target]["mean_pred_rel_err"] = float(np.mean([r[f"pred_rel_err_{target}"] for r in rows]))
    fo = [r for r in rows if "it_feature_only" in r]
    if fo:
        itf = np.array([r["it_feature_only"] for r in fo], dtype=float)
        itpf = np.array([r["it_plain"] for r in fo], dtype=float)
        rep["test"]["feature_only"] = {"pairs": len(fo), "mean_it": float(itf.mean()),
                                       "iteration_ratio": float(itpf.mean() / max(itf.mean(), 1e-9)),
                                       "mean_objective_ratio": float(np.mean([r["obj_ratio_feature_only"]
                                                                              for r in fo]))}
    rf = [r for r in rows if "it_refined1" in r]
    if rf:
        rep["test"]["refined_no_network"] = {"pairs": len(rf)}
        for key in rf[0]:
            if key.startswith("it_refined"):
                rep["test"]["refined_no_network"]["mean_" + key[3:]] = float(np.mean([r[key] for r in rf]))
    ora = [r for r in rows if "it_warm_true_abar" in r]
    if ora:
        rep["test"]["abar_error_tolerance"] = {k: float(np.mean([r[k] for r in ora])) for k in ora[0]
                                               if k.startswith("it_abar_smooth_err")}
    if ora:
        rep["test"]["true_abar"] = {"pairs": len(ora),
                                  "mean_it": float(np.mean([r["it_warm_true_abar"] for r in ora])),
                                  "mean_it_vbar_solve": float(np.mean([r["it_true_vbar_solve"] for r in ora])),
                                  "mean_E0_over_E": float(np.mean([r["E0_true_abar"] / r["E_plain"] for r in ora]))}
    rep["rows"] = rows
    log(json.dumps(rep["test"]))
    if a.report:
        with open(a.report, "w") as fh:
            json.dump(rep, fh, indent=1)
    print(json.dumps(rep["test"]), flush=True)


if __name__ == "__main__":
    main()
