"""End-to-end learned warm-start experiment (run on the GPU):

    python -m lddmm_amd.learned.experiment --train 2000 --test 100 --epochs 300 \
        --out momentum_net.npz --report report.json

1. plain zero-momentum registrations of seeded deformed spheres (training
   seeds 10000.., held-out seeds 900000..) record (features, a0 = at[0]);
2. MomentumNet is trained on them with the package's HIP kernels;
3. on held-out pairs, the plain run's final varifold data term E* is the
   target; iterations / seconds to reach E* (1e-4 relative) are compared for
   plain (from its own per-iteration trace) and warm-started registrations.
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


def plain_with_trace(seed, cfg):
    sm = W.make_matching(seed, cfg)
    feats = W.features(sm)
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
    a = ap.parse_args(argv)
    logging.getLogger().setLevel(logging.WARNING)
    cfg = dict(W.DEFAULT)
    rep = {"config": cfg, "train_requested": a.train, "test_requested": a.test}

    t0 = time.perf_counter()
    samples = []
    for k in range(a.train):
        seed = 10000 + k
        sm, feats, sec = plain_with_trace(seed, cfg)
        samples.append({"features": feats.cpu().numpy(), "a0": sm.control['at'][0].cpu().numpy()})
        if k % 100 == 0:
            log(f"train registration {k}: {sec:.2f} s, data {sm.varifold_data():.4f}")
        if time.perf_counter() - t0 > a.time_budget:
            log(f"data-generation budget reached after {k + 1} registrations")
            break
    rep["train_pairs"] = len(samples)
    rep["generation_seconds"] = time.perf_counter() - t0

    t1 = time.perf_counter()
    net, losses = W.train(samples, epochs=a.epochs, batch=a.batch, lr=a.lr, log_every=0)
    torch.cuda.synchronize()
    rep["train_seconds"] = time.perf_counter() - t1
    rep["loss_first_last"] = [losses[0], losses[-1]]
    log(f"trained on {len(samples)} pairs in {rep['train_seconds']:.1f} s, loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    net.save(a.out)

    rows = []
    for k in range(a.test):
        seed = 900000 + k
        sm, _, sec_plain = plain_with_trace(seed, cfg)
        E = sm.varifold_data()
        it_plain = iterations_to(sm.data_trace, E)
        sec_plain_to = sec_plain * it_plain / max(len(sm.data_trace), 1)
        r = W.register(seed, net=net, target_data=E, cfg=cfg)
        rows.append({"seed": seed, "E_plain": E, "it_plain": it_plain, "it_plain_total": len(sm.data_trace),
                     "sec_plain_total": sec_plain, "sec_plain_to_E": sec_plain_to, "it_warm": r["iterations"],
                     "sec_warm": r["seconds"], "E_warm": r["data"], "reached": r["reached"],
                     "used_warm": r["used_warm"]})
        if k % 10 == 0:
            log(f"test {k}: plain {it_plain} it ({sec_plain:.2f} s) -> warm {r['iterations']} it "
                f"({r['seconds']:.2f} s), reached {r['reached']}")
    itp = np.array([r["it_plain"] for r in rows], dtype=float)
    itw = np.array([r["it_warm"] for r in rows], dtype=float)
    rep["test"] = {
        "pairs": len(rows),
        "mean_it_plain": float(itp.mean()), "mean_it_warm": float(itw.mean()),
        "iteration_ratio": float(itp.mean() / max(itw.mean(), 1e-9)),
        "mean_sec_plain_to_E": float(np.mean([r["sec_plain_to_E"] for r in rows])),
        "mean_sec_warm": float(np.mean([r["sec_warm"] for r in rows])),
        "reached_fraction": float(np.mean([r["reached"] for r in rows])),
        "warm_used_fraction": float(np.mean([r["used_warm"] for r in rows])),
    }
    rep["rows"] = rows
    log(json.dumps(rep["test"]))
    if a.report:
        with open(a.report, "w") as fh:
            json.dump(rep, fh, indent=1)
    print(json.dumps(rep["test"]), flush=True)


if __name__ == "__main__":
    main()
