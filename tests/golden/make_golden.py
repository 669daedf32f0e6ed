"""Writes tests/golden/oracle_small.npz: small seeded inputs and the FP64
oracle's outputs for the hot path (kernel sums in every mode, the surface
varifold and current terms with gradients, the first-order matching
objective and gradient, geodesic shooting).  The fixture pins the HIP path
and the oracle itself between sessions (tests/test_golden.py).

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "py-lddmm_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

from oracle import oracle as O  # noqa: E402
from lddmm_amd import synth  # noqa: E402

OUT = os.path.join(HERE, "oracle_small.npz")

SIG1 = np.array([1.3])
SIG2 = np.array([0.8, 2.5])
SIG_DIFF, SIG_DIST, T = 1.5, 0.8, 5


def inputs():
    r = np.random.default_rng(20261204)
    d = {}
    d["y"], d["x"] = r.normal(size=(37, 3)) * 2, r.normal(size=(53, 3)) * 2
    d["a"], d["p"] = r.normal(size=(53, 3)), r.normal(size=(37, 3))
    d["xs"], d["ps"], d["as"] = r.normal(size=(41, 3)) * 2, r.normal(size=(41, 3)), r.normal(size=(41, 3))
    (v, f), (v2, _) = synth.plain_matching_pair(seed=3, subdiv=1, radius=3.0, nbumps=5, width=1.0)
    d["v"], d["faces"], d["v2"] = v, f.astype(np.int64), v2
    d["at"] = r.normal(size=(T, v.shape[0], 3)) * 0.2
    d["a0"] = r.normal(size=(v.shape[0], 3)) * 0.3
    return d


def outputs(d):
    o = {}
    for tag, sig in (("s1", SIG1), ("s2", SIG2)):
        o[f"applyK_{tag}"] = O.applyK(d["y"], d["x"], d["a"], sig)
        o[f"applyK_k1_{tag}"] = O.applyK(d["y"], d["x"], d["a"][:, :1], sig)
        o[f"diffKT_plain_{tag}"] = O.applyDiffKT(d["y"], d["x"], d["p"], d["a"], sig)
        o[f"diffKT_sym_{tag}"] = O.applyDiffKT(d["xs"], d["xs"], d["ps"], d["as"], sig, sym=True)
        o[f"diffKT_lddmm_{tag}"] = O.applyDiffKT(d["xs"], d["xs"], d["ps"], d["as"], sig, lddmm=True,
                                                 regweight=0.6)
    fv1, fv2 = (d["v"], d["faces"]), (d["v2"], d["faces"])
    o["varifold_norm0"] = np.array(O.varifoldNorm0(fv2, [SIG_DIST]))
    o["varifold_normdef"] = np.array(O.varifoldNormDef(fv1, fv2, [SIG_DIST]))
    o["varifold_grad"] = O.varifoldNormGradient(fv1, fv2, [SIG_DIST])
    o["current_grad"] = O.currentNormGradient(fv1, fv2, [SIG_DIST])
    obj = O.matching_objective(d["v"], d["faces"], d["at"], fv2, [SIG_DIFF], [SIG_DIST])
    o["matching_objective"] = np.array(obj[0] if isinstance(obj, tuple) else obj)
    o["matching_gradient"] = O.matching_gradient(d["v"], d["faces"], d["at"], fv2, [SIG_DIFF], [SIG_DIST])
    xt, at = O.landmarkEPDiff(T, d["v"], d["a0"], [SIG_DIFF])
    o["epdiff_xt"], o["epdiff_at"] = xt, at
    return o


def main():
    d = inputs()
    o = outputs(d)
    np.savez_compressed(OUT, **{"in_" + k: v for k, v in d.items()}, **{"out_" + k: v for k, v in o.items()},
                        sig1=SIG1, sig2=SIG2, sig_diff=SIG_DIFF, sig_dist=SIG_DIST, T=T)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
