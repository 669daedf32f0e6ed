"""Committed golden vectors (tests/golden/oracle_small.npz, written by
tests/golden/make_golden.py from the FP64 oracle) for the hot path: kernel
sums in every mode with one and two scales, the surface varifold / current
terms and gradients, the first-order matching objective and gradient, and
geodesic shooting.

* CPU: the oracle still reproduces the committed vectors (rtol 1e-12), so an
  edit of the oracle that changes a result is caught even where the oracle's
  own pins would not notice.
* GPU: the HIP path reproduces them, FP64 within 1e-11 relative (norm-wise)
  and the compensated FP32 form within 1e-5.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "oracle_small.npz"))


def gin(k):
    return G["in_" + k]


def gout(k):
    return G["out_" + k]


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


SIGS = {"s1": G["sig1"], "s2": G["sig2"]}
SD, SV, T = float(G["sig_diff"]), float(G["sig_dist"]), int(G["T"])


def kernel_sum_cases():
    y, x, a, p = gin("y"), gin("x"), gin("a"), gin("p")
    xs, ps, as_ = gin("xs"), gin("ps"), gin("as")
    for tag in ("s1", "s2"):
        yield f"applyK_{tag}", ("K", y, x, a, SIGS[tag], {})
        yield f"applyK_k1_{tag}", ("K", y, x, a[:, :1], SIGS[tag], {})
        yield f"diffKT_plain_{tag}", ("D", y, x, p, a, SIGS[tag], {})
        yield f"diffKT_sym_{tag}", ("D", xs, xs, ps, as_, SIGS[tag], {"sym": True})
        yield f"diffKT_lddmm_{tag}", ("D", xs, xs, ps, as_, SIGS[tag], {"lddmm": True, "regweight": 0.6})


def test_oracle_reproduces_golden():
    for name, case in kernel_sum_cases():
        if case[0] == "K":
            _, y, x, a, s, kw = case
            got = O.applyK(y, x, a, s)
        else:
            _, y, x, p, a, s, kw = case
            got = O.applyDiffKT(y, x, p, a, s, **kw)
        np.testing.assert_allclose(got, gout(name), rtol=1e-12, atol=1e-14, err_msg=name)
    fv1, fv2 = (gin("v"), gin("faces")), (gin("v2"), gin("faces"))
    np.testing.assert_allclose(O.varifoldNorm0(fv2, [SV]), gout("varifold_norm0"), rtol=1e-12)
    np.testing.assert_allclose(O.varifoldNormDef(fv1, fv2, [SV]), gout("varifold_normdef"), rtol=1e-12)
    np.testing.assert_allclose(O.varifoldNormGradient(fv1, fv2, [SV]), gout("varifold_grad"), rtol=1e-11,
                               atol=1e-13)
    np.testing.assert_allclose(O.currentNormGradient(fv1, fv2, [SV]), gout("current_grad"), rtol=1e-11,
                               atol=1e-13)
    obj, _ = O.matching_objective(gin("v"), gin("faces"), gin("at"), fv2, [SD], [SV])
    np.testing.assert_allclose(obj, gout("matching_objective"), rtol=1e-12)
    np.testing.assert_allclose(O.matching_gradient(gin("v"), gin("faces"), gin("at"), fv2, [SD], [SV]),
                               gout("matching_gradient"), rtol=1e-11, atol=1e-13)
    xt, at = O.landmarkEPDiff(T, gin("v"), gin("a0"), [SD])
    np.testing.assert_allclose(xt, gout("epdiff_xt"), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(at, gout("epdiff_at"), rtol=1e-12, atol=1e-13)


@pytest.fixture(scope="module")
def env():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lddmm_amd import kernelFunctions_util as ku, surfaceDistances as sd, pointEvolution as evol
    from lddmm_amd.surfaces import Surface
    from lddmm_amd.kernelFunctions import Kernel
    from lddmm_amd.surfaceMatching import SurfaceMatching
    return dict(torch=torch, ku=ku, sd=sd, evol=evol, Surface=Surface, Kernel=Kernel, SM=SurfaceMatching)


TOL = {"float64": 1e-11, "float32": 1e-5}


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_hip_kernel_sums_match_golden(env, dtype):
    ku = env["ku"]
    for name, case in kernel_sum_cases():
        if case[0] == "K":
            _, y, x, a, s, kw = case
            got = ku.applyK(y, x, a, 'gauss', list(s), 10, dtype=dtype)
        else:
            _, y, x, p, a, s, kw = case
            got = ku.applyDiffKT(y, x, p, a, 'gauss', list(s), 10, dtype=dtype, **kw)
        assert rel(got, gout(name)) < TOL[dtype], (name, dtype, rel(got, gout(name)))


@pytest.mark.gpu
def test_hip_surface_terms_match_golden(env):
    S, K, sd = env["Surface"], env["Kernel"], env["sd"]
    f = gin("faces")
    fv1, fv2 = S((f, gin("v"))), S((f, gin("v2")))
    Kd = K(name='gauss', sigma=SV)
    assert abs(sd.varifoldNorm0(fv2, Kd) - gout("varifold_norm0")) <= 1e-11 * abs(gout("varifold_norm0"))
    assert abs(sd.varifoldNormDef(fv1, fv2, Kd) - gout("varifold_normdef")) <= 1e-11 * abs(gout("varifold_normdef"))
    assert rel(sd.varifoldNormGradient(fv1, fv2, Kd), gout("varifold_grad")) < 1e-11
    assert rel(sd.currentNormGradient(fv1, fv2, Kd), gout("current_grad")) < 1e-11


@pytest.mark.gpu
def test_hip_matching_and_geodesic_match_golden(env):
    torch, S, K = env["torch"], env["Surface"], env["Kernel"]
    f = gin("faces")
    opts = {'KparDiff': K(name='gauss', sigma=SD), 'KparDist': K(name='gauss', sigma=SV),
            'errorType': 'varifold', 'sigmaError': 1.0, 'timeStep': 1.0 / T, 'pk_dtype': 'float64',
            'regWeight': 1.0, 'mode': 'quiet'}
    sm = env["SM"](Template=S((f, gin("v"))), Target=S((f, gin("v2"))), options=opts)
    sm.control['at'] = torch.tensor(gin("at"), device='cuda')
    sm.obj = None
    obj = sm.objectiveFun()
    assert abs(obj - gout("matching_objective")) <= 1e-11 * abs(gout("matching_objective"))
    assert rel(sm.getGradient()['at'].cpu().numpy(), gout("matching_gradient")) < 1e-11
    xt, at = env["evol"].landmarkEPDiff(T, gin("v"), gin("a0"), K(name='gauss', sigma=SD))
    xt = xt.cpu().numpy() if hasattr(xt, "cpu") else xt
    at = at.cpu().numpy() if hasattr(at, "cpu") else at
    assert rel(xt, gout("epdiff_xt")) < 1e-11
    assert rel(at, gout("epdiff_at")) < 1e-11
