"""Vertex-weight gradients (with_weights=True) of the surface data terms on
the GPU against the FP64 oracle: varifold (base/surfaceDistances.py:455-467),
current (:130-137), current magnitude (:61-67) and measure (:269-276), with
non-uniform vertex weights on both surfaces.  The oracle's weight parts are
pinned by finite differences in the vertex weights (tests/test_oracle.py).
Tolerances: FP64 1e-11, FP32 compensated 2e-5 (norm-wise relative), for the
vertex part and the weight part separately.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {'float64': 1e-11, 'float32': 2e-5}


@pytest.fixture(scope="module")
def env():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lddmm_amd import synth, surfaceDistances as sd
    from lddmm_amd.surfaces import Surface
    from lddmm_amd.kernelFunctions import Kernel
    return dict(synth=synth, sd=sd, Surface=Surface, Kernel=Kernel)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def shapes(env):
    synth = env["synth"]
    v, f = synth.icosphere(3, radius=6.0)
    v2 = synth.bump_deform(v, f, 8, 1.5, 0.5, seed=21)
    rng = np.random.default_rng(4)
    return v, f, rng.uniform(0.5, 1.5, v.shape[0]), v2, f, rng.uniform(0.5, 1.5, v.shape[0])


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("kind", ["varifold", "current", "measure", "magnitude"])
@pytest.mark.parametrize("sigma", [[1.2], [0.8, 2.0]])
def test_weight_gradients(env, shapes, dtype, kind, sigma):
    sd, S = env["sd"], env["Surface"]
    v, f, w, v2, f2, w2 = shapes
    K = env["Kernel"](name='gauss', sigma=sigma)
    K.pk_dtype = dtype
    A, Bt = S((f, v), weights=w), S((f2, v2), weights=w2)
    if kind == "varifold":
        px, pxw = sd.varifoldNormGradient(A, Bt, K, with_weights=True)
        rx, rw = O.varifoldNormGradient((v, f, w), (v2, f2, w2), sigma, with_weights=True)
    elif kind == "current":
        px, pxw = sd.currentNormGradient(A, Bt, K, with_weights=True)
        rx, rw = O.currentNormGradient((v, f, w), (v2, f2, w2), sigma, with_weights=True)
    elif kind == "measure":
        px, pxw = sd.measureNormGradient(A, Bt, K, with_weights=True)
        rx, rw = O.measureNormGradient((v, f, w), (v2, f2, w2), sigma, with_weights=True)
    else:
        px, pxw = sd.currentMagnitudeGradient(A, K, with_weights=True)
        rx, rw = O.currentMagnitudeGradient((v, f, w), sigma, with_weights=True)
    assert px.shape == rx.shape and pxw.shape == rw.shape == (v.shape[0],)
    assert rel(px, rx) < TOL[dtype]
    assert rel(pxw, rw) < TOL[dtype]
