"""Varifold energy with the cross terms on the matrix cores
(lddmm_set_varifold_engine: 'mfma_cos' takes the normal cosines from
v_mfma_f32_16x16x4f32, 'mfma_cos_r2' also the squared centre distances)
against the FP64 oracle (|S|^2, base/surfaceDistances.py:286-302, and
|S0|^2 - 2 <S0, S1>, :328-353) and, at 20,000 faces, against the device FP64 path.
Tolerance: the float32 VALU path's 1e-5 relative for the cosine engine; 1e-4
for the distance engine, whose |a|^2 + |b|^2 - 2 a.b form loses bits to
cancellation (taken relative to an origin per block of 16 target faces).
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {"mfma_cos": 1e-5, "mfma_cos_r2": 1e-4}


@pytest.fixture(scope="module")
def mods():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lddmm_amd import _backend as B, synth, surfaceDistances as sd
    from lddmm_amd.surfaces import Surface
    from lddmm_amd.kernelFunctions import Kernel
    return B, synth, sd, Surface, Kernel


def energies(B, sd, K, S0, S1, engine):
    B.set_varifold_engine(engine)
    try:
        return sd.varifoldNorm0(S1, K), sd.varifoldNormDef(S0, S1, K)
    finally:
        B.set_varifold_engine('valu')


@pytest.mark.parametrize("engine", ["mfma_cos", "mfma_cos_r2"])
@pytest.mark.parametrize("subdiv,sig", [(2, 1.0), (3, 1.0), (4, 1.0), (4, 2.5)])
def test_energy_matches_oracle(mods, engine, subdiv, sig):
    B, synth, sd, Surface, Kernel = mods
    (v, f), (v2, _) = synth.plain_matching_pair(seed=5, subdiv=subdiv, radius=10.0, nbumps=8, width=3.0)
    S0, S1 = Surface((f, v)), Surface((f, v2))
    K = Kernel(name='gauss', sigma=sig)
    K.pk_dtype = 'float32'
    e0, eD = energies(B, sd, K, S0, S1, engine)
    r0, rD = O.varifoldNorm0((v2, f), [sig]), O.varifoldNormDef((v, f), (v2, f), [sig])
    print(f"[vf-mfma] {engine} subdiv {subdiv} sigma {sig}: rel err {abs(e0 - r0) / abs(r0):.2e}, "
          f"{abs(eD - rD) / abs(rD):.2e}", flush=True)
    assert abs(e0 - r0) / abs(r0) < TOL[engine]
    assert abs(eD - rD) / abs(rD) < TOL[engine]


@pytest.mark.parametrize("engine", ["mfma_cos", "mfma_cos_r2"])
def test_energy_20k_faces_matches_fp64(mods, engine):
    """Two seeded sphere hulls of 20,000 faces (radius 50, sigma 1; the
    second with radial bumps), ragged against the 128-face target blocks."""
    B, synth, sd, Surface, Kernel = mods
    v1, f1 = synth.sphere_hull(10_002, 50.0, seed=21)
    v2, f2 = synth.sphere_hull(10_002, 50.0, seed=22)
    v2 = synth.bump_deform(v2, f2, 30, 8.0, 2.0, seed=23)
    S0, S1 = Surface((f1, v1)), Surface((f2, v2))
    K64 = Kernel(name='gauss', sigma=1.0)
    K64.pk_dtype = 'float64'
    r0, rD = sd.varifoldNorm0(S1, K64), sd.varifoldNormDef(S0, S1, K64)
    K = Kernel(name='gauss', sigma=1.0)
    K.pk_dtype = 'float32'
    e0, eD = energies(B, sd, K, S0, S1, engine)
    v0, vD = energies(B, sd, K, S0, S1, 'valu')
    print(f"[vf-mfma] {engine} 20k: rel err {abs(e0 - r0) / abs(r0):.2e}, {abs(eD - rD) / abs(rD):.2e} "
          f"(valu {abs(v0 - r0) / abs(r0):.2e}, {abs(vD - rD) / abs(rD):.2e})", flush=True)
    assert abs(e0 - r0) / abs(r0) < TOL[engine]
    assert abs(eD - rD) / abs(rD) < TOL[engine]


def test_engine_switch_rejects_unknown(mods):
    B = mods[0]
    with pytest.raises(Exception):
        B.set_varifold_engine(7)
    B.set_varifold_engine('valu')
