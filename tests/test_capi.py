"""C ABI checks that need no GPU: the library loads, exports every entry point
declared in include/lddmm_amd.h, and rejects bad arguments before any device
work (those paths return before a HIP call)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lddmm_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lddmm_[a-zA-Z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("lddmm_applyK", "lddmm_applyDiffKT", "lddmm_applyktensor", "lddmm_surface_varifold",
              "lddmm_shoot_forward", "lddmm_shoot_adjoint", "lddmm_epdiff_forward",
              "lddmm_epdiff_adjoint"):
        assert s in syms


def test_library_exports_every_declared_symbol(lib):
    missing = [s for s in declared_symbols() if not hasattr(lib.LIB, s)]
    assert not missing, missing


def test_python_signature_table_matches_header(lib):
    declared = set(declared_symbols())
    for name in lib.SIGNATURES:
        assert name in declared, name


def test_bad_arguments_are_rejected(lib):
    B = lib
    null = ctypes.c_void_p(0)
    # null kernel description
    rc = B.LIB.lddmm_applyK(None, null, 1, null, 1, 3, null, 3, null, null)
    assert rc == B.EINVAL
    assert b"kernel" in B.LIB.lddmm_last_error()
    kd = B.kernel_desc([1.0])
    kd.nscale = 0
    assert B.LIB.lddmm_applyK(kd, null, 1, null, 1, 3, null, 3, null, null) == B.EINVAL
    kd = B.kernel_desc([1.0])
    kd.kind = 7
    assert B.LIB.lddmm_applyK(kd, null, 1, null, 1, 3, null, 3, null, null) == B.EUNSUPPORTED
    kd = B.kernel_desc([1.0])
    assert B.LIB.lddmm_applyK(kd, null, 1, null, 1, 5, null, 3, null, null) == B.EINVAL   # dim
    assert B.LIB.lddmm_applyK(kd, null, -1, null, 1, 3, null, 3, null, null) == B.EINVAL  # n < 0
    assert B.LIB.lddmm_applyDiffKT(kd, null, 4, null, 5, 3, null, null, 3, B.DIFFKT_LDDMM, 1.0,
                                   null, null) == B.EINVAL   # lddmm needs n == m
    assert B.LIB.lddmm_applyktensor(kd, null, 4, null, 5, 3, null, null, null, null, 9, null,
                                    null) == B.EUNSUPPORTED
    kd.sigma[0] = -1.0
    assert B.LIB.lddmm_applyK(kd, null, 1, null, 1, 3, null, 3, null, null) == B.EINVAL
    # n == 0 is a no-op
    kd = B.kernel_desc([1.0])
    assert B.LIB.lddmm_applyK(kd, null, 0, null, 0, 3, null, 3, null, null) == B.OK


def test_adjacency_builder(lib):
    B = lib
    f = np.array([[0, 1, 2], [0, 2, 3], [3, 2, 1]], dtype=np.int32)
    ptr = np.zeros(5, dtype=np.int32)
    idx = np.zeros(9, dtype=np.int32)
    rc = B.LIB.lddmm_surface_adjacency(f.ctypes.data, 3, 4, ptr.ctypes.data, idx.ctypes.data)
    assert rc == B.OK
    assert ptr.tolist() == [0] + np.cumsum(np.bincount(f.ravel(), minlength=4)).tolist()
    for v in range(4):
        got = sorted((e // 3, e % 3) for e in idx[ptr[v]:ptr[v + 1]])
        want = sorted((fi, c) for fi in range(3) for c in range(3) if f[fi, c] == v)
        assert got == want
    bad = np.array([[0, 1, 9]], dtype=np.int32)
    assert B.LIB.lddmm_surface_adjacency(bad.ctypes.data, 1, 4, ptr.ctypes.data, idx.ctypes.data) == B.EINVAL


def test_curve_adjacency_and_validation(lib):
    B = lib
    s = np.array([[0, 1], [1, 2], [2, 0], [2, 3]], dtype=np.int32)
    ptr = np.zeros(5, dtype=np.int32)
    idx = np.zeros(8, dtype=np.int32)
    assert B.LIB.lddmm_curve_adjacency(s.ctypes.data, 4, 4, ptr.ctypes.data, idx.ctypes.data) == B.OK
    for v in range(4):
        got = sorted((e // 2, e % 2) for e in idx[ptr[v]:ptr[v + 1]])
        assert got == sorted((si, c) for si in range(4) for c in range(2) if s[si, c] == v)
    bad = np.array([[0, 4]], dtype=np.int32)
    assert B.LIB.lddmm_curve_adjacency(bad.ctypes.data, 1, 4, ptr.ctypes.data, idx.ctypes.data) == B.EINVAL
    # argument checks return before any device work
    kd = B.kernel_desc([1.0])
    c = B.LddmmCurve()
    c.nv, c.ns, c.dim = 4, 4, 4
    e = ctypes.c_void_p(1)
    assert B.LIB.lddmm_curve_varifold(kd, ctypes.byref(c), None, 1.0, e, None, None) == B.EINVAL   # dim
    c.dim = 2
    assert B.LIB.lddmm_curve_varifold(kd, ctypes.byref(c), None, 1.0, e, None, None) == B.EINVAL   # nulls
    c.vertices, c.segments = 8, 8        # never dereferenced: the weight check comes first
    assert B.LIB.lddmm_curve_varifold(kd, ctypes.byref(c), None, -1.0, e, None, None) == B.EINVAL
    assert b"weight" in B.LIB.lddmm_last_error()


def test_population_create_validation(lib):
    """Descriptor checks run before any device call."""
    B = lib
    h = ctypes.c_void_p()
    d = B.LddmmPopulationDesc()
    d.npairs = 0
    assert B.LIB.lddmm_population_create(ctypes.byref(d), ctypes.byref(h)) == B.EINVAL
    assert b"npairs" in B.LIB.lddmm_last_error()
    voff = np.array([0, 5, 3], dtype=np.int64)
    foff = np.array([0, 6, 8], dtype=np.int64)
    d.npairs = 2
    d.voff, d.foff, d.tfoff = voff.ctypes.data, foff.ctypes.data, foff.ctypes.data
    assert B.LIB.lddmm_population_create(ctypes.byref(d), ctypes.byref(h)) == B.EINVAL
    assert b"non-decreasing" in B.LIB.lddmm_last_error()
    assert B.LIB.lddmm_population_create(ctypes.byref(d), None) == B.EINVAL
    assert B.LIB.lddmm_population_destroy(None) == B.OK


def test_kernel_desc_mirrors_kernelspec(lib):
    kd = lib.kernel_desc([2.5, 5.0], kp=-1, prec='float32')
    assert kd.nscale == 2 and kd.sigma[1] == 5.0 and kd.prec == lib.PREC_F32C
    with pytest.raises(ValueError):
        lib.kernel_desc([1.0], prec='float16')
    assert lib.kernel_kind('laplacian') == lib.KERNEL_LAPLACIAN
    assert lib.kernel_kind('gaussian') == lib.KERNEL_GAUSS
    with pytest.raises(NotImplementedError):
        lib.kernel_kind('poly')
    kd = lib.kernel_for('laplacian', [1.0], 4)
    assert kd.kind == lib.KERNEL_LAPLACIAN and kd.order == 4
    with pytest.raises(ValueError):
        lib.kernel_for('laplacian', [1.0], 5)


def test_laplacian_descriptor_checks(lib):
    """Order range and the gauss-only data terms are checked before any device call."""
    B = lib
    null = ctypes.c_void_p(0)
    kd = B.kernel_for('laplacian', [1.0], 3)
    kd.order = 7
    assert B.LIB.lddmm_applyK(kd, null, 1, null, 1, 3, null, 3, null, null) == B.EINVAL
    assert b"order" in B.LIB.lddmm_last_error()
    kd.order = -1
    assert B.LIB.lddmm_applyK(kd, null, 1, null, 1, 3, null, 3, null, null) == B.EINVAL
    kd = B.kernel_for('laplacian', [1.0], 3)
    assert B.LIB.lddmm_applyK(kd, null, 0, null, 0, 3, null, 3, null, null) == B.OK
    s = B.LddmmSurface()
    e = ctypes.c_void_p(1)
    assert B.LIB.lddmm_surface_varifold(kd, ctypes.byref(s), None, e, None, None) == B.EUNSUPPORTED
    assert b"gauss" in B.LIB.lddmm_last_error()
    c = B.LddmmCurve()
    assert B.LIB.lddmm_curve_varifold(kd, ctypes.byref(c), None, 1.0, e, None, None) == B.EUNSUPPORTED


def test_population_schedule_covers_every_tile_chunk_within_bounds(lib):
    """Host-only: for random populations, active subsets and launch shapes the
    segment table covers every (tile, chunk) of every active pair exactly once
    and every item stays inside its pair's rows and sources (checked by
    lddmm_population_plan_check against the bounds the kernels index)."""
    B = lib
    rng = np.random.default_rng(5)
    for trial in range(40):
        P = int(rng.integers(1, 40))
        nv = rng.integers(0, 3000, size=P)
        nv[rng.random(P) < 0.1] = 0
        nf = np.where(nv > 0, 2 * nv - 4, 0).clip(min=0)
        ntf = np.where(rng.random(P) < 0.1, 0, rng.integers(1, 6000, size=P))
        voff = np.concatenate([[0], np.cumsum(nv)]).astype(np.int64)
        foff = np.concatenate([[0], np.cumsum(nf)]).astype(np.int64)
        tfoff = np.concatenate([[0], np.cumsum(ntf)]).astype(np.int64)
        act = (rng.random(P) < 0.7).astype(np.uint8)
        for fam in (0, 1):
            for R in (1, 2):
                for f10 in (10, 20, 40):
                    G = int(rng.choice([1, 7, 256, 1024]))
                    n = ctypes.c_int64()
                    rc = B.LIB.lddmm_population_plan_check(P, voff.ctypes.data, foff.ctypes.data,
                                                           tfoff.ctypes.data, act.ctypes.data, fam, R, f10,
                                                           G, ctypes.byref(n))
                    assert rc == B.OK, (trial, fam, R, f10, G, B.LIB.lddmm_last_error())
                    assert n.value >= 0


def test_affine_descriptor_checks(lib):
    """The affine part is accepted by the first-order and geodesic shooting
    entry points only; kinds and weights are checked before any device call."""
    B = lib
    null = ctypes.c_void_p(0)
    kd = B.kernel_desc([1.0])
    kd.affine = B.AFFINE_EUCLIDEAN
    kd.w1, kd.w2 = 0.5, 0.1
    assert B.LIB.lddmm_applyK(kd, null, 1, null, 1, 3, null, 3, null, null) == B.EUNSUPPORTED
    assert b"affine" in B.LIB.lddmm_last_error()
    assert B.LIB.lddmm_epdiff_forward(kd, 3, 0, 3, null, null, null) == B.OK                 # n == 0
    assert B.LIB.lddmm_epdiff_forward(kd, 3, 4, 3, null, null, null) == B.EINVAL             # null arrays
    assert B.LIB.lddmm_epdiff_adjoint(kd, 3, 4, 3, null, null, null, null, null, null, null, null) == B.EINVAL
    assert B.LIB.lddmm_applyDiffKT(kd, null, 1, null, 1, 3, null, null, 3, 0, 1.0, null, null) == B.EUNSUPPORTED
    assert B.LIB.lddmm_shoot_forward(kd, null, 3, 0, 3, null, 1.0, None, null) == B.OK      # n == 0
    kd.affine = 5
    assert B.LIB.lddmm_shoot_forward(kd, null, 3, 4, 3, null, 1.0, None, null) == B.EINVAL
    kd.affine = B.AFFINE_FULL
    kd.w1 = float('nan')
    assert B.LIB.lddmm_shoot_forward(kd, null, 3, 4, 3, null, 1.0, None, null) == B.EINVAL
    from lddmm_amd.kernelFunctions import Kernel
    K = Kernel(name='gauss', sigma=1.0)
    K.setAffine('euclidean', w1=2.0, w2=0.5, center=[1.0, 2.0, 3.0])
    kd = B.kernel_of(K, with_affine=True)
    assert kd.affine == B.AFFINE_EUCLIDEAN and kd.w1 == 2.0 and list(kd.center) == [1.0, 2.0, 3.0]
    with pytest.raises(NotImplementedError):
        B.kernel_of(K)


def test_internal_cost_descriptor_checks(lib):
    B = lib
    s = B.LddmmSurface()
    e = ctypes.c_void_p(1)
    d = B.internal_cost_desc([('elastic', 50.0), ('h1', 1.0)])
    assert d.ncost == 2 and list(d.kinds)[:2] == [1, 0] and d.weights[0] == 50.0
    d.kinds[1] = 9
    assert B.LIB.lddmm_surface_internal_cost(ctypes.byref(s), None, ctypes.byref(d), e, None, None, None) == B.EINVAL
    assert b"kind" in B.LIB.lddmm_last_error()
    d.ncost = 0
    assert B.LIB.lddmm_surface_internal_cost(ctypes.byref(s), None, ctypes.byref(d), e, None, None, None) == B.EINVAL
    d = B.internal_cost_desc([('displacement', 1.0)])
    s.nv, s.nf = 4, 2                      # no arrays: rejected before any device work
    assert B.LIB.lddmm_surface_internal_cost(ctypes.byref(s), None, ctypes.byref(d), e, None, None, None) == B.EINVAL
    kd = B.kernel_desc([1.0])
    g = ctypes.c_void_p(1)
    assert B.LIB.lddmm_shoot_adjoint_ic(kd, ctypes.byref(s), ctypes.byref(d), g, g, 3, g, 1.0, 1, None, g,
                                        None) == B.EINVAL            # adjacency required
    with pytest.raises(NotImplementedError):
        B.internal_cost_desc([('bending', 1.0)])


def test_exp2_polynomial_matches_its_generator():
    """The float64 2^f coefficients in common.hpp are the ones exp2_coeffs.py
    fits, and their Horner evaluation stays within 5e-16 relative on [-1/2, 1/2]."""
    import importlib.util
    here = os.path.join(ROOT, "py-lddmm_amd", "csrc")
    spec = importlib.util.spec_from_file_location("exp2_coeffs", os.path.join(here, "exp2_coeffs.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    cs, _ = gen.coefficients()
    src = open(os.path.join(here, "common.hpp")).read()
    block = src[src.index("kExp2Poly[11] = {"):]
    block = block[block.index("{") + 1:block.index("}")]
    in_header = [float(t) for t in block.replace("\n", " ").split(",") if t.strip()]
    assert in_header == cs
    assert gen.worst_relative_error(cs, points=2001) < 5e-16
