"""Host-side pieces that need no GPU: shape generators, Surface geometry,
legacy-VTK reading / writing."""
import os

import numpy as np
import pytest

from lddmm_amd import synth, vtk_io
from lddmm_amd.surfaces import Surface
from oracle import oracle as O

HIPPO = "/root/reference/py-lddmm/TestData/Hippo.vtk"


@pytest.mark.parametrize("s", [0, 1, 2, 4])
def test_icosphere_counts(s):
    v, f = synth.icosphere(s, 10.0)
    assert v.shape[0] == 10 * 4 ** s + 2 and f.shape[0] == 20 * 4 ** s
    assert O.surface_volume(v, f) > 0


def test_sphere_hull_face_count_and_orientation():
    v, f = synth.sphere_hull(1002, 50.0, seed=3)
    assert f.shape[0] == 2 * 1002 - 4
    assert O.surface_volume(v, f) > 0


def test_bump_deform_is_seeded():
    v, f = synth.icosphere(2, 10.0)
    a = synth.bump_deform(v, f, 20, 3.0, 1.0, seed=5)
    b = synth.bump_deform(v, f, 20, 3.0, 1.0, seed=5)
    c = synth.bump_deform(v, f, 20, 3.0, 1.0, seed=6)
    assert np.array_equal(a, b) and not np.array_equal(a, c)


def test_surface_geometry_matches_oracle():
    v, f = synth.icosphere(2, 3.0)
    w = np.linspace(1, 2, v.shape[0])
    S = Surface((f, v), weights=w)
    c, sf, fw = O.surface_geometry(v, f, w)
    np.testing.assert_allclose(S.centers, c)
    np.testing.assert_allclose(S.surfel, sf)
    np.testing.assert_allclose(S.face_weights, fw)
    S.flipFaces()
    assert S.surfVolume() < 0
    S.getEdges()
    assert S.bdry.max() == 0 and S.edges.shape[0] == 3 * f.shape[0] // 2


def test_vtk_roundtrip(tmp_path):
    v, f = synth.icosphere(1, 2.0)
    p = os.path.join(tmp_path, "s.vtk")
    vtk_io.write_polydata(p, v, f, weights=np.arange(v.shape[0], dtype=float))
    v2, f2, w2 = vtk_io.read_polydata(p)
    np.testing.assert_allclose(v2, v, rtol=1e-12)
    assert np.array_equal(f2, f)
    np.testing.assert_array_equal(w2, np.arange(v.shape[0]))


@pytest.mark.skipif(not os.path.exists(HIPPO), reason="reference test data not present")
def test_reads_reference_test_surface():
    S = Surface(HIPPO)
    assert S.vertices.shape == (1001, 3) and S.faces.shape == (1998, 3)
    S.getEdges()
    assert S.bdry.max() == 0


def test_closed_test_and_lazy_edge_tables():
    """Surface.is_closed equals getEdges' boundary test (bdry.max() == 0) on
    closed, open and doubled meshes; the edge tables appear on first access
    with getEdges' values."""
    v, f = synth.icosphere(2, 3.0)
    rng = np.random.default_rng(0)
    cases = [f, f[1:], f[rng.permutation(f.shape[0])[:-7]], np.concatenate([f, f])]
    expect = [True, False, False, True]       # the doubled mesh: every edge has four faces
    for faces, want in zip(cases, expect):
        S = Surface((faces, v))
        closed = S.is_closed()
        T = Surface((faces, v))
        T.getEdges()
        assert closed == bool(T.bdry.max() == 0) == want
        assert np.array_equal(S.edges, T.edges)         # built on first access
        assert np.array_equal(S.edgeFaces, T.edgeFaces) and np.array_equal(S.faceEdges, T.faceEdges)
        assert np.array_equal(S.bdry, T.bdry) and np.array_equal(S.vertices_bdry, T.vertices_bdry)
        assert S.is_closed() == closed                  # from the tables once they exist
    assert Surface().edges is None
    with pytest.raises(AttributeError):
        Surface().no_such_table
