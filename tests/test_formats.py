"""Surface and point-set file formats (CPU only).

Reads the reference's own TestData surface (tests/golden/ERC.vtk, copied
from py-lddmm/TestData) and round-trips every writer through its reader:
legacy VTK (saveVTK / saveVTK2 with fields and face labels), BYU, OFF, OBJ,
ASCII and binary STL, .lmk landmarks (including the version-6 centre / scale
lines), the 'N dim' vector format, savePoints and saveTrajectories.
"""
import os
import struct

import numpy as np
import pytest

from lddmm_amd import vtk_io
from lddmm_amd.pointSets import PointSet, savePoints, saveTrajectories, loadlmk, savelmk, readVector
from lddmm_amd.surfaces import Surface, vtkFields
from lddmm_amd import synth

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_reads_reference_testdata_vtk():
    s = Surface(os.path.join(GOLDEN, "ERC.vtk"))
    assert s.vertices.shape == (751, 3) and s.faces.shape == (1498, 3)
    np.testing.assert_allclose(s.vertices[0], [16.790265, 23.758322, 7.617858])
    assert np.all(s.weights == 1) and np.all(s.component == 0)
    s.getEdges()
    # closed triangulated sphere: V - E + F = 2, every edge shared by two faces
    assert s.vertices.shape[0] - s.edges.shape[0] + s.faces.shape[0] == 2
    assert s.bdry.max() == 0


@pytest.fixture
def mesh():
    v, f = synth.icosphere(1, radius=2.0)
    return Surface((f, v))


def test_vtk_round_trip_with_fields(tmp_path, mesh):
    rng = np.random.default_rng(0)
    mesh.updateWeights(rng.uniform(0.5, 1.5, mesh.vertices.shape[0]))
    mesh.component = np.arange(mesh.faces.shape[0]) % 3
    p = str(tmp_path / "a.vtk")
    nv = mesh.vertices.shape[0]
    mesh.saveVTK(p, scalars=rng.normal(size=nv), vectors=rng.normal(size=(nv, 3)),
                 normals=rng.normal(size=(nv, 3)), tensors=rng.normal(size=(nv, 3, 3)))
    r = Surface(p)
    np.testing.assert_array_equal(r.vertices, mesh.vertices)
    np.testing.assert_array_equal(r.faces, mesh.faces)
    np.testing.assert_array_equal(r.weights, mesh.weights)
    np.testing.assert_array_equal(r.component, mesh.component)
    vf = vtkFields()
    vf.scalars += ['a', np.arange(nv, dtype=float), 'weights', mesh.weights]
    mesh.saveVTK2(p, vf)
    with open(p) as fh:
        text = fh.read()
    assert text.startswith("# vtk DataFile Version 3.0\nSurface Data\nASCII\nDATASET POLYDATA\n")
    order = [text.index(k) for k in ("POINTS", "POLYGONS", "CELL_DATA", "SCALARS labels", "POINT_DATA",
                                     "SCALARS a")]
    assert order == sorted(order)
    np.testing.assert_array_equal(Surface(p).weights, mesh.weights)


@pytest.mark.parametrize("ext", [".byu", ".off"])
def test_byu_off_round_trip(tmp_path, mesh, ext):
    p = str(tmp_path / ("a" + ext))
    (mesh.savebyu if ext == ".byu" else mesh.saveOFF)(p)
    r = Surface(p)
    np.testing.assert_array_equal(r.vertices, mesh.vertices)
    np.testing.assert_array_equal(r.faces, mesh.faces)


def test_obj_reader(tmp_path, mesh):
    p = str(tmp_path / "a.obj")
    with open(p, "w") as fh:
        fh.write("# comment\nmtllib x.mtl\n")
        for q in mesh.vertices:
            fh.write(f"v {float(q[0])!r} {float(q[1])!r} {float(q[2])!r}\n")
        fh.write("vn 0 0 1\n")
        for k, t in enumerate(mesh.faces):
            if k % 2:
                fh.write(f"f {t[0] + 1}/1/1 {t[1] + 1}/1/1 {t[2] + 1}/1/1\n")
            else:
                n = mesh.vertices.shape[0]
                fh.write(f"f {t[0] - n} {t[1] - n} {t[2] - n}\n")     # negative (relative) indices
    r = Surface(p)
    np.testing.assert_array_equal(r.vertices, mesh.vertices)
    np.testing.assert_array_equal(r.faces, mesh.faces)


@pytest.mark.parametrize("binary", [False, True])
def test_stl_reader_merges_points(tmp_path, mesh, binary):
    tri = mesh.vertices[mesh.faces].astype(np.float32).astype(np.float64)
    p = str(tmp_path / "a.stl")
    if binary:
        with open(p, "wb") as fh:
            fh.write(b"\0" * 80 + struct.pack("<I", len(tri)))
            for t in tri:
                fh.write(struct.pack("<3f", 0, 0, 1) + struct.pack("<9f", *t.ravel()) + b"\0\0")
    else:
        with open(p, "w") as fh:
            fh.write("solid s\n")
            for t in tri:
                fh.write("facet normal 0 0 1\nouter loop\n")
                for q in t:
                    fh.write(f"vertex {float(q[0])!r} {float(q[1])!r} {float(q[2])!r}\n")
                fh.write("endloop\nendfacet\n")
            fh.write("endsolid s\n")
    r = Surface(p)
    assert r.vertices.shape == mesh.vertices.shape and r.faces.shape == mesh.faces.shape
    np.testing.assert_allclose(r.vertices[r.faces], tri, rtol=0, atol=0)
    np.testing.assert_allclose(r.surfVolume(), mesh.surfVolume(), rtol=1e-6)


def test_landmarks_and_vectors(tmp_path):
    rng = np.random.default_rng(1)
    x = rng.normal(size=(7, 3))
    p = str(tmp_path / "a.lmk")
    savelmk(x, p)
    y, lab = loadlmk(p)
    np.testing.assert_array_equal(y, x)
    assert lab[0].strip() == '"L-0"'
    # version 6: the last two lines are the centre and the scale
    with open(p) as fh:
        lines = fh.read().splitlines()
    lines[0] = "Landmarks-6.0"
    lines = lines[:-1] + ["1.0,2.0,3.0", "2.0,2.0,0.5"]
    with open(p, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    y6, _ = loadlmk(p)
    np.testing.assert_allclose(y6, x * [2.0, 2.0, 0.5] + [1.0, 2.0, 3.0], rtol=1e-15)
    q = str(tmp_path / "a.txt")
    with open(q, "w") as fh:
        fh.write("7 3\n" + "".join(f"{float(a)!r} {float(b)!r} {float(c)!r} {float(w)!r}\n" for (a, b, c), w in zip(x, range(7))))
    v, w = readVector(q)
    np.testing.assert_array_equal(v, x)
    np.testing.assert_array_equal(w.ravel(), np.arange(7))
    ps = PointSet(q)
    np.testing.assert_array_equal(ps.vertices, x)


def test_save_points_and_trajectories(tmp_path):
    rng = np.random.default_rng(2)
    x = rng.normal(size=(9, 2))
    p = str(tmp_path / "p.vtk")
    savePoints(p, x, vectors={'v': rng.normal(size=(9, 3))}, scalars={'weights': np.arange(9.0)})
    y, sc, vec = vtk_io.read_points(p)
    np.testing.assert_array_equal(y[:, :2], x)
    assert np.all(y[:, 2] == 0)
    np.testing.assert_array_equal(sc['weights'], np.arange(9.0))
    assert vec['v'].shape == (9, 3)
    ps = PointSet(rng.normal(size=(5, 3)), weights=np.arange(1.0, 6.0))
    ps.save(p)
    np.testing.assert_array_equal(vtk_io.read_points(p)[1]['weights'], np.arange(1.0, 6.0))
    xt = rng.normal(size=(4, 6, 2))
    saveTrajectories(p, xt)
    with open(p) as fh:
        text = fh.read()
    assert "POINTS 24 float" in text and "LINES 18 54" in text and "POINT_DATA 24" in text


def test_obj_point_cloud_weights(tmp_path, mesh):
    p = str(tmp_path / "c.obj")
    with open(p, "w") as fh:
        for q in mesh.vertices:
            fh.write(f"v {q[0]} {q[1]} {q[2]}\n")
    ps = PointSet(p, maxPoints=30)
    assert ps.vertices.shape == (30, 3)
    assert np.all(ps.weights > 0) and np.all(np.isfinite(ps.weights))
