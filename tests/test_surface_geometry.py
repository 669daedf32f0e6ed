"""Vertex areas and normals of lddmm_amd.surfaces.Surface (base/surfaces.py:364-465).

The mixed Voronoi split is pinned by building each corner's region of a
triangle from the geometry itself: in a face with no obtuse angle the region
of corner i is the quadrilateral (v_i, midpoint of e_ij, circumcentre,
midpoint of e_ik); in a face obtuse at corner i, corner j's region is the
triangle (v_j, midpoint of e_ij, point of edge jk equidistant from v_i and
v_j), and corner i keeps the rest.  Areas come from the polygon's vector area.
"""
import numpy as np
import pytest

from lddmm_amd import synth
from lddmm_amd.surfaces import Surface


def poly_area(*pts):
    s = np.zeros(3)
    for a, b in zip(pts, pts[1:] + pts[:1]):
        s += np.cross(a, b)
    return np.linalg.norm(s) / 2


def circumcentre(a, b, c):
    ab, ac = b - a, c - a
    n = np.cross(ab, ac)
    return a + (np.dot(ac, ac) * np.cross(n, ab) + np.dot(ab, ab) * np.cross(ac, n)) / (2 * np.dot(n, n))


def corner_regions(P):
    """Areas of the three corner regions of triangle P [3, 3]."""
    area = poly_area(*P)
    ang = []
    for i in range(3):
        u, w = P[(i + 1) % 3] - P[i], P[(i + 2) % 3] - P[i]
        ang.append(np.dot(u, w))
    obt = [i for i in range(3) if ang[i] < 0]
    if not obt:
        O = circumcentre(*P)
        return [poly_area(P[i], (P[i] + P[(i + 1) % 3]) / 2, O, (P[i] + P[(i + 2) % 3]) / 2) for i in range(3)]
    i = obt[0]
    out = [0.0] * 3
    for j in range(3):
        if j == i:
            continue
        k = 3 - i - j
        d = P[k] - P[j]
        s = -np.dot(P[j] - P[i], P[j] - P[i]) / (2 * np.dot(P[j] - P[i], d))
        q = P[j] + s * d
        assert 0 <= s <= 1
        assert abs(np.linalg.norm(q - P[i]) - np.linalg.norm(q - P[j])) < 1e-9
        out[j] = poly_area(P[j], (P[i] + P[j]) / 2, q)
    out[i] = area - sum(out)
    return out


def brute_vertex_area(V, F):
    AV = np.zeros(V.shape[0])
    for f in F:
        for c, a in zip(f, corner_regions(V[f])):
            AV[c] += a
    return AV


def flat_ellipsoid():
    v, f = synth.icosphere(2, 1.0)
    rng = np.random.default_rng(3)
    v = v * np.array([3.0, 1.0, 0.35]) + 0.02 * rng.normal(size=v.shape)
    return v, f


def test_single_triangles_closed_form():
    # equilateral: a third each
    P = np.array([[0, 0, 0], [1, 0, 0], [0.5, np.sqrt(3) / 2, 0]])
    AV, AF = Surface((np.array([[0, 1, 2]]), P)).computeVertexArea()
    np.testing.assert_allclose(AV, AF[0] / 3, rtol=1e-14)
    # obtuse at vertex 0: v0=(0,0), v1=(2,0), v2=(-1,1); corner 1 gets the right
    # triangle (v1, (1,0), (1,1/3)) = 1/6, corner 2 the right triangle with legs
    # |v0-v2|/2 = sqrt(2)/2 and sqrt(2)/4 (tan of the angle at v2 = 1/2) = 1/8
    P = np.array([[0, 0, 0], [2, 0, 0], [-1, 1, 0]], dtype=float)
    AV, AF = Surface((np.array([[0, 1, 2]]), P)).computeVertexArea()
    np.testing.assert_allclose(AF, [1.0], rtol=1e-14)
    np.testing.assert_allclose(AV, [1 - 1 / 6 - 1 / 8, 1 / 6, 1 / 8], rtol=1e-13)
    # right angle at vertex 0: the circumcentre is the hypotenuse's midpoint
    # (1, 1/2); the regions are the rectangle [0,1]x[0,1/2] and two triangles
    P = np.array([[0, 0, 0], [2, 0, 0], [0, 1, 0]], dtype=float)
    AV, _ = Surface((np.array([[0, 1, 2]]), P)).computeVertexArea()
    np.testing.assert_allclose(AV, [0.5, 0.25, 0.25], rtol=1e-13)


@pytest.mark.parametrize("mesh", ["icosphere", "flat_ellipsoid"])
def test_vertex_area_matches_geometric_regions(mesh):
    v, f = synth.icosphere(2, 2.0) if mesh == "icosphere" else flat_ellipsoid()
    S = Surface((f, v))
    AV, AF = S.computeVertexArea()
    ref = brute_vertex_area(v, f)
    np.testing.assert_allclose(AV, ref, rtol=1e-10)
    np.testing.assert_allclose(AV.sum(), AF.sum(), rtol=1e-12)
    np.testing.assert_allclose(AF.sum(), S.surfArea(), rtol=1e-12)
    if mesh == "flat_ellipsoid":       # the case split is exercised
        e = [v[f[:, (i + 1) % 3]] - v[f[:, i]] for i in range(3)]
        dots = [-(e[i] * e[(i + 2) % 3]).sum(1) for i in range(3)]
        assert (np.min(dots, axis=0) < 0).sum() > 20
    AVs, _ = S.computeVertexArea(simpleMethod=True)
    ref = np.zeros(v.shape[0])
    for k in range(3):
        np.add.at(ref, f[:, k], AF / 3)
    np.testing.assert_allclose(AVs, ref, rtol=1e-13)


def test_vertex_normals():
    # planar square: every normal is the plane's normal
    P = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], dtype=float)
    S = Surface((np.array([[0, 1, 2], [0, 2, 3]]), P))
    np.testing.assert_allclose(S.computeVertexNormals(), np.tile([0, 0, 1.0], (4, 1)), atol=1e-15)
    # sphere: outward, close to radial; closed surface: area vectors sum to zero
    v, f = synth.icosphere(3, 2.0)
    S = Surface((f, v))
    n = S.computeVertexNormals()
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, rtol=1e-14)
    assert (n * v / 2.0).sum(1).min() > 0.999
    aw = S.computeAreaWeightedVertexNormals()
    np.testing.assert_allclose(aw.sum(0), 3 * S.surfel.sum(0), atol=1e-12)
    assert np.abs(aw.sum(0)).max() < 1e-12
