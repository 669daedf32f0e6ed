"""Oracle pins of the learned warm start's ray-cast feature (CPU) and host
checks of its wrapper."""
import numpy as np
import pytest

from oracle import features as OF
from lddmm_amd import synth as S


def test_raycast_bump_pair_recovers_bump_heights():
    # the bump target is v + h(v) n(v) vertex by vertex, so the ray from a
    # template vertex along its normal passes through the target vertex
    (v, f), (v2, _) = S.plain_matching_pair(seed=3, subdiv=3)
    n = S.vertex_normals(v, f)
    h = ((v2 - v) * n).sum(1)
    assert np.abs(h).max() > 0.5
    t, hit = OF.normal_raycast(v, n, v2, f)
    assert hit.all()
    np.testing.assert_allclose(t, h, atol=1e-9, rtol=0)


def test_raycast_single_triangle_closed_forms():
    tv = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0]])
    tf = np.array([[0, 1, 2]])
    o = np.array([[0.2, 0.2, 1.0], [0.2, 0.2, 1.0], [0.2, 0.2, 1.0], [2.0, 2.0, 1.0], [0.25, 0.25, -3.0]])
    d = np.array([[0, 0, -1.0], [0, 0, 1.0], [1.0, 0, 0], [0, 0, 1.0], [0, 0, 2.0]])
    t, hit = OF.normal_raycast(o, d, tv, tf)
    np.testing.assert_array_equal(hit, [1, 1, 0, 0, 1])
    np.testing.assert_allclose(t, [1.0, -1.0, 0.0, 0.0, 1.5], atol=1e-15)


def test_raycast_matches_cramer_brute_force():
    rng = np.random.default_rng(0)
    tv, tf = S.sphere_hull(60, 3.0, seed=1)
    o = rng.uniform(-1.2, 1.2, (80, 3))      # inside the hull (inradius > 2)
    d = rng.normal(size=(80, 3))
    t, hit = OF.normal_raycast(o, d, tv, tf)
    for i in range(len(o)):
        ts = [OF.ray_triangle(o[i], d[i], *tv[fc]) for fc in tf]
        ts = [x for x in ts if x is not None]
        # a line through the inside of a closed hull meets it twice
        assert hit[i] == 1 and len(ts) >= 2
        best = min(ts, key=abs)
        assert abs(t[i] - best) < 1e-10 * max(1.0, abs(best))


def test_raycast_wrapper_rejects_bad_faces():
    from lddmm_amd.learned import features as F
    with pytest.raises(ValueError):
        F.normal_raycast(np.zeros((1, 3)), np.ones((1, 3)), np.zeros((3, 3)), np.array([[0, 1, 3]]))


def test_local_frame_is_orthonormal():
    from lddmm_amd.learned import features as F
    v, f = S.icosphere(2, 5.0)
    fr = F.local_frame(S.vertex_normals(v, f))
    np.testing.assert_allclose(np.einsum('nij,nkj->nik', fr, fr), np.broadcast_to(np.eye(3), fr.shape), atol=1e-12)
    np.testing.assert_allclose(np.linalg.det(fr), 1.0, atol=1e-12)
