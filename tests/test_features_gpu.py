"""GPU parity of the warm start's ray-cast feature (lddmm_normal_raycast)
against the float64 numpy oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def F():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lddmm_amd.learned import features
    return features


@pytest.mark.parametrize("subdiv", [3, 4])
def test_raycast_bump_pair(F, subdiv):
    from oracle import features as OF
    from lddmm_amd import synth as S
    (v, f), (v2, _) = S.plain_matching_pair(seed=11, subdiv=subdiv)
    n = S.vertex_normals(v, f)
    print(f"[raycast] bump pair subdiv {subdiv}: {len(v)} rays x {len(f)} faces", flush=True)
    t, hit = F.normal_raycast(v, n, v2, f)
    t_r, hit_r = OF.normal_raycast(v, n, v2, f)
    np.testing.assert_array_equal(hit.cpu().numpy(), hit_r)
    np.testing.assert_allclose(t.cpu().numpy(), t_r, atol=1e-12, rtol=1e-12)
    np.testing.assert_allclose(t.cpu().numpy(), ((v2 - v) * n).sum(1), atol=1e-9)


def test_raycast_random_rays_and_misses(F):
    from oracle import features as OF
    from lddmm_amd import synth as S
    rng = np.random.default_rng(4)
    tv, tf = S.sphere_hull(700, 4.0, seed=2)
    # origins inside and outside the hull (some lines miss it), 300 rays: ragged last tile
    o = rng.uniform(-7, 7, (300, 3))
    d = rng.normal(size=(300, 3))
    print(f"[raycast] random: {len(o)} rays x {len(tf)} faces", flush=True)
    t, hit = F.normal_raycast(o, d, tv, tf)
    t_r, hit_r = OF.normal_raycast(o, d, tv, tf)
    assert 0 < hit_r.sum() < len(o)
    np.testing.assert_array_equal(hit.cpu().numpy(), hit_r)
    np.testing.assert_allclose(t.cpu().numpy(), t_r, atol=1e-12, rtol=1e-12)


def test_raycast_empty_inputs(F):
    t, hit = F.normal_raycast(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros((3, 3)), np.zeros((0, 3), np.int32))
    assert t.numel() == 0
    t, hit = F.normal_raycast(np.zeros((5, 3)), np.ones((5, 3)), np.zeros((0, 3)), np.zeros((0, 3), np.int32))
    assert int(hit.sum()) == 0 and float(t.abs().sum()) == 0.0


def test_normal_heights_channels(F):
    import torch
    from lddmm_amd import synth as S
    (v, f), (v2, _) = S.plain_matching_pair(seed=5, subdiv=4)
    ch, fr = F.normal_heights(v, f, v2, f)
    n = S.vertex_normals(v, f)
    np.testing.assert_allclose(ch[:, 0].cpu().numpy(), ((v2 - v) * n).sum(1), atol=1e-9)
    # smoothed channels: normalised Gaussian averages of the heights (dense float64 reference)
    h = ((v2 - v) * n).sum(1)
    D2 = ((v[:, None] - v[None]) ** 2).sum(-1)
    for k, s in enumerate(F.SMOOTH_SCALES):
        G = np.exp(-D2 / (2 * s * s))
        np.testing.assert_allclose(ch[:, k + 1].cpu().numpy(), G @ h / G.sum(1), atol=1e-9)
    np.testing.assert_allclose(fr[:, 0].cpu().numpy(), n, atol=1e-15)


def test_raycast_out_of_range_face_is_never_hit(F):
    """Called through the C ABI directly (the Python wrapper rejects such faces):
    a face with an index outside [0, nv) is degenerate, the others still hit."""
    import torch
    from lddmm_amd import _backend as B
    tv = torch.tensor([[0.0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=torch.float64, device="cuda")
    tf = torch.tensor([[0, 1, 5], [0, 1, 2], [-1, 1, 2]], dtype=torch.int32, device="cuda")
    o = torch.tensor([[0.2, 0.2, 1.0], [3.0, 3.0, 1.0]], dtype=torch.float64, device="cuda")
    d = torch.tensor([[0, 0, -1.0], [0, 0, -1.0]], dtype=torch.float64, device="cuda")
    t = torch.empty(2, dtype=torch.float64, device="cuda")
    hit = torch.empty(2, dtype=torch.int32, device="cuda")
    B.check(B.LIB.lddmm_normal_raycast(B.ptr(o), B.ptr(d), 2, B.ptr(tv), 3, B.ptr(tf), 3, B.ptr(t), B.ptr(hit),
                                       B.stream_handle()), "normal_raycast")
    assert hit.cpu().tolist() == [1, 0]
    assert abs(float(t[0]) - 1.0) < 1e-15 and float(t[1]) == 0.0


def test_device_vertex_normals_match_host(F):
    """device_vertex_normals (face cross products summed per vertex by the
    CSR face gather) against synth.vertex_normals on a deformed icosphere and
    an ellipsoid hull, with the vertices replaced by a device tensor."""
    import torch
    from lddmm_amd import synth as S
    from lddmm_amd.surfaces import Surface
    (v, f), (v2, _) = S.plain_matching_pair(seed=5, subdiv=3)
    tv, tf = S.sphere_hull(900, 4.0, seed=3)
    for verts, faces, moved in ((v, f, v2), (tv, tf, tv * np.array([1.5, 1.0, 0.7]))):
        surf = Surface((faces, verts))
        got = F.device_vertex_normals(surf, torch.tensor(moved, device='cuda')).cpu().numpy()
        np.testing.assert_allclose(got, S.vertex_normals(moved, faces), rtol=0, atol=1e-13)


def test_surface_normal_heights_match_array_path(F):
    """The device-resident feature path (surface_normal_heights: device
    normals, cached device target, device frames) against normal_heights on
    the arrays: equal up to the rounding of the normals' face sums."""
    from lddmm_amd import synth as S
    from lddmm_amd.surfaces import Surface
    (v, f), (v2, _) = S.plain_matching_pair(seed=7, subdiv=4)
    print(f"[features] device path: {len(v)} template vertices", flush=True)
    ch, fr = F.normal_heights(v, f, v2, f)
    ch_d, fr_d = F.surface_normal_heights(Surface((f, v)), Surface((f, v2)))
    np.testing.assert_allclose(ch_d.cpu().numpy(), ch.cpu().numpy(), rtol=0, atol=1e-11)
    np.testing.assert_allclose(fr_d.cpu().numpy(), fr.cpu().numpy(), rtol=0, atol=1e-12)
    # the host path's frames are local_frame of the host normals
    np.testing.assert_allclose(fr.cpu().numpy(), F.local_frame(S.vertex_normals(v, f)), rtol=0, atol=0)
    with pytest.raises(ValueError):
        F.normal_raycast(v, S.vertex_normals(v, f), v2, F.surface_on_device(Surface((f, v2)))[1].long())
