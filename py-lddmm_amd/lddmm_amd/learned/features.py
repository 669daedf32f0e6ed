"""Input features of the learned warm start that need no vertex
correspondence between template and target.

normal_heights: signed distance from each template vertex to the target
surface along the template's unit vertex normal (lddmm_normal_raycast), then
the same heights smoothed over the template at two scales (normalised
Gaussian averages, lddmm_applyK).  The three channels are invariant under
rigid motions of the pair; the network's output is read in each vertex's
local frame (normal, two tangents), so the prediction turns with the shape.
"""
import numpy as np
import torch

from .. import _backend as B
from .. import kernelFunctions_util as ku
from ..synth import vertex_normals

SMOOTH_SCALES = (2.5, 5.0)


def local_frame(normals):
    """(n, t1, t2) per vertex, [N, 3, 3] with rows n, t1, t2: t1 is the first
    coordinate axis not nearly parallel to n, made orthogonal to n; t2 = n x t1."""
    n = np.asarray(normals, dtype=np.float64)
    ax = np.where(np.abs(n[:, :1]) < 0.9, np.array([[1.0, 0.0, 0.0]]), np.array([[0.0, 1.0, 0.0]]))
    t1 = ax - (ax * n).sum(1, keepdims=True) * n
    t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
    t2 = np.cross(n, t1)
    return np.stack([n, t1, t2], 1)


def normal_raycast(orig, dirs, tv, tf):
    """Device call of lddmm_normal_raycast.  orig, dirs [n, 3], tv [nv, 3]
    (float64, numpy or device), tf [nf, 3] integer faces: a host array
    (checked here: every index must lie in [0, nv)) or the int32 device faces
    of a Surface (surface_on_device: checked when the Surface built its
    adjacency).  Returns (t [n] float64, hit [n] int32) device tensors."""
    nv = int(np.asarray(tv).shape[0]) if not isinstance(tv, torch.Tensor) else int(tv.shape[0])
    on_device = isinstance(tf, torch.Tensor)
    if on_device:
        if tf.dtype != torch.int32 or tf.ndim != 2 or tf.shape[1] != 3:
            raise ValueError("normal_raycast: device faces must be int32 [nf, 3]")
    else:
        tf = np.asarray(tf)
        if tf.size and (tf.min() < 0 or tf.max() >= nv):
            raise ValueError("normal_raycast: face index out of range")
    o = B.to_dev(orig).contiguous()
    d = B.to_dev(dirs).contiguous()
    if o.shape != d.shape or o.ndim != 2 or o.shape[1] != 3:
        raise ValueError("normal_raycast: origins and directions must be [n, 3]")
    v = B.to_dev(tv).contiguous()
    if on_device:
        if tf.device != o.device:
            raise ValueError("normal_raycast: device faces must be on the rays' device")
        f = tf.contiguous()
    else:
        f = torch.from_numpy(np.ascontiguousarray(tf, dtype=np.int32)).to(o.device)
    n = o.shape[0]
    t = torch.empty(n, dtype=torch.float64, device=o.device)
    hit = torch.empty(n, dtype=torch.int32, device=o.device)
    B.check(B.LIB.lddmm_normal_raycast(B.ptr(o), B.ptr(d), n, B.ptr(v), v.shape[0], B.ptr(f), f.shape[0],
                                       B.ptr(t), B.ptr(hit), B.stream_handle()), "normal_raycast")
    return t, hit


def local_frame_device(n):
    """local_frame on the device: [N, 3, 3] with rows n, t1, t2 for unit
    normals n [N, 3] (float64 device tensor)."""
    e1 = torch.tensor([1.0, 0.0, 0.0], dtype=n.dtype, device=n.device)
    e2 = torch.tensor([0.0, 1.0, 0.0], dtype=n.dtype, device=n.device)
    ax = torch.where(n[:, :1].abs() < 0.9, e1, e2)
    t1 = ax - (ax * n).sum(1, keepdim=True) * n
    t1 = t1 / t1.norm(dim=1, keepdim=True)
    return torch.stack([n, t1, torch.cross(n, t1, dim=1)], 1)


def surface_on_device(surface):
    """(vertices [nv, 3] float64, faces [nf, 3] int32) device tensors of a
    Surface, cached by the Surface until its vertices or faces change."""
    _, keep = surface.device()
    return keep[0], keep[2]


def smooth(x0, values, sigma):
    """Normalised Gaussian average over the points x0 of per-point values
    [N, c]: sum_j K(x_i, x_j) values_j / sum_j K(x_i, x_j)."""
    w = torch.cat([values, torch.ones_like(values[:, :1])], 1)
    s = ku.applyK(x0, x0, w, 'gauss', [sigma], 10)
    return s[:, :-1] / s[:, -1:]


def device_vertex_normals(surface, x):
    """Unit vertex normals [N, 3] of `surface` with its vertices replaced by x
    [N, 3] (device tensor), as synth.vertex_normals computes them on the
    host: per vertex the sum of its faces' cross products (x2 - x1) x (x3 - x1),
    here added in a fixed order by the CSR face gather, then normalised."""
    from .. import surfaceDistances as sd
    faces = surface.topology()[0].long()
    x1, x2, x3 = x[faces[:, 0]], x[faces[:, 1]], x[faces[:, 2]]
    nf = torch.cross(x2 - x1, x3 - x1, dim=1)
    s = sd._gather(surface.device(vertices=x), nf, torch.zeros_like(nf))
    return s / s.norm(dim=1, keepdim=True).clamp_min(1e-300)


def surface_normal_heights(template, target):
    """normal_heights of two Surfaces with every step on the device: the
    template's unit vertex normals from its faces (device_vertex_normals),
    the ray cast against the target's cached device vertices and faces, the
    smoothed channels and the local frames (local_frame_device).  Equal to
    normal_heights of the surfaces' arrays up to the rounding of the normals'
    face sums (the host adds them in another order)."""
    x0, _ = surface_on_device(template)
    nrm = device_vertex_normals(template, x0)
    tv, tf = surface_on_device(target)
    h, _ = normal_raycast(x0, nrm, tv, tf)
    h = h[:, None]
    ch = [h] + [smooth(x0, h, s) for s in SMOOTH_SCALES]
    return torch.cat(ch, 1), local_frame_device(nrm)


def normal_heights(template_vertices, template_faces, target_vertices, target_faces):
    """Channels [N, 3] (float64 device): the ray-cast normal height h and
    h smoothed at SMOOTH_SCALES; plus the local frames [N, 3, 3] (device)."""
    x0 = np.asarray(template_vertices, dtype=np.float64)
    nrm = vertex_normals(x0, template_faces)
    h, _ = normal_raycast(x0, nrm, target_vertices, target_faces)
    x0d = B.to_dev(x0)
    h = h[:, None]
    ch = [h] + [smooth(x0d, h, s) for s in SMOOTH_SCALES]
    return torch.cat(ch, 1), B.to_dev(local_frame(nrm))
