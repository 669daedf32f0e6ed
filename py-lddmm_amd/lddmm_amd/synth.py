"""Seeded synthetic shapes for the benchmark workloads (numpy, float64).

The reference builds test spheres by marching cubes + VTK decimation
(py-lddmm base/surfaceExamples.py:115-129), which needs vtk / skimage; the
workloads here need exact vertex counts, so they are built from closed forms:

* icosphere(s, R): icosahedron subdivided s times, projected on the sphere of
  radius R; V = 10*4^s + 2 vertices, F = 20*4^s outward-oriented faces.
* sphere_hull(n, R, seed): n seeded points (uniform on the sphere, or a
  Fibonacci lattice) and their convex hull: F = 2n - 4 outward faces.
* bump_deform(v, f, nbumps, width, amp_sd, seed): v + sum_b A_b
  exp(-|v - c_b|^2 / (2 width^2)) n(v), centres c_b uniform on the shape's
  mean-radius sphere, A_b ~ N(0, amp_sd^2), n(v) the outward unit vertex normal.
  "width" is the Gaussian sigma of each bump and amp_sd its amplitude standard
  deviation.
"""
import functools

import numpy as np


def _orient_outward(v, f):
    """Flip faces whose normal points towards the centroid (star-shaped meshes)."""
    c = v.mean(0)
    x1, x2, x3 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    nrm = np.cross(x2 - x1, x3 - x1)
    out = ((x1 + x2 + x3) / 3 - c)
    bad = (nrm * out).sum(1) < 0
    f = f.copy()
    f[bad] = f[bad][:, [0, 2, 1]]
    return f


def icosphere(subdiv, radius=1.0):
    """Icosahedron subdivided `subdiv` times, projected to the sphere of the
    given radius: V = 10 * 4^subdiv + 2 vertices, F = 20 * 4^subdiv outward faces.
    The unit mesh of each subdivision level is built once and copied."""
    v, f = _unit_icosphere(int(subdiv))
    return v * radius, f.copy()


@functools.lru_cache(maxsize=8)
def _unit_icosphere(subdiv):
    t = (1.0 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = np.array(f, dtype=np.int64)
    for _ in range(subdiv):
        cache = {}

        def mid(i, j):
            key = (i, j) if i < j else (j, i)
            if key not in cache:
                m = verts[i] + verts[j]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = np.array(nf, dtype=np.int64)
    V = np.array(verts)
    f = _orient_outward(V, faces)
    V.flags.writeable = False
    f.flags.writeable = False
    return V, f


def vertex_normals(v, f):
    x1, x2, x3 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    nf = np.cross(x2 - x1, x3 - x1)
    nv = np.zeros_like(v)
    for j in range(3):
        np.add.at(nv, f[:, j], nf)
    return nv / np.maximum(np.linalg.norm(nv, axis=1, keepdims=True), 1e-300)


def bump_deform(v, f, nbumps, width, amp_sd, seed, center=None):
    rng = np.random.default_rng(seed)
    c = v.mean(0) if center is None else np.asarray(center, dtype=np.float64)
    R = np.linalg.norm(v - c, axis=1).mean()
    cb = rng.normal(size=(nbumps, 3))
    cb = c + R * cb / np.linalg.norm(cb, axis=1, keepdims=True)
    A = rng.normal(0.0, amp_sd, size=nbumps)
    n = vertex_normals(v, f)
    d2 = ((v[:, None, :] - cb[None]) ** 2).sum(-1)
    h = (A[None] * np.exp(-d2 / (2 * width ** 2))).sum(1)
    return v + h[:, None] * n


def fibonacci_sphere(n, radius=1.0):
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    th = np.pi * (1 + 5 ** 0.5) * i
    return radius * np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)


def sphere_hull(n, radius=1.0, seed=None, fibonacci=False):
    """Convex hull of n points on a sphere: exactly 2n - 4 outward faces."""
    from scipy.spatial import ConvexHull
    if fibonacci:
        p = fibonacci_sphere(n, radius)
    else:
        rng = np.random.default_rng(seed)
        p = rng.normal(size=(n, 3))
        p = radius * p / np.linalg.norm(p, axis=1, keepdims=True)
    h = ConvexHull(p)
    return p, _orient_outward(p, np.asarray(h.simplices, dtype=np.int64))


def ellipsoid(n, axes, seed=None):
    v, f = sphere_hull(n, 1.0, seed=seed, fibonacci=True)
    return v * np.asarray(axes, dtype=np.float64)[None], f


def plain_matching_pair(seed=0, subdiv=4, radius=10.0, nbumps=20, width=3.0, amp_sd=1.0):
    """Template icosphere and bump-deformed target (plain surface-matching workload)."""
    v, f = icosphere(subdiv, radius)
    return (v, f), (bump_deform(v, f, nbumps, width, amp_sd, seed), f)


def uniform_cloud(n, lo=0.0, hi=50.0, seed=0, dim=3):
    return np.random.default_rng(seed).uniform(lo, hi, size=(n, dim))


def closed_curve(n, dim=2, radius=5.0, nmodes=4, amp=0.15, seed=0):
    """Seeded closed curve of n vertices: a circle of the given radius (in the
    xy plane) with random low-frequency radial modes; in 3-D the z coordinate
    carries its own modes.  Returns (vertices [n, dim], segments [n, 2])."""
    rng = np.random.default_rng(seed)
    t = np.linspace(0, 2 * np.pi, n, endpoint=False)
    r = np.ones(n)
    for k in range(1, nmodes + 1):
        r += amp / k * (rng.normal() * np.cos(k * t) + rng.normal() * np.sin(k * t))
    v = np.stack([radius * r * np.cos(t), radius * r * np.sin(t)], 1)
    if dim == 3:
        z = np.zeros(n)
        for k in range(1, nmodes + 1):
            z += amp * radius / k * (rng.normal() * np.cos(k * t) + rng.normal() * np.sin(k * t))
        v = np.concatenate([v, z[:, None]], 1)
    f = np.stack([np.arange(n), (np.arange(n) + 1) % n], 1)
    return v, f


def curve_matching_pair(seed=0, n=400, dim=2, radius=5.0):
    """Template = closed_curve(seed), target = closed_curve(seed + 1) with the
    same vertex count and segment list."""
    v, f = closed_curve(n, dim, radius, seed=seed)
    v2, _ = closed_curve(n, dim, radius, seed=seed + 1)
    return (v, f), (v2, f)


def population_pairs(npairs=256, seed=0, nmin=500, nmax=20000):
    """Seeded template-target pairs of the batched population workload.

    Per pair: a vertex count N log-uniform in [nmin, nmax]; the template is a
    Fibonacci-lattice sphere hull (radius U[5, 20]) or an ellipsoid hull with
    axes U[5, 20]^3 (one in two each); the target is the template with
    U{10..50} Gaussian bumps of width 0.3 R and amplitude N(0, (0.1 R)^2)
    along the normals, R the template's mean radius (the plain workload's
    width 3 / amplitude 1 at R = 10).  Returns [((v, f), (v2, f)), ...]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(npairs):
        n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        if rng.random() < 0.5:
            v, f = sphere_hull(n, radius=rng.uniform(5.0, 20.0), fibonacci=True)
        else:
            v, f = ellipsoid(n, rng.uniform(5.0, 20.0, size=3))
        R = np.linalg.norm(v - v.mean(0), axis=1).mean()
        nb = int(rng.integers(10, 51))
        v2 = bump_deform(v, f, nb, 0.3 * R, 0.1 * R, int(rng.integers(1 << 30)))
        out.append(((v, f), (v2, f)))
    return out
