"""Float64 numpy restatement of the learned warm start's input features
(test infrastructure only; nothing in the product imports it).

These functions have no counterpart in the reference.  Pinned by
tests/test_oracle.py: the ray from a template vertex of the bump workload
along its vertex normal meets the target exactly at the bump height (the
target vertex is v + h(v) n(v) by construction); a ray through the centre
of a unit-distance triangle gives t = +-1; a ray parallel to every triangle
misses; brute force over single triangles agrees with the vectorised form.
"""
import numpy as np

BARY_TOL = 1e-9


def normal_raycast(orig, dirs, tv, tf, chunk=256):
    """For each ray i (orig[i] + t dirs[i], t real) the intersection with the
    triangle mesh (tv, tf) of smallest |t|, lowest face index on ties.
    Moller-Trumbore with barycentric slack BARY_TOL.  Returns (t [n], hit [n]);
    t = 0 and hit = 0 where the line meets no triangle."""
    orig = np.asarray(orig, dtype=np.float64)
    dirs = np.asarray(dirs, dtype=np.float64)
    tv = np.asarray(tv, dtype=np.float64)
    tf = np.asarray(tf, dtype=np.int64)
    v0 = tv[tf[:, 0]]
    e1 = tv[tf[:, 1]] - v0
    e2 = tv[tf[:, 2]] - v0
    n = orig.shape[0]
    t_out = np.zeros(n)
    hit = np.zeros(n, dtype=np.int32)
    for r0 in range(0, n, chunk):
        o = orig[r0:r0 + chunk, None, :]
        d = dirs[r0:r0 + chunk, None, :]
        p = np.cross(d, e2[None])
        det = (e1[None] * p).sum(-1)
        ok = det != 0
        inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
        s = o - v0[None]
        u = (s * p).sum(-1) * inv
        q = np.cross(s, e1[None])
        v = (d * q).sum(-1) * inv
        t = (e2[None] * q).sum(-1) * inv
        ok &= (u >= -BARY_TOL) & (u <= 1 + BARY_TOL) & (v >= -BARY_TOL) & (u + v <= 1 + BARY_TOL)
        at = np.where(ok, np.abs(t), np.inf)
        j = np.argmin(at, axis=1)                 # first index of the minimum
        rows = np.arange(at.shape[0])
        found = np.isfinite(at[rows, j])
        t_out[r0:r0 + chunk] = np.where(found, t[rows, j], 0.0)
        hit[r0:r0 + chunk] = found
    return t_out, hit


def ray_triangle(o, d, a, b, c):
    """Single ray / single triangle, written out from the plane equation
    (solve o + t d = a + u (b - a) + v (c - a) by Cramer's rule); None when
    the line misses the triangle or is parallel to its plane."""
    M = np.stack([b - a, c - a, -d], 1)
    if abs(np.linalg.det(M)) == 0:
        return None
    u, v, t = np.linalg.solve(M, o - a)
    if u < -BARY_TOL or v < -BARY_TOL or u + v > 1 + BARY_TOL:
        return None
    return t
