"""Polygonal curves with the reference's data layout (py-lddmm base/curves.py:21-160).

`Curve(curve=(faces, vertices))`, `Curve(curve=vertices_array, isOpen=...)`
or `Curve(curve=other_curve)`; attributes vertices [nv, dim] (dim 2 or 3),
faces [ns, 2] (segments), weights [nv], centers, linel, line_weights,
component.  Geometry follows computeCentersLengths (base/curves.py:117-125).
`device()` gives the C-ABI view used by the curve data terms.  File I/O,
plotting and resampling are not part of this package.
"""
import numpy as np


class Curve:
    def __init__(self, curve=None, isOpen=False, c=0.0001):
        self._topo = None
        if isinstance(curve, (list, tuple)):
            if len(curve) and isinstance(curve[0], Curve):
                self.concatenate(curve)
            else:
                self.vertices = np.array(curve[1], dtype=np.float64)
                self.faces = np.array(curve[0], dtype=np.int64).reshape(-1, 2)
                self.weights = np.ones(self.vertices.shape[0])
                self.component = np.zeros(self.faces.shape[0], dtype=int)
                self.computeCentersLengths()
        elif isinstance(curve, np.ndarray):
            # polygon through the rows of `curve` (base/curves.py:38-47)
            self.vertices = np.array(curve, dtype=np.float64)
            n = self.vertices.shape[0]
            self.faces = np.zeros([n, 2], dtype=np.int64)
            self.weights = np.ones(n)
            self.component = np.zeros(n, dtype=int)
            self.faces[:n - 1, 0] = np.arange(n - 1)
            self.faces[:n - 1, 1] = np.arange(1, n)
            if not isOpen and n > 0:
                self.faces[n - 1] = (n - 1, 0)
            # an open polygon keeps the reference's last row (0, 0): a zero-length segment
            self.computeCentersLengths()
        elif isinstance(curve, Curve):
            self.vertices = np.copy(curve.vertices)
            self.faces = np.copy(curve.faces)
            self.weights = np.copy(curve.weights)
            self.component = np.copy(curve.component)
            self.computeCentersLengths()
            self._topo = curve._topo
        elif isinstance(curve, str):
            raise NotImplementedError("lddmm_amd.curves: reading curve files is not built")
        else:
            self.vertices = np.empty(0)
            self.faces = np.empty((0, 2), dtype=np.int64)
            self.weights = np.empty(0)
            self.component = np.empty(0, dtype=int)
            self.centers = np.empty(0)
            self.linel = np.empty(0)
            self.line_weights = np.empty(0)

    def concatenate(self, fvl):
        """base/curves.py:88-115 without duplicate removal."""
        self.vertices = np.concatenate([fv.vertices for fv in fvl])
        self.weights = np.concatenate([fv.weights for fv in fvl])
        off = np.cumsum([0] + [fv.vertices.shape[0] for fv in fvl])
        self.faces = np.concatenate([fv.faces + o for fv, o in zip(fvl, off)])
        self.component = np.concatenate([np.full(fv.faces.shape[0], k, dtype=int) for k, fv in enumerate(fvl)])
        self.computeCentersLengths()

    def computeCentersLengths(self):
        x1 = self.vertices[self.faces[:, 0], :]
        x2 = self.vertices[self.faces[:, 1], :]
        self.centers = (x1 + x2) / 2
        self.linel = x2 - x1
        self.line_weights = (self.weights[self.faces[:, 0]] + self.weights[self.faces[:, 1]]) / 2

    def updateVertices(self, x0):
        self.vertices = np.array(x0, dtype=np.float64)
        self.computeCentersLengths()

    def updateWeights(self, w0):
        self.weights = np.copy(w0)
        self.line_weights = (self.weights[self.faces[:, 0]] + self.weights[self.faces[:, 1]]) / 2

    def flipFaces(self):
        self.faces = self.faces[:, [1, 0]]
        self._topo = None
        self.computeCentersLengths()

    def length(self):
        return np.sqrt((self.linel ** 2).sum(axis=1)).sum()

    def computeUnitFaceNormals(self):
        a = np.sqrt((self.linel ** 2).sum(axis=1))
        normals = np.zeros(self.faces.shape)
        normals[:, 0] = -self.linel[:, 1] / a
        normals[:, 1] = self.linel[:, 0] / a
        return normals

    # ------------------------------------------------------------ device
    def topology(self):
        """int32 segments and CSR adjacency on the device (cached per topology)."""
        import torch
        from . import _backend as B
        if self._topo is None or self._topo[0].shape[0] != self.faces.shape[0]:
            s32 = np.ascontiguousarray(self.faces, dtype=np.int32)
            nv = self.vertices.shape[0]
            ptr = np.zeros(nv + 1, dtype=np.int32)
            idx = np.zeros(max(2 * s32.shape[0], 1), dtype=np.int32)
            B.check(B.LIB.lddmm_curve_adjacency(s32.ctypes.data, s32.shape[0], nv, ptr.ctypes.data,
                                                idx.ctypes.data), "curve_adjacency")
            self._topo = (B.upload(s32), B.upload(ptr), B.upload(idx))
        return self._topo

    def device(self, vertices=None):
        """LddmmCurve view (+ keep-alive tensors); `vertices` may be a device
        tensor [nv, dim] replacing this curve's vertices for the call."""
        from . import _backend as B
        segs, ptr, idx = self.topology()
        v = B.to_dev(self.vertices if vertices is None else vertices)
        w = None
        if not np.all(self.weights == 1):
            w = B.to_dev(np.asarray(self.weights, dtype=np.float64))
        c = B.LddmmCurve()
        c.vertices = v.data_ptr()
        c.nv = v.shape[0]
        c.dim = v.shape[1] if v.dim() == 2 else 3
        c.segments = segs.data_ptr()
        c.ns = segs.shape[0]
        c.vweights = None if w is None else w.data_ptr()
        c.adj_ptr = ptr.data_ptr()
        c.adj_idx = idx.data_ptr()
        return c, (v, w, segs, ptr, idx)
