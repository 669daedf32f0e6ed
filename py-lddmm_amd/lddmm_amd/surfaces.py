"""Triangulated surfaces (py-lddmm base/surfaces.py:226-351, :603-605, :1645-1651).

`Surface((faces, vertices))`, `Surface(surf=other)` and `Surface('file')`
(.vtk legacy ASCII POLYDATA, the format of the reference's TestData, .byu,
.off, .obj, .stl) build
the same attributes as the reference: vertices, faces, weights, centers,
surfel, face_weights, component.  `device()` returns the C-ABI view of the
surface (float64 vertices, int32 faces, vertex -> face adjacency) on the GPU,
rebuilt only when the vertices or topology change.
"""
import logging
import os

import numpy as np

from . import vtk_io
from .vtk_io import vtkFields  # noqa: F401  (reference name, base/surfaces.py:53-58)


class Surface:
    def __init__(self, surf=None, weights=None):
        self._dev = None
        self._topo = None
        if type(surf) in (list, tuple):
            if len(surf) > 0 and isinstance(surf[0], Surface):
                self.concatenate(surf)
            elif len(surf) > 0 and isinstance(surf[0], str):
                self.concatenate([Surface(surf=name) for name in surf])
            else:
                self.vertices = np.array(surf[1], dtype=np.float64, copy=True)
                self.faces = np.array(surf[0], dtype=np.int64, copy=True)
                self.component = np.zeros(self.faces.shape[0], dtype=int)
                if weights is None:
                    self.weights = np.ones(self.vertices.shape[0], dtype=int)
                    self.face_weights = np.ones(self.faces.shape[0], dtype=int)
                else:
                    self.updateWeights(weights)
                self.computeCentersAreas()
        elif isinstance(surf, str):
            self.read(surf)
        elif isinstance(surf, Surface):
            self.vertices = np.copy(surf.vertices)
            self.surfel = np.copy(surf.surfel)
            self.faces = np.copy(surf.faces)
            self.centers = np.copy(surf.centers)
            self.component = np.copy(surf.component)
            if weights is None:
                self.weights = np.copy(surf.weights)
                self.face_weights = np.copy(surf.face_weights)
            else:
                self.updateWeights(weights)
            self._topo = surf._topo
        else:
            self.vertices = np.empty(0)
            self.centers = np.empty(0)
            self.faces = np.empty(0, dtype=np.int64)
            self.surfel = np.empty(0)
            self.component = np.empty(0)
            self.weights = np.empty(0)
            self.face_weights = np.empty(0)

    # The edge tables (edges, edgeFaces, faceEdges, bdry, vertices_bdry) are
    # built by getEdges on first access, so a matching that only needs to know
    # whether a surface is closed (is_closed) does not pay for them.
    _EDGE_TABLES = ('edges', 'edgeFaces', 'faceEdges', 'bdry', 'vertices_bdry')

    def __getattr__(self, name):
        if name in Surface._EDGE_TABLES and 'faces' in self.__dict__:
            if self.faces.ndim != 2:               # an empty Surface() has no tables
                return None
            self.getEdges()
            return self.__dict__[name]
        raise AttributeError(name)

    # ------------------------------------------------------------------ I/O
    def read(self, filename):
        """Dispatch on the extension (base/surfaces.py:303-323)."""
        ext = os.path.splitext(filename)[1]
        if ext in ('.byu', '.g'):
            self.readbyu(filename)
        elif ext == '.off':
            self.readOFF(filename)
        elif ext == '.vtk':
            self.readVTK(filename)
        elif ext == '.stl':
            self.readSTL(filename)
        elif ext == '.obj':
            self.readOBJ(filename)
        else:
            raise NameError('Unknown Surface Extension: ' + filename)

    def _set_mesh(self, v, f, w=None, component=None):
        self.vertices = np.asarray(v, dtype=np.float64)
        self.faces = np.asarray(f, dtype=np.int64)
        self.weights = np.ones(self.vertices.shape[0]) if w is None else np.asarray(w, dtype=np.float64)
        self.component = np.zeros(self.faces.shape[0], dtype=int) if component is None else component
        self._topo = None
        self.computeCentersAreas()

    def readVTK(self, fileName):
        """Legacy ASCII VTK: vertices, faces, 'weights' point scalars and
        'labels' cell scalars (the face components), base/surfaces.py:1880-1926."""
        v, f, w, lab = vtk_io.read_polydata(fileName, with_labels=True)
        self._set_mesh(v, f, w, lab)

    def readbyu(self, byufile):
        """base/surfaces.py:1697-1747"""
        self._set_mesh(*vtk_io.read_byu(byufile))

    def readOFF(self, offfile):
        """base/surfaces.py:1658-1695"""
        self._set_mesh(*vtk_io.read_off(offfile))

    def readOBJ(self, fileName):
        """base/surfaces.py:1948-1978"""
        self._set_mesh(*vtk_io.read_obj(fileName))

    def readSTL(self, fileName):
        """base/surfaces.py:1981-2009"""
        self._set_mesh(*vtk_io.read_stl(fileName))

    def saveVTK(self, fileName, scalars=None, normals=None, tensors=None, scal_name='scalars',
                vectors=None, vect_name='vectors'):
        """base/surfaces.py:1794-1812"""
        vtk_io.write_polydata(fileName, self.vertices, self.faces, weights=self.weights,
                              scalars=scalars, scal_name=scal_name, vectors=vectors, vect_name=vect_name,
                              normals=normals, tensors=tensors, component=self.component)

    def saveVTK2(self, fileName, vtkFields=None):
        """base/surfaces.py:1815-1877"""
        vtk_io.write_polydata2(fileName, self.vertices, self.faces, self.component, vtkFields)

    def savebyu(self, byufile):
        """base/surfaces.py:1750-1787"""
        vtk_io.write_byu(byufile, self.vertices, self.faces)

    def saveOFF(self, offfile):
        vtk_io.write_off(offfile, self.vertices, self.faces)

    def concatenate(self, fvl):
        nv = 0
        vs, fs, ws, cs = [], [], [], []
        for k, fv in enumerate(fvl):
            vs.append(fv.vertices)
            fs.append(fv.faces + nv)
            ws.append(fv.weights)
            cs.append(np.full(fv.faces.shape[0], k, dtype=int))
            nv += fv.vertices.shape[0]
        self.vertices = np.concatenate(vs)
        self.faces = np.concatenate(fs).astype(np.int64)
        self.weights = np.concatenate(ws)
        self.component = np.concatenate(cs)
        self.computeCentersAreas()

    # ------------------------------------------------------------- geometry
    def computeCentersAreas(self):
        """base/surfaces.py:326-335"""
        x1 = self.vertices[self.faces[:, 0], :]
        x2 = self.vertices[self.faces[:, 1], :]
        x3 = self.vertices[self.faces[:, 2], :]
        self.centers = (x1 + x2 + x3) / 3
        self.surfel = np.cross(x2 - x1, x3 - x1) / 2
        w = self.weights
        self.face_weights = (w[self.faces[:, 0]] + w[self.faces[:, 1]] + w[self.faces[:, 2]]) / 3
        self._dev = None

    def updateWeights(self, w0):
        self.weights = np.copy(w0)
        w = self.weights
        self.face_weights = (w[self.faces[:, 0]] + w[self.faces[:, 1]] + w[self.faces[:, 2]]) / 3
        self._dev = None

    def updateVertices(self, x0):
        """base/surfaces.py:345-351"""
        self.vertices = np.array(x0, dtype=np.float64, copy=True)
        x1 = self.vertices[self.faces[:, 0], :]
        x2 = self.vertices[self.faces[:, 1], :]
        x3 = self.vertices[self.faces[:, 2], :]
        self.centers = (x1 + x2 + x3) / 3
        self.surfel = np.cross(x2 - x1, x3 - x1) / 2
        self._dev = None

    def computeVertexArea(self, simpleMethod=False):
        """Vertex areas AV [nv] and face areas AF [nf] (base/surfaces.py:364-440).

        simpleMethod: each corner gets a third of its face.  Otherwise the mixed
        Voronoi split of each face: when no angle is obtuse, corner i gets
        (|e_ij|^2 cot k + |e_ik|^2 cot j) / 8, the part of the face nearer to it
        than to the other corners; when the angle at corner i is obtuse, each
        other corner j gets the right triangle that the perpendicular bisector
        of edge ij cuts off the face, |e_ij|^2 / (8 cot j), and corner i keeps
        the rest.  Either way the three parts add up to the face area."""
        V, F = self.vertices, self.faces
        x12 = V[F[:, 1]] - V[F[:, 0]]
        x13 = V[F[:, 2]] - V[F[:, 0]]
        x23 = V[F[:, 2]] - V[F[:, 1]]
        AF = np.sqrt((np.cross(x12, x13) ** 2).sum(axis=1)) / 2
        if simpleMethod:
            u1 = u2 = u3 = AF / 3
        else:
            l12, l13, l23 = (x12 ** 2).sum(1), (x13 ** 2).sum(1), (x23 ** 2).sum(1)
            n12, n13, n23 = np.sqrt(l12), np.sqrt(l13), np.sqrt(l23)
            with np.errstate(divide='ignore', invalid='ignore'):
                c1 = (x12 * x13).sum(1) / (n12 * n13)
                c2 = -(x12 * x23).sum(1) / (n12 * n23)
                c3 = (x13 * x23).sum(1) / (n13 * n23)
                cot1, cot2, cot3 = (c / np.sqrt(1 - c ** 2) for c in (c1, c2, c3))
                u1 = (l12 * cot3 + l13 * cot2) / 8
                u2 = (l12 * cot3 + l23 * cot1) / 8
                u3 = (l13 * cot2 + l23 * cot1) / 8
                o1 = c1 < 0
                o2 = ~o1 & (c2 < 0)
                o3 = ~o1 & ~o2 & (c3 < 0)
                u2 = np.where(o1, l12 / (8 * cot2), u2)
                u3 = np.where(o1, l13 / (8 * cot3), u3)
                u1 = np.where(o1, AF - u2 - u3, u1)
                u1 = np.where(o2, l12 / (8 * cot1), u1)
                u3 = np.where(o2, l23 / (8 * cot3), u3)
                u2 = np.where(o2, AF - u1 - u3, u2)
                u1 = np.where(o3, l13 / (8 * cot1), u1)
                u2 = np.where(o3, l23 / (8 * cot2), u2)
                u3 = np.where(o3, AF - u1 - u2, u3)
        AV = np.zeros(V.shape[0])
        np.add.at(AV, F[:, 0], u1)
        np.add.at(AV, F[:, 1], u2)
        np.add.at(AV, F[:, 2], u3)
        lonely = np.flatnonzero(np.abs(AV) < 1e-10)
        if lonely.size:
            logging.info(f'Warning: vertices {lonely[:10].tolist()} have no face; use removeIsolated')
        return AV, AF

    def computeAreaWeightedVertexNormals(self):
        """Sum of the incident faces' area vectors at each vertex (base/surfaces.py:456-465)."""
        self.computeCentersAreas()
        normals = np.zeros(self.vertices.shape)
        for c in range(3):
            np.add.at(normals, self.faces[:, c], self.surfel)
        return normals

    def computeVertexNormals(self):
        """Unit vertex normals: the normalised sum of the incident faces' area
        vectors (base/surfaces.py:442-454)."""
        normals = self.computeAreaWeightedVertexNormals()
        return normals / np.sqrt((normals ** 2).sum(axis=1))[:, None]

    def flipFaces(self):
        """base/surfaces.py:603-605"""
        self.faces = self.faces[:, [0, 2, 1]]
        self._topo = None
        self.computeCentersAreas()

    def surfVolume(self):
        """sum_f det(v[f]) / 6  (base/surfaces.py:1645-1651), each determinant as
        the triple product x1 . (x2 x x3)."""
        x1 = self.vertices[self.faces[:, 0]]
        x2 = self.vertices[self.faces[:, 1]]
        x3 = self.vertices[self.faces[:, 2]]
        return float((x1 * np.cross(x2, x3)).sum() / 6)

    def surfArea(self):
        return float(np.sqrt((self.surfel ** 2).sum(axis=1)).sum())

    def is_closed(self):
        """True when no edge belongs to exactly one face (getEdges' boundary
        test, bdry.max() == 0), from one sort of the undirected edge keys."""
        if 'bdry' in self.__dict__:
            return bool(self.bdry.size == 0 or self.bdry.max() == 0)
        f = self.faces
        if f.shape[0] == 0:
            return True
        a = f.ravel()
        b = f[:, [1, 2, 0]].ravel()
        nv = max(int(self.vertices.shape[0]), int(f.max()) + 1)
        keys = np.sort(np.minimum(a, b).astype(np.int64) * nv + np.maximum(a, b))
        lone = np.ones(keys.shape[0], dtype=bool)       # key equal to neither neighbour
        eq = keys[1:] == keys[:-1]
        lone[1:] &= ~eq
        lone[:-1] &= ~eq
        return not bool(lone.any())

    def getEdges(self):
        """Edges, edge -> faces and boundary flags (sparse; the reference builds
        a dense V x V table, base/surfaces.py:61-116).  One stable sort of an
        int64 key per (face, side), lo * nv + hi, gives the edges in the
        lexicographic order of (lo, hi), each edge's first and last face (in
        face order) and its multiplicity."""
        f = self.faces
        nf = f.shape[0]
        a = np.concatenate([f[:, 0], f[:, 1], f[:, 2]]).astype(np.int64)
        b = np.concatenate([f[:, 1], f[:, 2], f[:, 0]]).astype(np.int64)
        lo, hi = np.minimum(a, b), np.maximum(a, b)
        nv = max(int(self.vertices.shape[0]), int(hi.max()) + 1 if hi.size else 1)
        keys = lo * nv + hi
        order = np.argsort(keys, kind='stable')
        ks = keys[order]
        first = np.ones(len(ks), dtype=bool)
        first[1:] = ks[1:] != ks[:-1]
        start = np.flatnonzero(first)
        cnt = np.diff(np.append(start, len(ks)))
        uk = ks[start]
        self.edges = np.stack([uk // nv, uk % nv], 1).astype(f.dtype)
        inv = np.empty(len(ks), dtype=np.int64)
        inv[order] = np.cumsum(first) - 1
        self.faceEdges = inv.reshape(3, nf).T
        fid = np.tile(np.arange(nf), 3)
        self.edgeFaces = -np.ones((len(uk), 2), dtype=int)
        self.edgeFaces[:, 0] = fid[order[start]]
        two = cnt >= 2                  # second face: the last one in face order
        self.edgeFaces[two, 1] = fid[order[start[two] + cnt[two] - 1]]
        self.bdry = (cnt == 1).astype(int)
        vb = np.zeros(self.vertices.shape[0], dtype=int)
        vb[self.edges[cnt == 1].ravel()] = 1
        self.vertices_bdry = vb
        return self.edges

    # ---------------------------------------------------------- device view
    def topology(self):
        """int32 faces and CSR adjacency on the device (cached per topology)."""
        import torch
        from . import _backend as B
        if self._topo is None or self._topo[0].shape[0] != self.faces.shape[0]:
            f32 = np.ascontiguousarray(self.faces, dtype=np.int32)
            nv = self.vertices.shape[0]
            ptr = np.zeros(nv + 1, dtype=np.int32)
            idx = np.zeros(3 * f32.shape[0], dtype=np.int32)
            B.check(B.LIB.lddmm_surface_adjacency(f32.ctypes.data, f32.shape[0], nv, ptr.ctypes.data,
                                                  idx.ctypes.data), "surface_adjacency")
            self._topo = (B.upload(f32), B.upload(ptr), B.upload(idx))
        return self._topo

    def device(self, vertices=None):
        """LddmmSurface view (+ keep-alive tensors).  `vertices` may be a device
        tensor [nv, 3] that replaces this surface's vertices for the call."""
        from . import _backend as B
        faces, ptr, idx = self.topology()
        if vertices is None:
            if self._dev is None or self._dev[0] is None:
                w_ = self._dev[1] if self._dev is not None else B.to_dev(np.asarray(self.weights, dtype=np.float64))
                self._dev = (B.to_dev(self.vertices), w_)
            v, w = self._dev
        else:
            v = B.to_dev(vertices)
            if self._dev is None:
                self._dev = (None, B.to_dev(np.asarray(self.weights, dtype=np.float64)))
            w = self._dev[1]
        s = B.LddmmSurface()
        s.vertices = v.data_ptr()
        s.nv = v.shape[0]
        s.faces = faces.data_ptr()
        s.nf = faces.shape[0]
        s.vweights = w.data_ptr() if not np.all(self.weights == 1) else None
        s.adj_ptr = ptr.data_ptr()
        s.adj_idx = idx.data_ptr()
        return s, (v, w, faces, ptr, idx)
