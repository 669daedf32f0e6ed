"""Weighted point sets with the reference's data layout (py-lddmm
base/pointSets.py:8-78): `PointSet(data=array | PointSet | filename,
weights=None, maxPoints=None)` holds `vertices` [N, d] and `weights` [N]
(ones by default; OBJ point clouds get the reference's k-nearest-neighbour
weights).  The module functions savePoints, saveTrajectories, readVector,
loadlmk and savelmk restate base/pointSets.py:80-236.  Plotting is not part
of this package."""
import os

import numpy as np

from . import vtk_io
from .vtk_io import savePoints, saveTrajectories, loadlmk, savelmk  # noqa: F401  (reference names)


def readVector(filename):
    """base/pointSets.py:80-101"""
    return vtk_io.read_vector(filename)


class PointSet:
    def __init__(self, data=None, weights=None, maxPoints=None):
        if isinstance(data, PointSet):
            self.vertices = np.copy(data.vertices)
            self.weights = np.copy(data.weights) if weights is None else np.asarray(weights, dtype=np.float64)
        elif isinstance(data, np.ndarray):
            self.vertices = np.array(data, dtype=np.float64)
            self.weights = np.ones(data.shape[0]) if weights is None else np.asarray(weights, dtype=np.float64)
        elif isinstance(data, str):
            self.read(data, maxPoints=maxPoints)
            if weights is not None:
                self.weights = np.asarray(weights, dtype=np.float64)
        else:
            self.vertices = np.empty(0)
            self.weights = np.empty(0)

    @property
    def points(self):
        return self.vertices

    def updateVertices(self, pts):
        self.vertices = np.copy(pts)

    def set_weights(self, knn=5):
        """pi eps^2 / #(neighbours within eps), eps = mean distance to the
        knn-th neighbour (base/pointSets.py:41-45)."""
        from scipy.spatial.distance import pdist, squareform
        d = np.sort(squareform(pdist(self.vertices)), axis=1)
        eps = d[:, knn].mean() + 1e-10
        self.weights = np.pi * eps ** 2 / (d < eps).sum(axis=1)

    def read(self, filename, maxPoints=None):
        """.obj point clouds (vertices, optional random subset of maxPoints,
        knn weights) or the 'N dim' vector format (base/pointSets.py:48-72)."""
        ext = os.path.splitext(filename)[1]
        if ext in ('.obj', '.OBJ'):
            v, _ = vtk_io.read_obj(filename)
            if maxPoints is not None and maxPoints < v.shape[0]:
                v = v[np.random.default_rng().choice(v.shape[0], maxPoints, replace=False)]
            self.vertices = v
            self.set_weights(5)
        else:
            v, w = vtk_io.read_vector(filename)
            self.vertices, self.weights = v, w

    def saveVTK(self, filename):
        self.save(filename)

    def save(self, filename):
        """base/pointSets.py:75-78"""
        savePoints(filename, self.vertices, scalars={'weights': np.asarray(self.weights).ravel()})
