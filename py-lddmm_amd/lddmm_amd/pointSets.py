"""Weighted point sets with the reference's data layout (py-lddmm
base/pointSets.py:8-30): `PointSet(data=array, weights=None)` holds
`vertices` [N, d] and `weights` [N] (ones by default).  File readers and
plotting are not part of this package."""
import numpy as np


class PointSet:
    def __init__(self, data=None, weights=None, maxPoints=None):
        if isinstance(data, PointSet):
            self.vertices = np.copy(data.vertices)
            self.weights = np.copy(data.weights) if weights is None else np.asarray(weights, dtype=np.float64)
        elif isinstance(data, np.ndarray):
            self.vertices = np.array(data, dtype=np.float64)
            self.weights = np.ones(data.shape[0]) if weights is None else np.asarray(weights, dtype=np.float64)
        elif isinstance(data, str):
            raise NotImplementedError("lddmm_amd.pointSets: reading point-set files is not built")
        else:
            self.vertices = np.empty(0)
            self.weights = np.empty(0)

    @property
    def points(self):
        return self.vertices

    def updateVertices(self, pts):
        self.vertices = np.copy(pts)
