"""The reference cell of a simulation (reference ``nofluidx3d/TetraCell.py``): its tetrahedral mesh
in simulation units and the per-tetra reference quantities the force laws start from.

``TetraCell(path, radius=18.0, recenter=True)`` takes a legacy-VTK file (or an in-memory
``UnstructuredGrid``).  With ``recenter`` the mesh is moved so that node 0 is the origin and scaled
so that node 1 lies at |x| = ``radius`` (``TetraCell.py:15-20``).  Node coordinates are then held
as the FP32 values the device keeps (stored in float64 arrays), so that the FP64 precompute, the
device state and the oracle all start from one set of positions.  ``edgeVectors`` (T, 9: R1, R2,
R3 with R_k = x_k - x_0) and the signed ``volumes`` R1 . (R2 x R3) / 6 follow
(``TetraCell.py:26-41``).
"""

from __future__ import annotations

import numpy as np

from nofluidx3d_amd.util import vtk


def _normalised(grid, radius):
    """node 0 at the origin, |x of node 1| = radius"""
    origin = np.array(grid.points[0], dtype=np.float64)
    grid.translate(-origin)
    grid.scale(radius / abs(float(grid.points[1][0])), inplace=True)


def _device_values(points):
    """the coordinates as the FP32 device state holds them, in float64"""
    return np.asarray(points, dtype=np.float64).astype(np.float32).astype(np.float64)


class TetraCell:
    def __init__(self, path, radius=18.0, recenter=True):
        grid = path.copy() if isinstance(path, vtk.UnstructuredGrid) else vtk.read(path)
        if recenter:
            _normalised(grid, radius)
        grid.points = _device_values(grid.points)
        self.mesh = grid
        self.points = grid.points  # the reference positions (frames later replace mesh.points)
        self.radius = radius
        self.tetras = grid.cells_dict[vtk.VTK_TETRA]
        self._set_reference()

    @classmethod
    def from_arrays(cls, points, tets, radius=18.0, recenter=True):
        return cls(vtk.from_tetra(points, tets), radius=radius, recenter=recenter)

    def _set_reference(self):
        edges = self._edges(self.points)  # (T, 3, 3): edge k of tetra t in edges[t, k]
        self.edgeVectors = edges.reshape(len(edges), 9)
        self.volumes = np.linalg.det(edges) / 6.0

    def _edges(self, points):
        corners = np.asarray(points, dtype=np.float64)[self.tetras]  # (T, 4, 3)
        return corners[:, 1:, :] - corners[:, :1, :]

    def calcEdgeVectors(self, points):
        """(R1, R2, R3), each (T, 3): the edges from node 0 of every tetra at ``points``."""
        e = self._edges(points)
        return e[:, 0], e[:, 1], e[:, 2]

    def calcVolumes(self, R1, R2, R3):
        """signed volumes R1 . (R2 x R3) / 6"""
        return np.linalg.det(np.stack([R1, R2, R3], axis=1)) / 6.0

    def scale(self, scaling):
        self.mesh.scale(scaling, inplace=True)
        self.mesh.points = _device_values(self.mesh.points)
        self.points = self.mesh.points
        self._set_reference()

    def addTetraData(self, name, dataArray):
        """Attach a per-tetra array as cell data; cells before the tetra (the mesh's other cell
        types come first) get zeros (``TetraCell.py:53-56``)."""
        values = np.asarray(dataArray)
        padded = np.zeros(self.mesh.n_cells, dtype=values.dtype)
        padded[self.mesh.n_cells - values.size:] = values
        self.mesh.cell_data.set_array(padded, name)

    def calcVolumeChange(self):
        """``volumeChange`` = V / V0 - 1 at the mesh's current positions."""
        current = np.linalg.det(self._edges(self.mesh.points)) / 6.0
        self.addTetraData("volumeChange", current / self.volumes - 1.0)
