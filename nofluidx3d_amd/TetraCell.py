"""Tetrahedral cell mesh with its FP64 reference quantities (reference ``nofluidx3d/TetraCell.py``).

``TetraCell(path, radius=18.0, recenter=True)`` reads a legacy-VTK unstructured grid (or takes an
in-memory ``UnstructuredGrid``), translates point 0 to the origin, scales so that point 1's
x-coordinate has magnitude ``radius`` (``TetraCell.py:15-20``), and computes the edge vectors
``[p1-p0, p2-p0, p3-p0]`` (T, 9) and the signed volumes (``TetraCell.py:26-41``) in FP64.
"""

from __future__ import annotations

import numpy as np

from nofluidx3d_amd.util import vtk


class TetraCell:
    def __init__(self, path, radius=18.0, recenter=True):
        if isinstance(path, vtk.UnstructuredGrid):
            self.mesh = path.copy()
        else:
            self.mesh = vtk.read(path)
        if recenter:
            transvec = -self.mesh.points[0]  # point 0 of the reference mesh marks the origin
            self.mesh.translate(transvec)
            radiusMesh = np.abs(self.mesh.points[1][0])  # x-value of point 1
            self.mesh.scale(radius / radiusMesh, inplace=True)
        self.points = self.mesh.points
        self.radius = radius
        self.tetras = self.mesh.cells_dict[vtk.VTK_TETRA]
        R1, R2, R3 = self.calcEdgeVectors(self.points)
        self.edgeVectors = np.stack([R1, R2, R3], 1).reshape(len(R1), 9)
        self.volumes = self.calcVolumes(R1, R2, R3)

    @classmethod
    def from_arrays(cls, points, tets, radius=18.0, recenter=True):
        return cls(vtk.from_tetra(points, tets), radius=radius, recenter=recenter)

    def calcEdgeVectors(self, points):
        p0 = points[self.tetras[:, 0]]
        p1 = points[self.tetras[:, 1]]
        p2 = points[self.tetras[:, 2]]
        p3 = points[self.tetras[:, 3]]
        return p1 - p0, p2 - p0, p3 - p0

    def calcVolumes(self, R1, R2, R3):
        return np.einsum("ij,ij->i", R1, np.cross(R2, R3)) / 6

    def scale(self, scaling):
        self.mesh.scale(scaling, inplace=True)
        self.points = self.mesh.points
        R1, R2, R3 = self.calcEdgeVectors(self.points)
        self.edgeVectors = np.stack([R1, R2, R3], 1).reshape(len(R1), 9)
        self.volumes = self.calcVolumes(R1, R2, R3)

    def addTetraData(self, name, dataArray):
        # pad at the front to the mesh's cell count (non-tetra cells come first), TetraCell.py:53-56
        dataArrayPadded = np.pad(dataArray, pad_width=(self.mesh.n_cells - dataArray.size, 0))
        self.mesh.cell_data.set_array(dataArrayPadded, name)

    def calcVolumeChange(self):
        curVolumes = self.calcVolumes(*self.calcEdgeVectors(self.mesh.points))
        self.addTetraData("volumeChange", curVolumes / self.volumes - 1)
