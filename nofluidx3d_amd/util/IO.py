"""Frame output (reference ``nofluidx3d/util/IO.py:13-39``).

``writeVTK`` stores the current node positions in the cell's mesh, attaches the per-tetra
``vonMises`` and ``pressure`` (pressure with the reference's sign: minus the mean stress) and
optionally ``volumeChange``, and saves ``<vtkfilespath><time>.vtk``.  Non-finite positions end the
program with the reference's message, as there is nothing left to simulate.
"""

import sys

import numpy as np


def _abort(kind):
    rule = "+" + "-" * 75 + "+"
    print("\n\n")
    print(rule)
    print(f"| !!! Encountered {kind} when reading positions from VRAM, exiting program !!! |")
    print(rule)
    sys.exit()


def writeVTK(simulation, writeStress=True, writePressure=True, writeVolumeChange=False):
    x = simulation.positions()
    finite = np.isfinite(x)
    if not finite.all():
        _abort("NaN" if np.isnan(x).any() else "INF")
    cell = simulation.cell
    cell.mesh.points = x
    outputs = ((writeStress, "vonMises", simulation.vonMises, 1.0), (writePressure, "pressure", simulation.pressure, -1.0))
    for wanted, name, bridge, sign in outputs:
        if wanted:
            cell.addTetraData(name, sign * bridge.readFromGPU())
    if writeVolumeChange:
        cell.calcVolumeChange()
    cell.mesh.save(f"{simulation.vtkfilespath}{simulation.time}.vtk")
