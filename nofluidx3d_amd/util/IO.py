"""Frame output (reference ``nofluidx3d/util/IO.py:13-39``): read positions back, stop on NaN/Inf,
attach per-tetra ``vonMises``/``pressure`` (pressure written with a minus sign, as the reference
does), optionally ``volumeChange``, and save ``<vtkfilespath><time>.vtk``."""

import numpy as np


def _banner(what):
    print("\n\n")
    print("+---------------------------------------------------------------------------+")
    print(f"| !!! Encountered {what} when reading positions from VRAM, exiting program !!! |")
    print("+---------------------------------------------------------------------------+")


def writeVTK(simulation, writeStress=True, writePressure=True, writeVolumeChange=False):
    positions = simulation.points.readFromGPU()
    positions = positions.reshape((3, simulation.numPoints)).transpose()
    if np.any(np.isnan(positions)):
        _banner("NaN")
        exit()
    if np.any(np.isinf(positions)):
        _banner("INF")
        exit()
    simulation.cell.mesh.points = positions
    if writeStress:
        stress = simulation.vonMises.readFromGPU()
        simulation.cell.addTetraData("vonMises", stress)
    if writePressure:
        pressure = -simulation.pressure.readFromGPU()
        simulation.cell.addTetraData("pressure", pressure)
    if writeVolumeChange:
        simulation.cell.calcVolumeChange()
    simulation.cell.mesh.save(simulation.vtkfilespath + f"{simulation.time}.vtk")
