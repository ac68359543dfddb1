"""Utilities: Bridge (host/device buffer pair), IO (frame output), vtk (legacy VTK I/O)."""
