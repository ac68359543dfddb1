"""Seeded tetrahedral cell meshes (the reference ships no meshes; it reads them from VTK).

A cell is a ball of radius ``radius`` (µm).  Points are a cubic lattice clipped to the ball,
jittered uniformly by ``±jitter·spacing`` per coordinate from a PCG64 stream, then
Delaunay-tetrahedralised (``scipy.spatial.Delaunay``).  Low-quality tetrahedra (slivers from
near-cospherical lattice cells and flat hull elements) are dropped, because the explicit
integrator runs at dt = 1 with unit node mass and a sliver's stiffness grows like 1/quality.

Node 0 is the exact centre and node 1 the exact +x pole, un-jittered: the reference's
``TetraCell`` recentres on point 0 and scales so that ``|x(point 1)| = RadiusSIM``
(reference ``nofluidx3d/TetraCell.py:15-20``).  Every tetrahedron is oriented positively
(det[p1-p0, p2-p0, p3-p0] > 0, the VTK_TETRA convention); the reference's force formula
``-V0·J·r⁻¹·σ`` (``kernels/Interactions/tetra.cl:159-162``) is only restoring for that
orientation.  Nodes (after 0 and 1) and tetrahedra are sorted along a Morton curve so that
the per-tetra node gathers on the GPU hit nearby cache lines.
"""

from __future__ import annotations

import numpy as np

DEFAULT_MIN_QUALITY = 0.1


def tetra_quality(points: np.ndarray, tets: np.ndarray) -> np.ndarray:
    """Normalised volume 6·sqrt(2)·|V| / l_rms³ (1 for a regular tetrahedron, 0 for a flat one)."""
    p = points[tets]
    e = (
        p[:, 1] - p[:, 0],
        p[:, 2] - p[:, 0],
        p[:, 3] - p[:, 0],
        p[:, 2] - p[:, 1],
        p[:, 3] - p[:, 1],
        p[:, 3] - p[:, 2],
    )
    l2 = sum((x * x).sum(1) for x in e) / 6.0
    vol = np.einsum("ij,ij->i", e[0], np.cross(e[1], e[2])) / 6.0
    return 6.0 * np.sqrt(2.0) * np.abs(vol) / l2**1.5


def signed_volumes(points: np.ndarray, tets: np.ndarray) -> np.ndarray:
    p = points[tets]
    return (
        np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]))
        / 6.0
    )


def _morton_keys(xyz: np.ndarray, bits: int = 10) -> np.ndarray:
    lo = xyz.min(0)
    span = np.maximum(xyz.max(0) - lo, 1e-30)
    q = np.clip(((xyz - lo) / span * ((1 << bits) - 1)).astype(np.uint64), 0, (1 << bits) - 1)
    key = np.zeros(len(xyz), dtype=np.uint64)
    for b in range(bits):
        for d in range(3):
            key |= ((q[:, d] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + d)
    return key


def generate_sphere_mesh(
    radius: float,
    spacing: float,
    seed: int,
    jitter: float = 0.1,
    min_quality: float = DEFAULT_MIN_QUALITY,
    morton: bool = True,
):
    """Return ``(points (N,3) float64 [µm], tets (T,4) int64)`` for a seeded ball mesh.

    The lattice spacing is adjusted to ``radius / round(radius / spacing)`` so that the +x pole
    is a lattice point; for the benchmark cells (8 µm / 0.8 µm, 10 µm / 0.2 µm) this is exact.
    """
    from scipy.spatial import Delaunay

    m = max(int(round(radius / spacing)), 2)
    s = radius / m
    g = np.arange(-m, m + 1, dtype=np.float64) * s
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    P = P[np.einsum("ij,ij->i", P, P) <= radius * radius * (1 + 1e-12)]
    # centre and +x pole first, exactly on their lattice sites
    is_c = np.all(np.abs(P) < 1e-9 * s, axis=1)
    is_p = (np.abs(P[:, 0] - radius) < 1e-9 * s) & (np.abs(P[:, 1]) < 1e-9 * s) & (
        np.abs(P[:, 2]) < 1e-9 * s
    )
    rest = P[~(is_c | is_p)]
    rng = np.random.Generator(np.random.PCG64(seed))
    rest = rest + rng.uniform(-jitter * s, jitter * s, rest.shape)
    P = np.concatenate([np.zeros((1, 3)), np.array([[radius, 0.0, 0.0]]), rest], 0)

    tets = Delaunay(P).simplices.astype(np.int64)
    vol = signed_volumes(P, tets)
    neg = vol < 0
    tets[neg, 2], tets[neg, 3] = tets[neg, 3].copy(), tets[neg, 2].copy()
    tets = tets[tetra_quality(P, tets) >= min_quality]

    used = np.zeros(len(P), dtype=bool)
    used[tets.ravel()] = True
    used[:2] = True
    if morton:
        keys = _morton_keys(P[2:])
        order_rest = 2 + np.argsort(keys, kind="stable")
        order = np.concatenate([[0, 1], order_rest])
    else:
        order = np.arange(len(P))
    order = order[used[order]]
    remap = -np.ones(len(P), dtype=np.int64)
    remap[order] = np.arange(len(order))
    P = P[order]
    tets = remap[tets]
    if morton:
        tk = _morton_keys(P[tets].mean(1))
        tets = tets[np.argsort(tk, kind="stable")]
    return np.ascontiguousarray(P), np.ascontiguousarray(tets)
