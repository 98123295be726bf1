"""Mesh helpers: the structured Kuhn-split cuboids of the benchmark configurations.

The reference meshes arbitrary geometries with its own mesher (out of scope);
the configurations of this path all use cuboids of cubes, each split into the
6 Kuhn tetrahedra that share the cube's main diagonal.  Node numbering is
x-fastest: index = i + (nx+1)*(j + (ny+1)*k), coordinates in nm.
"""
import itertools

import numpy as np

# each Kuhn simplex walks from corner (0,0,0) to (1,1,1) along one axis order
_PERMS = list(itertools.permutations(range(3)))


def kuhn_cuboid(nx, ny, nz, h, origin=(0.0, 0.0, 0.0)):
    """Return (coords (N,3) float64, tets (6*nx*ny*nz, 4) int32)."""
    nx, ny, nz = int(nx), int(ny), int(nz)
    if min(nx, ny, nz) < 1:
        raise ValueError("a cuboid needs at least one cube per direction")
    sx, sy = nx + 1, ny + 1
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1),
                          indexing="ij")
    coords = np.stack([i.ravel() * h, j.ravel() * h, k.ravel() * h], axis=1).astype(np.float64)
    coords += np.asarray(origin, dtype=np.float64)
    ck, cj, ci = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    base = np.stack([ci.ravel(), cj.ravel(), ck.ravel()], axis=1)
    tets = np.zeros((len(base), 6, 4), dtype=np.int64)
    for p, perm in enumerate(_PERMS):
        c = base.copy()
        tets[:, p, 0] = c[:, 0] + sx * (c[:, 1] + sy * c[:, 2])
        for s in range(3):
            c[:, perm[s]] += 1
            tets[:, p, s + 1] = c[:, 0] + sx * (c[:, 1] + sy * c[:, 2])
    return coords, tets.reshape(-1, 4).astype(np.int32)


def cuboid_sizes(nx, ny, nz):
    """(nodes, tets, boundary nodes, boundary triangles) of a Kuhn cuboid."""
    n = (nx + 1) * (ny + 1) * (nz + 1)
    interior = max(nx - 1, 0) * max(ny - 1, 0) * max(nz - 1, 0)
    return n, 6 * nx * ny * nz, n - interior, 4 * (nx * ny + ny * nz + nz * nx)
