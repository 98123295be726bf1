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


def read_nmesh(path):
    """Read an nmesh ASCII file (format version 1.0, src/mesh.ml:6426-6616).

    Returns (coords (N,3), tets (T,4) int32, regions (T,) int32)."""
    with open(path) as f:
        lines = [ln for ln in f if not ln.startswith("#") and ln.strip()]
    pos = 0
    n = int(lines[pos].split()[0]); pos += 1
    coords = np.array([[float(v) for v in lines[pos + i].split()] for i in range(n)])
    pos += n
    t = int(lines[pos].split()[0]); pos += 1
    rows = np.array([[int(v) for v in lines[pos + i].split()] for i in range(t)], dtype=np.int64)
    if coords.ndim != 2 or coords.shape[1] != 3 or rows.shape[1] != 5:
        raise ValueError("%s: only 3-d tetrahedral nmesh files are supported" % path)
    return coords, rows[:, 1:5].astype(np.int32), rows[:, 0].astype(np.int32)


def write_nmesh(path, coords, tets, regions=None):
    """Write an nmesh ASCII file (version 1.0) without surface/periodic data."""
    coords = np.asarray(coords, dtype=np.float64)
    tets = np.asarray(tets)
    regions = np.ones(len(tets), dtype=int) if regions is None else np.asarray(regions)
    with open(path, "w") as f:
        f.write("# PYFEM mesh file version 1.0\n")
        f.write("# dim = 3 \t nodes = %d \t simplices = %d \t surfaces = 0 \t periodic = 0\n"
                % (len(coords), len(tets)))
        f.write("%d\n" % len(coords))
        for c in coords:
            f.write("".join("\t%25.12f " % v for v in c) + "\n")
        f.write("%d\n" % len(tets))
        for r, t in zip(regions, tets):
            f.write("\t%d " % r + "".join("%d\t" % v for v in t) + "\n")
        f.write("0\n0\n")


def cuboid_sizes(nx, ny, nz):
    """(nodes, tets, boundary nodes, boundary triangles) of a Kuhn cuboid."""
    n = (nx + 1) * (ny + 1) * (nz + 1)
    interior = max(nx - 1, 0) * max(ny - 1, 0) * max(nz - 1, 0)
    return n, 6 * nx * ny * nz, n - interior, 4 * (nx * ny + ny * nz + nz * nx)
