"""Sub-blocks of a generated structured mesh, for comparing operator rows of a full-size
GPU context with the oracle assembled on a small piece of the same mesh.

A row of C, D, G or of the raw Laplacian at node i depends only on the elements that
contain i, so the oracle on a voxel sub-block holds the full mesh's rows for the nodes
whose 8 adjacent voxels all lie in the sub-block (with C and G scaled by the full
magnet volume, passed as ``V``).
"""
import numpy as np

from mammos_mumag_amd import generate


def submesh(mesh, lo, hi):
    """Voxels [lo[a], hi[a]) per axis (x, y, z) of ``mesh`` (a generate.StructuredMesh).
    Returns (sub, full_index, interior): the sub-block as a StructuredMesh, the full-mesh
    index of each sub node, and the mask of sub nodes whose 8 voxels are in the block."""
    nx, ny, nz = mesh.shape
    sub = generate.StructuredMesh(mesh.xs[lo[0]:hi[0] + 1], mesh.ys[lo[1]:hi[1] + 1], mesh.zs[lo[2]:hi[2] + 1],
                                  mesh.voxel_tag[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]].copy(),
                                  mesh.tags_props, None, {})
    k, j, i = np.meshgrid(np.arange(lo[2], hi[2] + 1), np.arange(lo[1], hi[1] + 1), np.arange(lo[0], hi[0] + 1),
                          indexing="ij")
    full = (i + nx * (j + ny * k)).ravel().astype(np.int64)
    inner = ((i > lo[0]) & (i < hi[0]) & (j > lo[1]) & (j < hi[1]) & (k > lo[2]) & (k < hi[2])).ravel()
    return sub, full, inner


def dirichlet_mask(mesh):
    """Nodes on the six faces of the air box (u = 0, scripts/trafo.py:4-16)."""
    nx, ny, nz = mesh.shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return ((i == 0) | (i == nx - 1) | (j == 0) | (j == ny - 1) | (k == 0) | (k == nz - 1)).ravel()


def uniform_anisotropy_energy(mesh, mvec, V):
    """Closed form of E = m.C m / 2 for a uniform unit m over the magnet: the exchange of a
    uniform state is zero, and the anisotropy term integrates to
    -(mu0 / V) sum_voxels K1 vol (m . u)^2 with mu0 = 4 pi 1e-7 (scripts/materials.py:21-34)."""
    mu0 = 4e-7 * np.pi
    h = mesh.meta["h"]
    tags, counts = np.unique(mesh.voxel_tag, return_counts=True)
    e = 0.0
    for t, cnt in zip(tags.tolist(), counts.tolist()):
        p = mesh.tags_props[t]
        if p["Js"] <= 0:
            continue
        u = np.array([np.sin(p["theta"]) * np.cos(p["phi"]), np.sin(p["theta"]) * np.sin(p["phi"]), np.cos(p["theta"])])
        e += p["K1"] * cnt * h ** 3 * float(np.dot(u, mvec)) ** 2
    return -mu0 * e / V


def magnet_volume_and_moment(mesh):
    """V (magnet volume, nm^3) and sum_voxels Js vol (T nm^3), from the voxel tags."""
    h = mesh.meta["h"]
    tags, counts = np.unique(mesh.voxel_tag, return_counts=True)
    V = M = 0.0
    for t, cnt in zip(tags.tolist(), counts.tolist()):
        p = mesh.tags_props[t]
        if p["Js"] > 0:
            V += cnt * h ** 3
            M += p["Js"] * cnt * h ** 3
    return V, M
