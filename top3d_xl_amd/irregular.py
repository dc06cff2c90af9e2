"""Seeded irregular voxel domains with boundary conditions (the irregular
GLOBAL workload: 384x768x384 bounding box, ~30% occupied).

The domain is built as follows (all choices are this generator's; the
reference only consumes such models through .TopVoxel files,
TOP3D_XL.m:448-528):

* bounding box (nely, nelx, nelz) = (384, 768, 384) by default;
* union U of `n_ellipsoids` ellipsoids with semi-axes uniform in [20, 120]
  voxels (scaled with the box), uniformly random orientations, and centres
  uniform in the box shrunk by half the ellipsoid's largest semi-axis on
  every face (so most of each ellipsoid lies inside the box);
* the solid shell S: the bounding box with a centred cuboid hollow removed,
  the hollow spanning the middle `shell_inner` fraction of each axis (a
  thick-walled box; the default 0.4 removes 6.4% of the volume), which
  brings the occupancy to ~0.3 (0.25-0.35 over seeds);
* the domain is the largest 6-connected component of U intersected with S
  (face connectivity, so every kept voxel shares a face with another and
  the FE mesh is connected through faces).
* fixations: the surface nodes (nodes of a solid voxel that also touch a
  void voxel or the box boundary) in the lowest 5% of the domain's z extent,
  all three DOFs.  Supports sit on the surface: the reference applies a
  fixation only at boundary nodes (Solving_SetupKeWithFixedDOFs intersects
  fixingCond_ with nodesOnBoundary, TOP3D_XL.m:1369);
* load: a total force of magnitude 1, direction drawn from the seed, split
  evenly over the surface nodes (nodes of a solid voxel that also touch a
  void voxel or the box boundary) in the top 5% of the domain's z extent.

Node ids are 0-based on the unpadded (nely+1)x(nelx+1)x(nelz+1) grid, the
convention of t3x_create and of .TopVoxel files.
"""
from __future__ import annotations

import numpy as np

DEFAULT_SHAPE = (384, 768, 384)


def _ellipsoid_union(shape, n, axes_range, rng):
    ny, nx, nz = shape
    U = np.zeros(shape, dtype=bool)
    for _ in range(n):
        u = rng.uniform(0.0, 1.0, size=3)
        a = rng.uniform(axes_range[0], axes_range[1], size=3)
        m = np.minimum(0.5 * a.max(), np.array(shape) / 2.0)
        c = m + u * (np.array(shape) - 2.0 * m)
        # uniformly random rotation from a normalised Gaussian quaternion
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        r = a.max()
        lo = np.maximum(np.floor(c - r).astype(int), 0)
        hi = np.minimum(np.ceil(c + r).astype(int) + 1, [ny, nx, nz])
        if np.any(hi <= lo):
            continue
        gy = np.arange(lo[0], hi[0]) + 0.5 - c[0]
        gx = np.arange(lo[1], hi[1]) + 0.5 - c[1]
        for kz in range(lo[2], hi[2]):
            dz = kz + 0.5 - c[2]
            DY, DX = np.meshgrid(gy, gx, indexing="ij")
            P = np.stack([DY, DX, np.full_like(DY, dz)], axis=-1) @ R  # body-frame coordinates
            inside = ((P / a) ** 2).sum(axis=-1) <= 1.0
            U[lo[0]:hi[0], lo[1]:hi[1], kz] |= inside
    return U


def _shell(shape, inner):
    # normalised coordinates in [-1, 1]; the hollow is max(|y|,|x|,|z|) < inner
    c = [np.abs((np.arange(n) + 0.5) / n * 2 - 1) for n in shape]
    hollow = (c[0][:, None, None] < inner) & (c[1][None, :, None] < inner) & (c[2][None, None, :] < inner)
    return ~hollow


def _largest_component(mask):
    from scipy import ndimage
    lab, n = ndimage.label(mask)  # 6-connectivity (face neighbours)
    if n == 0:
        return mask
    counts = np.bincount(lab.ravel())
    counts[0] = 0
    return lab == int(np.argmax(counts))


def _node_id(jy, jx, jz, ny, nx):
    return jy + (ny + 1) * (jx + (nx + 1) * jz)


def generate(seed=0, shape=DEFAULT_SHAPE, n_ellipsoids=40, axes=(20.0, 120.0), shell_inner=0.4, slab=0.05):
    """Returns dict(voxels, fix_nodes, fix_state, load_nodes, load_vec, direction, occupancy)."""
    rng = np.random.default_rng(seed)
    ny, nx, nz = shape
    scale = min(shape) / 384.0
    U = _ellipsoid_union(shape, n_ellipsoids, (axes[0] * scale, axes[1] * scale), rng)
    vox = _largest_component(U & _shell(shape, shell_inner))
    zs = np.flatnonzero(vox.any(axis=(0, 1)))
    z0, z1 = int(zs[0]), int(zs[-1]) + 1  # cell layers [z0, z1)
    ext = z1 - z0
    nslab = max(1, int(round(slab * ext)))
    # node masks: nodes of solid cells
    nodes = np.zeros((ny + 1, nx + 1, nz + 1), dtype=bool)
    for dy in (0, 1):
        for dx in (0, 1):
            for dz in (0, 1):
                nodes[dy:dy + ny, dx:dx + nx, dz:dz + nz] |= vox
    # surface nodes: touch a void cell or the box boundary
    pad = np.pad(vox, 1, constant_values=False)
    allsolid = np.ones((ny + 1, nx + 1, nz + 1), dtype=bool)
    for dy in (0, 1):
        for dx in (0, 1):
            for dz in (0, 1):
                allsolid &= pad[dy:dy + ny + 1, dx:dx + nx + 1, dz:dz + nz + 1]
    surface = nodes & ~allsolid
    jz = np.arange(nz + 1)
    fixed = surface & (jz <= z0 + nslab)[None, None, :]
    loaded = surface & (jz >= z1 - nslab)[None, None, :] & ~fixed
    fy, fx, fz = np.nonzero(fixed)
    fix_nodes = _node_id(fy, fx, fz, ny, nx)
    order = np.argsort(fix_nodes)
    fix_nodes = fix_nodes[order].astype(np.int32)
    ly, lx, lz = np.nonzero(loaded)
    load_nodes = _node_id(ly, lx, lz, ny, nx)
    load_nodes = np.sort(load_nodes).astype(np.int32)
    d = rng.standard_normal(3)
    d /= np.linalg.norm(d)
    load_vec = np.tile(d / len(load_nodes), (len(load_nodes), 1))
    return dict(voxels=vox, fix_nodes=fix_nodes, fix_state=np.ones((len(fix_nodes), 3), np.uint8),
                load_nodes=load_nodes, load_vec=load_vec, direction=d, occupancy=float(vox.mean()),
                z_extent=(z0, z1))
