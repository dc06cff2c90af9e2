"""The .TopVoxel voxel-model format (version 1.0) of TOP3D_XL.

Reader and writer for the file CreateVoxelFEAmodel parses
(TOP3D_XL.m:448-528).  The format is a whitespace-separated token stream:

    <4 header words> <version word> 1.0
    Resolution: nelx nely nelz
    <2 words> <0|1: density values follow the voxel ids>
    <2 words> numSolidVoxels
        voxel id [density]      x numSolidVoxels  (1-based, column-major (y,x,z))
    Passive elements: n   voxel ids ...
    Fixations: n          node fx fy fz        x n (1-based node id of the
                                                    (nely+1)x(nelx+1)x(nelz+1) grid)
    Loads: n              node Fx Fy Fz        x n
    <2 words> k           k additional load cases, each
        Loads: n idx      node Fx Fy Fz        x n

Node and voxel ids are returned 0-based, with duplicates removed keeping the
first occurrence and rows sorted by id, as the reference does before
AdaptBCExternalMdl (499-518).  The padding to the multigrid resolution
(AdaptBCExternalMdl, 1439-1489) happens inside t3x_create, which takes ids of
the unpadded grid.
"""
from __future__ import annotations

import numpy as np

HEADER = "# TopVoxel voxel model"
VERSION = "1.0"


class FormatError(ValueError):
    pass


def _dedupe_sorted(ids, rows):
    ids = np.asarray(ids, dtype=np.int64)
    if ids.size == 0:
        return ids.astype(np.int32), np.asarray(rows).reshape(0, 3)
    _, first = np.unique(ids, return_index=True)  # first occurrence of each id, ids ascending
    return ids[first].astype(np.int32), np.asarray(rows)[first]


def read_topvoxel(path):
    """Parse a .TopVoxel file; returns a dict with
    voxels (nely, nelx, nelz) bool, density (numSolid,) or None,
    passive (0-based voxel ids), fix_nodes (0-based), fix_state (n x 3 uint8),
    load_nodes (0-based), load_vec (n x 3), extra_loads [(idx, nodes, vec)]."""
    with open(path, "r") as f:
        tok = f.read().split()
    pos = 0

    def take(n=1):
        nonlocal pos
        if pos + n > len(tok):
            raise FormatError("Incompatible Mesh Data Format!")
        v = tok[pos:pos + n]
        pos += n
        return v

    def expect(word):
        if take()[0] != word:
            raise FormatError("Incompatible Mesh Data Format!")

    take(4)  # file head
    take(1)  # version head
    if float(take()[0]) != 1.0:
        raise FormatError("Unsupported Data!")
    expect("Resolution:")
    nelx, nely, nelz = (int(v) for v in take(3))
    take(2)
    with_density = int(take()[0])
    take(2)
    nsolid = int(take()[0])
    if with_density:
        vals = np.array(take(2 * nsolid), dtype=np.float64).reshape(nsolid, 2)
        solid = vals[:, 0].astype(np.int64) - 1
        density = vals[:, 1].copy()
    else:
        solid = np.array(take(nsolid), dtype=np.int64) - 1
        density = None
    expect("Passive")
    expect("elements:")
    npas = int(take()[0])
    passive = np.sort(np.array(take(npas), dtype=np.int64) - 1).astype(np.int32)
    expect("Fixations:")
    nfix = int(take()[0])
    fx = np.array(take(4 * nfix), dtype=np.float64).reshape(nfix, 4)
    expect("Loads:")
    nload = int(take()[0])
    ld = np.array(take(4 * nload), dtype=np.float64).reshape(nload, 4)
    extra = []
    if pos < len(tok):
        take(2)
        nadd = int(take()[0])
        for _ in range(nadd):
            expect("Loads:")
            n = int(take()[0])
            idx = int(take()[0])
            a = np.array(take(4 * n), dtype=np.float64).reshape(n, 4)
            nodes, vec = _dedupe_sorted(a[:, 0].astype(np.int64) - 1, a[:, 1:4])
            extra.append((idx, nodes, vec))
    vox = np.zeros(nelx * nely * nelz, dtype=bool)
    if solid.size and (solid.min() < 0 or solid.max() >= vox.size):
        raise FormatError("Incompatible Mesh Data Format!")
    vox[solid] = True
    fix_nodes, fix_state = _dedupe_sorted(fx[:, 0].astype(np.int64) - 1, fx[:, 1:4])
    load_nodes, load_vec = _dedupe_sorted(ld[:, 0].astype(np.int64) - 1, ld[:, 1:4])
    return dict(voxels=vox.reshape((nely, nelx, nelz), order="F"), density=density, passive=passive,
                fix_nodes=fix_nodes, fix_state=(np.asarray(fix_state) == 1).astype(np.uint8),
                load_nodes=load_nodes, load_vec=np.asarray(load_vec, dtype=np.float64), extra_loads=extra)


def load_cases(model):
    """The load cases of a parsed model in loadingCond_ order: case 1 from
    "Loads:", then each additional case at its index idx (490).  The indices
    must fill 2..n+1: a gap leaves an empty case, which the reference rejects
    with 'No Loads!' (1309)."""
    cases = [(model["load_nodes"], model["load_vec"])]
    extra = sorted(model.get("extra_loads", []), key=lambda t: t[0])
    if [t[0] for t in extra] != list(range(2, 2 + len(extra))):
        raise FormatError("No Loads!")
    cases += [(nodes, vec) for _, nodes, vec in extra]
    return cases


def write_topvoxel(path, voxels, fix_nodes, fix_state, load_nodes, load_vec, passive=None, density=None,
                   extra_loads=()):
    """Write a version-1.0 .TopVoxel file (0-based ids in, 1-based ids out);
    extra_loads: further load cases (nodes, vec), written as cases 2, 3, ..."""
    vox = np.asarray(voxels, dtype=bool)
    nely, nelx, nelz = vox.shape
    solid = np.flatnonzero(vox.reshape(-1, order="F"))
    passive = np.zeros(0, np.int64) if passive is None else np.asarray(passive, dtype=np.int64)
    fix_state = np.asarray(fix_state).reshape(-1, 3)
    load_vec = np.asarray(load_vec, dtype=np.float64).reshape(-1, 3)
    with open(path, "w") as f:
        f.write(f"{HEADER}\nVersion: {VERSION}\n")
        f.write(f"Resolution: {nelx} {nely} {nelz}\n")
        f.write(f"Density values: {1 if density is not None else 0}\n")
        f.write(f"Solid voxels: {solid.size}\n")
        if density is not None:
            for i, d in zip(solid, np.asarray(density, dtype=np.float64)):
                f.write(f"{i + 1} {d:.16e}\n")
        else:
            f.write("\n".join(str(i + 1) for i in solid))
            f.write("\n")
        f.write(f"Passive elements: {passive.size}\n")
        if passive.size:
            f.write("\n".join(str(i + 1) for i in passive))
            f.write("\n")
        f.write(f"Fixations: {len(fix_nodes)}\n")
        for n, s in zip(fix_nodes, fix_state):
            f.write(f"{int(n) + 1} {int(s[0])} {int(s[1])} {int(s[2])}\n")
        f.write(f"Loads: {len(load_nodes)}\n")
        for n, v in zip(load_nodes, load_vec):
            f.write(f"{int(n) + 1} {v[0]:.16e} {v[1]:.16e} {v[2]:.16e}\n")
        f.write(f"Additional Loads: {len(extra_loads)}\n")
        for k, (nodes, vec) in enumerate(extra_loads):
            vec = np.asarray(vec, dtype=np.float64).reshape(-1, 3)
            f.write(f"Loads: {len(nodes)} {k + 2}\n")
            for n, v in zip(nodes, vec):
                f.write(f"{int(n) + 1} {v[0]:.16e} {v[1]:.16e} {v[2]:.16e}\n")
