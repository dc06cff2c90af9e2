"""CPU tests of the .TopVoxel reader/writer and the irregular-domain generator.

The reference ships no .TopVoxel file (its datasets are an external
download, README.md:33), so the reader is pinned by a hand-written file that
follows the token layout of CreateVoxelFEAmodel (TOP3D_XL.m:448-528) line by
line, including density values, passive elements, duplicate BC rows and an
additional load case, and by write -> read round trips.
"""
import numpy as np
import pytest

from top3d_xl_amd.irregular import generate
from top3d_xl_amd.topvoxel import FormatError, read_topvoxel, write_topvoxel

HAND_WRITTEN = """# Topology Voxel Model
Version: 1.0
Resolution: 3 2 4
Density values: 1
Solid voxels: 5
1 1.0
2 0.5
7 0.25
8 1.0
24 0.125
Passive elements: 2
8
2
Fixations: 3
4 1 1 1
1 1 0 1
4 0 0 0
Loads: 2
60 0.0 0.0 -0.5
59 1e-1 0 -0.5
Additional Loads: 1
Loads: 1 2
60 1.0 0.0 0.0
"""


def test_read_hand_written(tmp_path):
    p = tmp_path / "m.TopVoxel"
    p.write_text(HAND_WRITTEN)
    m = read_topvoxel(str(p))
    vox = m["voxels"]
    assert vox.shape == (2, 3, 4)  # nely, nelx, nelz from "Resolution: nelx nely nelz"
    flat = vox.reshape(-1, order="F")
    assert np.flatnonzero(flat).tolist() == [0, 1, 6, 7, 23]
    assert m["density"].tolist() == [1.0, 0.5, 0.25, 1.0, 0.125]
    assert m["passive"].tolist() == [1, 7]  # sorted, 0-based
    # duplicates keep their first row (unique(), 499-518); rows sorted by node id
    assert m["fix_nodes"].tolist() == [0, 3]
    assert m["fix_state"].tolist() == [[1, 0, 1], [1, 1, 1]]
    assert m["load_nodes"].tolist() == [58, 59]
    assert np.allclose(m["load_vec"], [[0.1, 0.0, -0.5], [0.0, 0.0, -0.5]])
    assert len(m["extra_loads"]) == 1 and m["extra_loads"][0][0] == 2


def test_bad_file(tmp_path):
    p = tmp_path / "bad.TopVoxel"
    p.write_text(HAND_WRITTEN.replace("Resolution:", "Resolutions:"))
    with pytest.raises(FormatError):
        read_topvoxel(str(p))


def test_round_trip(tmp_path):
    g = generate(3, (24, 48, 24))
    p = tmp_path / "g.TopVoxel"
    write_topvoxel(str(p), g["voxels"], g["fix_nodes"], g["fix_state"], g["load_nodes"], g["load_vec"])
    m = read_topvoxel(str(p))
    assert np.array_equal(m["voxels"], g["voxels"])
    assert np.array_equal(m["fix_nodes"], g["fix_nodes"])
    assert np.array_equal(m["fix_state"], g["fix_state"])
    assert np.array_equal(m["load_nodes"], g["load_nodes"])
    assert np.array_equal(m["load_vec"], g["load_vec"])  # %.16e round-trips doubles exactly


def test_generator_invariants():
    from scipy import ndimage
    shp = (48, 96, 48)
    g = generate(1, shp)
    g2 = generate(1, shp)
    assert np.array_equal(g["voxels"], g2["voxels"])  # deterministic in the seed
    vox = g["voxels"]
    _, ncomp = ndimage.label(vox)
    assert ncomp == 1  # one face-connected component
    assert 0.15 < g["occupancy"] < 0.45
    ny, nx, nz = shp
    # every BC node is a node of a solid voxel
    nodes = np.zeros((ny + 1, nx + 1, nz + 1), bool)
    for d in np.ndindex(2, 2, 2):
        nodes[d[0]:d[0] + ny, d[1]:d[1] + nx, d[2]:d[2] + nz] |= vox
    flatn = nodes.reshape(-1, order="F")
    assert flatn[g["fix_nodes"]].all() and flatn[g["load_nodes"]].all()
    assert not np.intersect1d(g["fix_nodes"], g["load_nodes"]).size
    # unit total load along the seeded direction
    assert np.allclose(g["load_vec"].sum(axis=0), g["direction"], atol=1e-12)
    assert abs(np.linalg.norm(g["direction"]) - 1) < 1e-12
    # fixations in the bottom slab, loads in the top slab of the domain's z extent
    jz_f = g["fix_nodes"] // ((ny + 1) * (nx + 1))
    jz_l = g["load_nodes"] // ((ny + 1) * (nx + 1))
    z0, z1 = g["z_extent"]
    assert jz_f.max() <= z0 + max(1, round(0.05 * (z1 - z0)))
    assert jz_l.min() >= z1 - max(1, round(0.05 * (z1 - z0)))


def test_oracle_accepts_generated_model():
    from oracle import Oracle
    g = generate(2, (24, 48, 24))
    o = Oracle(g["voxels"], (g["fix_nodes"], g["fix_state"]), (g["load_nodes"], g["load_vec"]))
    F = o.get("F").reshape(-1, 3)
    assert np.allclose(F.sum(axis=0), g["direction"], atol=1e-12)


def test_generator_supports_on_the_surface():
    """Every fixed node touches a void voxel or the box boundary: the
    reference applies fixations only at boundary nodes (1369)."""
    g = generate(4, (24, 48, 24))
    vox = g["voxels"]
    ny, nx, nz = vox.shape
    pad = np.pad(vox, 1, constant_values=False)
    ids = np.asarray(g["fix_nodes"])
    jy, jx, jz = ids % (ny + 1), (ids // (ny + 1)) % (nx + 1), ids // ((ny + 1) * (nx + 1))
    around = np.zeros(len(ids), int)
    for dy in (0, 1):
        for dx in (0, 1):
            for dz in (0, 1):
                around += pad[jy + dy, jx + dx, jz + dz]
    assert len(ids) > 0 and np.all((around > 0) & (around < 8))


@pytest.mark.usefixtures("oracle_lib")
def test_oracle_refuses_interior_fixation():
    from oracle import Oracle
    vox = np.ones((6, 12, 6), bool)
    uy, ux = 7, 13
    fix = np.array([3 + uy * (6 + ux * 3)])  # the centre node: 8 solid voxels around it
    load = np.array([0 + uy * (12 + ux * 0)])
    with pytest.raises(ValueError, match="inside the domain"):
        Oracle(vox, (fix, np.ones((1, 3), np.uint8)), (load, np.array([[0.0, 0.0, -1.0]])))
