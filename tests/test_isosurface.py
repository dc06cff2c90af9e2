"""Design export as a closed iso-surface (TOP3D_XL.m:206-222, 572-577).

Pinned by what the level set of the piecewise-linear interpolant fixes:
a full block closes into the box of its grid points (volume (n-1)^3, area
6 (n-1)^2: the caps of isocaps); a linear field cuts the box by a plane
(exact volume); a single sample of 1 among 0s encloses half the star of its
node in the 6-tetrahedra split (volume 1/2); every surface is watertight
(each welded edge on exactly two triangles) and outward oriented.
"""
import numpy as np
import pytest

from top3d_xl_amd import isosurface as iso
from top3d_xl_amd.results import cell_map_from_voxels


def edges_closed(tris):
    key = {}
    for t in np.round(tris, 12):
        pts = [tuple(p) for p in t]
        for a, b in ((0, 1), (1, 2), (2, 0)):
            e = (pts[a], pts[b]) if pts[a] < pts[b] else (pts[b], pts[a])
            key[e] = key.get(e, 0) + 1
    return all(v == 2 for v in key.values())


@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 5, 4)])
def test_full_block_is_the_box_of_its_grid_points(shape):
    t = iso.isosurface_with_caps(np.ones(shape))
    a, b, c = (s - 1 for s in shape)
    assert iso.enclosed_volume(t) == pytest.approx(a * b * c, rel=1e-12)
    assert iso.surface_area(t) == pytest.approx(2 * (a * b + b * c + a * c), rel=1e-12)
    assert edges_closed(t)


def test_linear_field_is_cut_by_a_plane():
    m, n, p = 4, 9, 5
    V = np.broadcast_to((np.arange(n) / (n - 1))[None, :, None], (m, n, p)).copy()  # 0..1 along x
    t = iso.isosurface_with_caps(V, 0.5)
    # region x >= 1 + (n-1)/2 of the box [1, n] x [1, m] x [1, p]
    assert iso.enclosed_volume(t) == pytest.approx((m - 1) * (p - 1) * (n - 1) / 2, rel=1e-12)
    assert edges_closed(t)


def test_single_sample_encloses_half_its_star():
    V = np.zeros((3, 3, 3))
    V[1, 1, 1] = 1.0
    t = iso.isosurface_with_caps(V, 0.5)
    assert iso.enclosed_volume(t) == pytest.approx(0.5, rel=1e-12)
    assert edges_closed(t)


def test_random_design_is_watertight_and_outward(tmp_path):
    rng = np.random.default_rng(3)
    V = rng.random((7, 10, 6))
    t = iso.isosurface_with_caps(V, 0.5)
    assert edges_closed(t)
    v5 = iso.enclosed_volume(t)
    assert 0 < v5 < 6 * 9 * 5
    assert iso.enclosed_volume(iso.isosurface_with_caps(V, 0.7)) < v5  # nested level sets
    p = tmp_path / "d.stl"
    iso.write_stl(p, t)
    tr, nr = iso.read_stl(p)
    assert tr.shape == t.shape
    np.testing.assert_allclose(tr, t, rtol=0, atol=1e-5)
    assert np.all(np.abs(np.linalg.norm(nr, axis=1) - 1) < 1e-6)


def test_design_stl_flips_y(tmp_path):
    # a slab of solid cells at the first rows (y = 1 side) ends up at the
    # largest y after the flip of 209 (rows 5, 6; the level crosses at 4.5)
    vox = np.zeros((6, 8, 5), bool)
    vox[:2] = True
    res = (6, 8, 5)
    cm = cell_map_from_voxels(vox, res)
    t = iso.write_design_stl(tmp_path / "DesignVolume.stl", np.ones(len(cm)), cm, res)
    assert t[:, :, 1].min() == pytest.approx(4.5) and t[:, :, 1].max() == pytest.approx(6.0)
