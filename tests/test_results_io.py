"""CPU tests of the result/trajectory writers (TOP3D_XL.m:187-204, 410-421, 564-570)."""
import numpy as np
import pytest

from oracle import Oracle
from top3d_xl_amd.irregular import generate
from top3d_xl_amd.results import cell_map_from_voxels, padded_volume, read_nii, write_nii, write_target_files


def test_cell_map_matches_oracle_eleMapBack():
    g = generate(4, (20, 40, 20))  # irregular and padded to the level-count multiple
    o = Oracle(g["voxels"], (g["fix_nodes"], g["fix_state"]), (g["load_nodes"], g["load_vec"]))
    lv = o.levels[0]
    res = (lv["resY"], lv["resX"], lv["resZ"])
    assert res != g["voxels"].shape  # exercises the padding
    assert np.array_equal(cell_map_from_voxels(g["voxels"], res), o.get("eleMapBack"))


def test_target_files_and_nii(tmp_path):
    r = dict(lssIts=np.array([27, 40]), cHist=np.array([53.539528792347, 35.0169977536]),
             sharpHist=np.array([0.4224, 0.41]), tHist=np.array([[0, 0, 0, 0, 0, 0.5], [0, 0, 0, 0, 0, 0.25]]))
    write_target_files(str(tmp_path), r)
    assert (tmp_path / "iters_Target.dat").read_text() == "27\n40\n"
    c = (tmp_path / "c_Target.dat").read_text().splitlines()
    assert c[0] == f"{53.539528792347:30.16e}" and len(c[0]) == 30
    assert float(c[1]) == 35.0169977536  # %.16e keeps every double digit
    assert (tmp_path / "timing_Target.dat").read_text() == f"{0.5:16.6e}\n{0.25:16.6e}\n"
    vox = np.zeros((3, 4, 5), bool)
    vox[1, 2, 3] = vox[0, 0, 0] = True
    cm = cell_map_from_voxels(vox, vox.shape)
    V = padded_volume(np.array([0.25, 0.75]), cm, vox.shape)
    assert V[0, 0, 0] == 0.25 and V[1, 2, 3] == 0.75 and V.sum() == 1.0
    write_nii(str(tmp_path / "d.nii"), V)
    assert (tmp_path / "d.nii").stat().st_size == 352 + 8 * V.size
    assert np.array_equal(read_nii(str(tmp_path / "d.nii")), V)


def test_design_outputs_written(tmp_path):
    from top3d_xl_amd.isosurface import enclosed_volume, read_stl
    from top3d_xl_amd.results import write_design_outputs
    vox = np.ones((4, 6, 4), bool)
    res = (4, 8, 4)  # padded in x
    r = dict(xPhys=np.ones(vox.sum()), cHist=[1.0, 0.5], sharpHist=[0.2, 0.1], lssIts=[7, 8],
             tHist=np.ones((2, 6)))
    write_design_outputs(str(tmp_path), r, vox, res)
    for f in ("iters_Target.dat", "c_Target.dat", "sharp_Target.dat", "timing_Target.dat", "DesignVolume.nii",
              "DesignVolume.stl"):
        assert (tmp_path / f).exists()
    V = read_nii(str(tmp_path / "DesignVolume.nii"))
    assert V.shape == res and V[:, :6].min() == 1.0 and V[:, 6:].max() == 0.0
    tris, _ = read_stl(str(tmp_path / "DesignVolume.stl"))
    # all-solid 4x6x4 cells: samples 1 on 4x6x4 grid points, 0 beyond x = 6:
    # the box [1,6.5] x [1,4] x [1,4] (the level crosses halfway to the void column)
    assert enclosed_volume(tris) == pytest.approx(5.5 * 3 * 3, rel=1e-6)
