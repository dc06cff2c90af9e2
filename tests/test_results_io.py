"""CPU tests of the result/trajectory writers (TOP3D_XL.m:187-204, 410-421, 564-570)."""
import numpy as np

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
