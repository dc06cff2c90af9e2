"""Several load cases (loadingCond_{1..n}, TOP3D_XL.m:481-493; F_ columns,
1315-1321; one warm-started PCG per column, 101/121/301/321; element
compliance ceList * objWeightingList_, 597, with weights 1/n, 493).

CPU: the oracle's multi-case restatement is pinned by exact invariants (two
identical cases reproduce the one-case run bit for bit; the solid
compliance of two cases is the mean of the single-case values; a node
loaded in any case is released from the fixings, 1301-1305) and the
.TopVoxel round trip of additional load cases.  GPU: the library against
the oracle through the C ABI, at the bars of the single-case parity tests.
"""
import math

import numpy as np
import pytest

from oracle import Oracle
from top3d_xl_amd import topvoxel as tv

SMALL_TOL = 1e-4  # as in test_gpu_parity: CG at tol 1e-3 on these tiny models


def model(nely=6, nelx=12, nelz=6):
    """Cuboid with explicit BCs: the x=0 face fixed; case A pulls down on the
    lower x=nelx edge region, case B pushes sideways on the upper one."""
    vox = np.ones((nely, nelx, nelz), bool)
    uy, ux = nely + 1, nelx + 1
    fix = [jy + uy * (0 + ux * jz) for jz in range(nelz + 1) for jy in range(nely + 1)]
    fs = np.ones((len(fix), 3), np.uint8)
    a = [jy + uy * (nelx + ux * jz) for jz in range(2) for jy in range(nely + 1)]
    b = [jy + uy * (nelx + ux * jz) for jz in range(nelz - 1, nelz + 1) for jy in range(nely + 1)]
    la = (np.array(a), np.tile([0.0, 0.0, -1.0], (len(a), 1)) / len(a))
    lb = (np.array(b), np.tile([0.0, 0.7, 0.2], (len(b), 1)) / len(b))
    return vox, (np.array(fix), fs), la, lb


@pytest.mark.usefixtures("oracle_lib")
def test_identical_cases_reproduce_one_case():
    vox, fix, la, _ = model()
    r1 = Oracle(vox, fix, la).run("GLOBAL", 0.12, 5, math.sqrt(3), nSnap=5)
    o2 = Oracle(vox, fix, la, extra_loads=[la])
    assert o2.num_cases == 2
    np.testing.assert_array_equal(o2.get("objWeight"), [0.5, 0.5])
    r2 = o2.run("GLOBAL", 0.12, 5, math.sqrt(3), nSnap=5)
    assert r2["cSolid"] == r1["cSolid"]
    np.testing.assert_array_equal(r2["cHist"], r1["cHist"])
    np.testing.assert_array_equal(r2["xSnap"], r1["xSnap"])
    np.testing.assert_array_equal(r2["lssIts"], np.repeat(r1["lssIts"], 2))


@pytest.mark.usefixtures("oracle_lib")
def test_two_cases_average_the_compliance():
    vox, fix, la, lb = model()
    ca = Oracle(vox, fix, la).run("GLOBAL", 0.12, 1, math.sqrt(3))
    cb = Oracle(vox, fix, lb).run("GLOBAL", 0.12, 1, math.sqrt(3))
    o = Oracle(vox, fix, la, extra_loads=[lb])
    r = o.run("GLOBAL", 0.12, 1, math.sqrt(3))
    # same solves (same operator, right-hand side and zero start), weighted sum
    assert abs(r["cSolid"] / (0.5 * (ca["cSolid"] + cb["cSolid"])) - 1) < 1e-13
    np.testing.assert_array_equal(r["lssIts"], [ca["lssIts"][0], cb["lssIts"][0]])
    # F_ columns are the per-case load vectors on the compact nodes
    Fa, Fb = o.get("F", 0), o.get("F", 1)
    np.testing.assert_array_equal(Fa, Oracle(vox, fix, la).get("F"))
    np.testing.assert_array_equal(Fb, Oracle(vox, fix, lb).get("F"))


@pytest.mark.usefixtures("oracle_lib")
def test_node_loaded_in_any_case_is_released():
    vox, fix, la, _ = model()
    node = int(fix[0][3])  # a fixed node on the x=0 face
    lc = (np.array([node]), np.array([[0.0, 0.0, 1.0]]))
    n1 = len(Oracle(vox, fix, la).get("fixNode"))
    o = Oracle(vox, fix, la, extra_loads=[lc])
    assert len(o.get("fixNode")) == n1 - 1
    assert np.abs(o.get("F", 1)).sum() == 1.0
    assert np.abs(o.get("F", 0)).sum() == pytest.approx(np.abs(Oracle(vox, fix, la).get("F")).sum())


def test_topvoxel_additional_loads_round_trip(tmp_path):
    vox, fix, la, lb = model()
    p = tmp_path / "two.TopVoxel"
    tv.write_topvoxel(p, vox, fix[0], fix[1], la[0], la[1], extra_loads=[lb])
    m = tv.read_topvoxel(p)
    cases = tv.load_cases(m)
    assert len(cases) == 2
    for (n, v), (n0, v0) in zip(cases, [la, lb]):
        np.testing.assert_array_equal(n, n0)
        np.testing.assert_allclose(v, v0, rtol=1e-15)
    m["extra_loads"] = [(3, cases[1][0], cases[1][1])]
    with pytest.raises(tv.FormatError):
        tv.load_cases(m)


@pytest.mark.gpu
@pytest.mark.usefixtures("oracle_lib")
@pytest.mark.parametrize("mode", ["GLOBAL", "LOCAL"])
def test_gpu_two_load_cases(mode):
    import top3d_xl_amd as t3x
    vox, fix, la, lb = model()
    print(f"\n[gpu-multiload] 6x12x6 two cases {mode}", flush=True)
    args = (0.12, 8, math.sqrt(3)) if mode == "GLOBAL" else (0.5, 6, math.sqrt(3), 2.0)
    g = t3x.Problem(vox, fix, la, extra_loads=[lb])
    assert g.num_cases == 2
    rg = g.run(mode, *args, nSnap=args[1])
    g.close()
    ro = Oracle(vox, fix, la, extra_loads=[lb]).run(mode, *args, nSnap=args[1])
    assert rg["loops"] == ro["loops"]
    assert len(rg["lssIts"]) == 2 * rg["loops"]
    assert np.abs(rg["lssIts"] - ro["lssIts"]).max() <= 2
    assert abs(rg["cSolid"] / ro["cSolid"] - 1) < 1e-7
    np.testing.assert_allclose(rg["cHist"], ro["cHist"], rtol=SMALL_TOL)
    assert np.abs(rg["xSnap"] - ro["xSnap"]).max() < SMALL_TOL


@pytest.mark.gpu
@pytest.mark.usefixtures("oracle_lib")
def test_gpu_identical_cases_reproduce_one_case():
    import top3d_xl_amd as t3x
    vox, fix, la, _ = model()
    print("\n[gpu-multiload] identical cases", flush=True)
    g1 = t3x.Problem(vox, fix, la)
    r1 = g1.run("GLOBAL", 0.12, 5, math.sqrt(3), nSnap=5)
    g1.close()
    g2 = t3x.Problem(vox, fix, la, extra_loads=[la])
    r2 = g2.run("GLOBAL", 0.12, 5, math.sqrt(3), nSnap=5)
    g2.close()
    # two identical solves weighted 1/2 each: the compliance is exact
    assert r2["cSolid"] == r1["cSolid"]
    np.testing.assert_array_equal(r2["cHist"], r1["cHist"])
    np.testing.assert_array_equal(r2["xSnap"], r1["xSnap"])
    np.testing.assert_array_equal(r2["lssIts"], np.repeat(r1["lssIts"], 2))
