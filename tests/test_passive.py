"""Passive elements (passiveElements_, TOP3D_XL.m:81, 150, 165, 274-276, 366, 381).

CPU: the oracle's OC with passive elements against a numpy restatement of
145-155 (xnew(passive) = 1 inside the bisection, so gt counts them), and the
structural rules of the design loop (passive xPhys is exactly 1 after every
update; the first solve, on the uniform initial design, is unaffected).
GPU: the HIP path against the oracle with a passive slab, both modes.
"""
import math

import numpy as np
import pytest

from oracle import Oracle


def passive_slab(shape):
    ny, nx, nz = shape
    ids = [iy + ny * (ix + nx * iz) for iz in range(nz) for ix in range(nx - 2, nx) for iy in range(ny)]
    return np.array(ids, dtype=np.int32)


def compact_index(shape, ids):
    # compact cell order of a full cuboid without padding = column-major voxel order
    return np.asarray(ids)


def oc_numpy(x, xPhys, dc, V0, passive):
    fval = xPhys.mean() - V0
    l1, l2, move = 0.0, 1e9, 0.2
    while (l2 - l1) / (l1 + l2) > 1e-6:
        lmid = 0.5 * (l2 + l1)
        xnew = np.maximum(0, np.maximum(x - move, np.minimum(1, np.minimum(x + move, x * np.sqrt(-dc / 1.0 / lmid)))))
        xnew[passive] = 1.0
        gt = fval + np.sum(xnew - x)
        if gt > 0:
            l1 = lmid
        else:
            l2 = lmid
    return xnew, np.abs(xnew - x).max()


@pytest.mark.usefixtures("oracle_lib")
def test_oracle_oc_passive_matches_numpy():
    shape = (8, 16, 8)
    o = Oracle(np.ones(shape, bool))
    pas = passive_slab(shape)
    o.set_passive(pas)
    rng = np.random.default_rng(1)
    n = o.nE
    x = np.clip(0.3 + 0.1 * rng.standard_normal(n), 0.01, 1)
    dc = -rng.random(n) - 0.01
    xo, cho, _ = o.oc(x, x, dc, 0.3)
    xn, chn = oc_numpy(x, x, dc, 0.3, compact_index(shape, pas))
    assert np.allclose(xo, xn, rtol=0, atol=1e-12) and abs(cho - chn) < 1e-12
    assert np.all(xo[pas] == 1.0)


@pytest.mark.usefixtures("oracle_lib")
@pytest.mark.parametrize("mode", ["GLOBAL", "LOCAL"])
def test_oracle_passive_design_rules(mode):
    shape = (8, 16, 8)
    pas = passive_slab(shape)
    args = ("GLOBAL", 0.3, 3, math.sqrt(3)) if mode == "GLOBAL" else ("LOCAL", 0.5, 3, math.sqrt(3), 2.0)
    o0 = Oracle(np.ones(shape, bool))
    r0 = o0.run(*args, nSnap=3)
    o = Oracle(np.ones(shape, bool))
    o.set_passive(pas)
    r = o.run(*args, nSnap=3)
    assert r["cHist"][0] == r0["cHist"][0]  # the first solve sees the uniform initial design
    assert np.all(r["xSnap"][:, pas] == 1.0)
    assert not np.allclose(r["xSnap"][-1], r0["xSnap"][-1])


@pytest.mark.gpu
@pytest.mark.usefixtures("oracle_lib")
@pytest.mark.parametrize("mode", ["GLOBAL", "LOCAL"])
def test_gpu_passive_matches_oracle(mode):
    import top3d_xl_amd as t3x
    shape = (10, 20, 10)
    pas = passive_slab(shape)
    args = ("GLOBAL", 0.3, 5, math.sqrt(3)) if mode == "GLOBAL" else ("LOCAL", 0.5, 4, math.sqrt(3), 2.0)
    g = t3x.Problem(np.ones(shape, bool))
    g.set_passive(pas)
    rg = g.run(*args, nSnap=5)
    g.close()
    o = Oracle(np.ones(shape, bool))
    o.set_passive(pas)
    ro = o.run(*args, nSnap=5)
    assert rg["loops"] == ro["loops"]
    assert np.all(rg["xSnap"][:, pas] == 1.0)
    # small-model trajectories carry an FP64 noise floor up to ~1e-5 (the
    # tiny-model bar of test_gpu_parity.py, 1e-4); a wrong passive rule moves
    # densities by O(0.1)
    np.testing.assert_allclose(rg["cHist"], ro["cHist"], rtol=1e-4)
    assert np.abs(rg["xSnap"] - ro["xSnap"]).max() < 1e-4


# ------------------------------------------- passive layers (1511-1615)
def layers_numpy(vox, fixed_nodes, loaded_nodes, nB, nL, nF):
    """Brute force of TopOpti_SetPassiveElements on the unpadded voxel grid:
    boundary nodes have fewer than 8 solid cells around them; layer 1 is the
    solid cells touching such a node (or, near loads / fixations, those of
    them touching a loaded / fixed node); each further layer is a 3x3x3
    binary dilation restricted to the solid cells."""
    from scipy import ndimage
    vox = np.asarray(vox, bool)
    ny, nx, nz = vox.shape
    cnt = np.zeros((ny + 1, nx + 1, nz + 1), int)
    for dy in (0, 1):
        for dx in (0, 1):
            for dz in (0, 1):
                cnt[dy:dy + ny, dx:dx + nx, dz:dz + nz] += vox
    bnode = (cnt > 0) & (cnt < 8)

    def touching(nodes):
        t = np.zeros_like(vox)
        for dy in (0, 1):
            for dx in (0, 1):
                for dz in (0, 1):
                    t |= nodes[dy:dy + ny, dx:dx + nx, dz:dz + nz]
        return t & vox

    bcell = touching(bnode)

    def grow(s, n):
        for _ in range(n - 1):
            s = ndimage.binary_dilation(s, np.ones((3, 3, 3), bool)) & vox
        return s

    out = np.zeros_like(vox)
    if nB > 0:
        out |= grow(bcell, nB)
    for n, ids in ((nL, loaded_nodes), (nF, fixed_nodes)):
        if n > 0:
            m = np.zeros((ny + 1) * (nx + 1) * (nz + 1), bool)
            m[np.asarray(ids)] = True
            out |= grow(bcell & touching(m.reshape((ny + 1, nx + 1, nz + 1), order="F")), n)
    return np.flatnonzero(out.reshape(-1, order="F"))


def layer_model():
    vox = np.ones((8, 14, 8), bool)
    vox[3:5, 5:9, 3:5] = False  # an interior cavity: its faces are boundary too
    ny, nx, nz = vox.shape
    uy, ux = ny + 1, nx + 1
    fix = np.array([jy + uy * (0 + ux * jz) for jz in range(nz + 1) for jy in range(ny + 1)])
    load = np.array([jy + uy * (nx + ux * jz) for jz in range(2) for jy in range(ny + 1)])
    return vox, (fix, np.ones((len(fix), 3), np.uint8)), (load, np.tile([0, 0, -1.0], (len(load), 1)) / len(load))


LAYERS = [(1, 0, 0), (2, 0, 0), (0, 1, 0), (0, 2, 3), (1, 3, 2)]


@pytest.mark.usefixtures("oracle_lib")
@pytest.mark.parametrize("layers", LAYERS)
def test_oracle_passive_layers_match_dilation(layers):
    vox, fix, load = layer_model()
    o = Oracle(vox, fix, load)
    o.set_passive_layers(*layers)
    got = np.flatnonzero(o.get("isPassive"))  # compact order == voxel order (no padding here)
    want = layers_numpy(vox, fix[0], load[0], *layers)
    ids = np.flatnonzero(vox.reshape(-1, order="F"))  # compact -> voxel id
    np.testing.assert_array_equal(ids[got], want)
    assert len(want) > 0


@pytest.mark.gpu
@pytest.mark.usefixtures("oracle_lib")
@pytest.mark.parametrize("layers", [(1, 0, 0), (0, 2, 3)])
def test_gpu_passive_layers_match_oracle(layers):
    import top3d_xl_amd as t3x
    vox, fix, load = layer_model()
    print(f"\n[gpu-passive] layers {layers}", flush=True)
    g = t3x.Problem(vox, fix, load)
    g.set_passive_layers(*layers)
    np.testing.assert_array_equal(g.passive(), layers_numpy(vox, fix[0], load[0], *layers))
    rg = g.run("GLOBAL", 0.3, 4, math.sqrt(3), nSnap=4)
    g.close()
    o = Oracle(vox, fix, load)
    o.set_passive_layers(*layers)
    ro = o.run("GLOBAL", 0.3, 4, math.sqrt(3), nSnap=4)
    assert rg["loops"] == ro["loops"]
    np.testing.assert_allclose(rg["cHist"], ro["cHist"], rtol=1e-4)
    assert np.abs(rg["xSnap"] - ro["xSnap"]).max() < 1e-4
