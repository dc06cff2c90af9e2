"""Whole-trajectory parity of the HIP path with the FP64 oracle on the
workloads beyond the built-in cuboid: a seeded irregular domain from the
config-4 generator (explicit fixations and surface load, V0 0.4) and the
config-1 cuboid with two load cases (F_ columns, weights 1/2).

Bars as for config 1 (test_gpu_parity): compliance within 1e-8 relative and
densities within 1e-6 max-abs per iteration, widened to 5x the FP64 noise
floor measured live from two oracle coarsest-level orderings where that
floor is larger (the same factor in exact arithmetic, as any fill-reducing
permutation of CHOLMOD, 1065).  The strict bars must hold over the first
n_strict iterations.

The irregular domain is run twice.  At the reference's tol 1e-3 (tol_, 30)
its thin features make the stopping iteration rounding-sensitive: the
oracle's two orderings already differ by 4.6e-8 in compliance and 8.9e-6 in
density at the first design iteration, and the GPU stops the first solve at
224 against the oracle's 223 PCG iterations; the bars there are 1e-6 and
1e-4 (20x and 10x that measured floor).  Solved to tol 1e-8 the stopping
point no longer matters and the strict bars hold at every iteration
(measured 3e-12 and 6e-11).
"""
import math

import numpy as np
import pytest

from oracle import Oracle

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("oracle_lib")]


def _parity(vox, fix, load, extra, mode, V0, nLoop, rHat, n_strict, label, tol=1e-3, max_it=800, c_bar=1e-8,
            x_bar=1e-6, layers=None):
    import top3d_xl_amd as t3x
    print(f"\n[gpu-trajectory] {label} {mode} {nLoop} iterations, tol {tol}", flush=True)
    g = t3x.Problem(vox, fix, load, extra_loads=extra, tol=tol, maxIT=max_it)
    if layers:
        g.set_passive_layers(*layers)
    rg = g.run(mode, V0, nLoop, math.sqrt(3), rHat, nSnap=nLoop)
    g.close()
    ros = []
    for v in (0, 1):
        o = Oracle(vox, fix, load, extra_loads=extra)
        o.set_tol(tol, max_it)
        o.set_order_variant(v)
        if layers:
            o.set_passive_layers(*layers)
        ros.append(o.run(mode, V0, nLoop, math.sqrt(3), rHat, nSnap=nLoop))
    ro = ros[0]
    k = min(rg["loops"], ro["loops"])
    assert rg["loops"] == ro["loops"]
    c_spread = np.abs(ros[1]["cHist"][:k] / ro["cHist"][:k] - 1)
    x_spread = np.abs(ros[1]["xSnap"][:k] - ro["xSnap"][:k]).max(axis=1)
    crel = np.abs(rg["cHist"][:k] / ro["cHist"][:k] - 1)
    xerr = np.abs(rg["xSnap"][:k] - ro["xSnap"][:k]).max(axis=1)
    print("gpu its", rg["lssIts"], "oracle its", ro["lssIts"], flush=True)
    print("compliance rel err", crel, "\n  oracle spread", c_spread, flush=True)
    print("xPhys max-abs err", xerr, "\n  oracle spread", x_spread, flush=True)
    assert np.all(crel <= np.maximum(c_bar, 5 * c_spread)), crel
    assert np.all(xerr <= np.maximum(x_bar, 5 * x_spread)), xerr
    assert np.all(crel[:n_strict] < c_bar) and np.all(xerr[:n_strict] < x_bar)


@pytest.mark.parametrize("tol,c_bar,x_bar", [(1e-3, 1e-6, 1e-4), (1e-8, 1e-8, 1e-6)])
def test_irregular_generated_domain_parity(tol, c_bar, x_bar):
    from top3d_xl_amd.irregular import generate
    gm = generate(5, (32, 64, 32))
    _parity(gm["voxels"], (gm["fix_nodes"], gm["fix_state"]), (gm["load_nodes"], gm["load_vec"]), (),
            "GLOBAL", 0.4, 6, 0.0, 6, "irregular seed 5, 32x64x32", tol=tol, max_it=5000, c_bar=c_bar, x_bar=x_bar)


def test_config1_two_load_cases_parity():
    nely, nelx, nelz = 50, 100, 50
    vox = np.ones((nely, nelx, nelz), bool)
    uy, ux = nely + 1, nelx + 1
    fix = np.array([jy + uy * (0 + ux * jz) for jz in range(nelz + 1) for jy in range(nely + 1)])
    ztop = int(math.floor(nelz / 6 + 1 + 0.5))  # the built-in load of 552-556
    la = np.array([jy + uy * (nelx + ux * jz) for jz in range(ztop) for jy in range(nely + 1)])
    lb = np.array([jy + uy * (nelx + ux * jz) for jz in range(nelz - ztop + 1, nelz + 1) for jy in range(nely + 1)])
    load_a = (la, np.tile([0.0, 0.0, -1.0], (len(la), 1)) / len(la))
    load_b = (lb, np.tile([0.0, 1.0, 0.0], (len(lb), 1)) / len(lb))
    _parity(vox, (fix, np.ones((len(fix), 3), np.uint8)), load_a, [load_b], "GLOBAL", 0.12, 6, 0.0, 4,
            "50x100x50 two load cases")


def test_irregular_local_passive_parity():
    """LOCAL mode (Heaviside, p-norm local volume, MMA) on the generated
    irregular domain with one boundary layer of passive elements (specify
    PassiveRegions_ = [1 0 0], 37), solved to tol 1e-8."""
    from top3d_xl_amd.irregular import generate
    gm = generate(7, (24, 48, 24))
    _parity(gm["voxels"], (gm["fix_nodes"], gm["fix_state"]), (gm["load_nodes"], gm["load_vec"]), (),
            "LOCAL", 0.5, 5, 4.0, 5, "irregular seed 7, 24x48x24, passive boundary layer", tol=1e-8,
            max_it=5000, layers=(1, 0, 0))
