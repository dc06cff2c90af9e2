"""Parity of the FP64 'Standard' V-cycle (typeVcycle_ = 'Standard',
TOP3D_XL.m:34, 759-777, with coarse-level K*u through the stored Galerkin
matrices, 727-744) with the FP64 CPU oracle, through the C ABI.

Bars as for the 'Adapted' FP64 path: V-cycle 1e-11 relative; PCG the same
iteration count on a well-conditioned field and iterates within the CG
rounding sensitivity; whole optimisation on config 1 (true(50,100,50),
GLOBAL, V0 0.12): compliance within 1e-8 relative and densities within 1e-6
max-abs, widened to 5x the FP64 noise floor measured live from three oracle
coarsest-level orderings.
"""
import math

import numpy as np
import pytest

from oracle import Oracle
from tests.test_gpu_parity import CASES, SMALL_TOL, _trajectory_parity, rel

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("oracle_lib")]


def make(name):
    import top3d_xl_amd as t3x
    vox, fix, load = CASES[name]()
    if fix is None:
        g, o = t3x.Problem(vox, vcycle=1), Oracle(vox)
    else:
        g, o = t3x.Problem(vox, fix, load, vcycle=1), Oracle(vox, fix, load)
    o.set_vcycle("Standard")
    return g, o


@pytest.fixture(scope="module", params=list(CASES))
def pair(request):
    print(f"\n[gpu-standard] case {request.param}", flush=True)
    g, o = make(request.param)
    rng = np.random.default_rng(5)
    E = 1e-6 + rng.random(o.nE) ** 3
    for m in (g, o):
        m.set_modulus(E)
        m.assemble()
    yield dict(g=g, o=o, E=E, rng=rng)
    g.close()


def test_standard_vcycle(pair):
    g, o, rng = pair["g"], pair["o"], pair["rng"]
    for _ in range(2):
        r = rng.standard_normal(3 * o.nN)
        assert rel(g.vcycle(r), o.vcycle(r)) < 1e-11


def test_standard_pcg(pair):
    g, o, rng = pair["g"], pair["o"], pair["rng"]
    E = 0.5 + rng.random(o.nE)
    for m in (g, o):
        m.set_modulus(E)
        m.assemble()
    b = o.get("F")
    xg, itg, _ = g.pcg(b)
    xo, ito, _ = o.pcg(b)
    assert itg == ito
    assert rel(xg, xo) < 1e-4
    for m in (g, o):
        m.set_modulus(pair["E"])
        m.assemble()


@pytest.mark.parametrize("name", ["cuboid_6x12x6", "irregular_5x9x5"])
def test_standard_run_global_small(oracle_lib, name):
    g, o = make(name)
    rg = g.run("GLOBAL", 0.12, 10, math.sqrt(3), nSnap=10)
    ro = o.run("GLOBAL", 0.12, 10, math.sqrt(3), nSnap=10)
    g.close()
    assert rg["loops"] == ro["loops"]
    assert np.abs(rg["lssIts"] - ro["lssIts"]).max() <= 3
    assert abs(rg["cSolid"] / ro["cSolid"] - 1) < 1e-7
    np.testing.assert_allclose(rg["cHist"], ro["cHist"], rtol=SMALL_TOL)
    assert np.abs(rg["xSnap"] - ro["xSnap"]).max() < SMALL_TOL


def test_config1_standard_global_parity(oracle_lib):
    """Config 1 with the 'Standard' cycle: first 10 iterations; the strict bar
    must hold over all 10 (the measured oracle spread is
    below 3e-12 in compliance and 3e-10 in densities)."""
    _trajectory_parity("GLOBAL", 0.12, 10, 0.0, 10, vcycle="Standard")
