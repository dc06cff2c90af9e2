"""The PCG iteration replayed from a captured HIP graph (t3x_params.graphs = 1)
gives bitwise the same optimisation as kernel-by-kernel
launches: every replay runs between the solver's own eager work (flag
read-backs every checkEvery iterations, the filter, OC/MMA, assembly), and
whole trajectories are compared, so each replay is checked against the
eager result."""
import math

import numpy as np
import pytest

from tests.test_multiload import model as two_case_model

pytestmark = pytest.mark.gpu

FORMS = {"fp64": dict(precision=0), "fp64_standard": dict(precision=0, vcycle=1),
         "mixed": dict(precision=1, vcycle=1, omega=0.4), "mixed_adapted": dict(precision=1),
         "mixed_f16": dict(precision=2, vcycle=1, omega=0.4)}


@pytest.mark.parametrize("form", list(FORMS))
def test_graph_replay_matches_eager(form):
    import top3d_xl_amd as t3x
    vox = np.ones((10, 20, 10), bool)
    print(f"\n[gpu-graphs] 10x20x10 {form}", flush=True)
    runs = []
    for graphs in (0, 1):
        P = t3x.Problem(vox, graphs=graphs, checkEvery=3, **FORMS[form])
        runs.append(P.run("GLOBAL", 0.12, 6, math.sqrt(3), nSnap=6))
        P.close()
    a, b = runs
    assert a["loops"] == b["loops"]
    np.testing.assert_array_equal(a["lssIts"], b["lssIts"])
    np.testing.assert_array_equal(a["cHist"], b["cHist"])
    np.testing.assert_array_equal(a["xSnap"], b["xSnap"])
    assert a["cSolid"] == b["cSolid"]


def test_graph_per_solution_vector():
    """Two load cases solve into two vectors: one graph each, LOCAL mode."""
    import top3d_xl_amd as t3x
    vox, fix, la, lb = two_case_model()
    print("\n[gpu-graphs] two load cases LOCAL", flush=True)
    runs = []
    for graphs in (0, 1):
        P = t3x.Problem(vox, fix, la, extra_loads=[lb], graphs=graphs)
        runs.append(P.run("LOCAL", 0.5, 4, math.sqrt(3), 2.0, nSnap=4))
        P.close()
    a, b = runs
    np.testing.assert_array_equal(a["lssIts"], b["lssIts"])
    np.testing.assert_array_equal(a["cHist"], b["cHist"])
    np.testing.assert_array_equal(a["xSnap"], b["xSnap"])
