"""Serving engine (include/t3x_serve.h) on the GPU.

* the admission footprint is an upper bound of what a run allocates;
* a request's result is bitwise the same alone (one request at a time) and in
  a mix (four running together, more waiting for memory), the bar being
  bit-equality because a request's kernels, launch shapes and reduction
  orders do not depend on what else runs;
* a served request equals a direct t3x_run with the reference's built-in
  cantilever BCs (TOP3D_XL.m:536-559) when its load is the built-in one.
"""
import math
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_footprint_bounds_allocation():
    import top3d_xl_amd as t3x
    from top3d_xl_amd.irregular import generate
    g = generate(5, (24, 48, 24))
    std = dict(precision=1, vcycle=1, omega=0.4)
    cases = [(np.ones((16, 32, 16), bool), None, "GLOBAL", {}), (np.ones((24, 48, 24), bool), None, "LOCAL", {}),
             (np.ones((24, 48, 24), bool), None, "GLOBAL", dict(precision=1)),
             (np.ones((24, 48, 24), bool), None, "GLOBAL", std),
             (np.ones((24, 48, 24), bool), None, "GLOBAL", dict(std, precision=2)), (g["voxels"], g, "GLOBAL", {})]
    for vox, bc, mode, prec in cases:
        if bc is None:
            P = t3x.Problem(vox, **prec)
        else:
            P = t3x.Problem(vox, (bc["fix_nodes"], bc["fix_state"]), (bc["load_nodes"], bc["load_vec"]), **prec)
        P.run(mode, 0.3 if mode == "GLOBAL" else 0.5, 2, math.sqrt(3), 4.0 if mode == "LOCAL" else 0.0)
        used = P.memory()
        P.close()
        est = t3x.footprint(*vox.shape, voxels=vox, mode=mode, **prec)
        print(f"  {vox.shape} {mode} prec {prec}: used {used / 2**20:.1f} MiB, estimate {est / 2**20:.1f} MiB")
        assert used <= est <= 1.5 * used + (8 << 20)


REQUESTS = [  # tag, (nely, nelx, nelz), mode, V0, nLoop, rHat, load direction, magnitude
    (0, (16, 32, 16), "GLOBAL", 0.12, 6, 0.0, (0.0, 0.0, -1.0), 1.0),
    (1, (24, 48, 24), "LOCAL", 0.5, 5, 4.0, (0.3, 0.0, -1.0), 2.0),
    (2, (12, 24, 12), "GLOBAL", 0.25, 8, 0.0, (0.0, 1.0, 0.0), 0.5),
    (3, (32, 64, 32), "GLOBAL", 0.1, 4, 0.0, (1.0, 0.0, -1.0), 1.0),
    (4, (16, 32, 16), "LOCAL", 0.45, 6, 5.0, (0.0, 0.0, -1.0), 1.5),
    (5, (24, 48, 24), "GLOBAL", 0.3, 5, 0.0, (-0.5, 0.5, -1.0), 1.0),
    (6, (12, 24, 12), "LOCAL", 0.6, 4, 4.0, (0.0, -1.0, 0.0), 1.0),
    (7, (32, 64, 32), "GLOBAL", 0.2, 3, 0.0, (0.0, 0.0, 1.0), 3.0),
]


def _serve(max_active, budget=0):
    import top3d_xl_amd as t3x
    S = t3x.Server(max_active=max_active, mem_budget=budget)
    out = {}
    peak = 0
    for tag, (ny, nx, nz), mode, V0, nLoop, rHat, d, mag in REQUESTS:
        S.submit(tag, ny, nx, nz, mode, V0, nLoop, math.sqrt(3), rHat, d, mag)
        if max_active == 1:  # alone: wait for this one before the next arrives
            while tag not in out:
                for c in S.poll(wait_ms=50):
                    out[c["tag"]] = c
    t0 = time.time()
    while len(out) < len(REQUESTS) and time.time() - t0 < 120:
        st = S.stats()
        peak = max(peak, st["mem_in_use"])
        for c in S.poll(wait_ms=20):
            out[c["tag"]] = c
    S.close()
    return out, peak


def test_alone_equals_mix():
    import top3d_xl_amd as t3x
    alone, _ = _serve(1)
    # a budget that holds the two largest footprints but not all requests at once
    fps = sorted(t3x.footprint(*r[1], mode=r[2]) for r in REQUESTS)
    budget = fps[-1] + fps[-2]
    mix, peak = _serve(4, budget)
    assert len(alone) == len(mix) == len(REQUESTS)
    assert 0 < peak <= budget
    for tag in alone:
        a, m = alone[tag], mix[tag]
        assert a["status"] == m["status"] == 0, (a["status"], m["status"])
        assert a["loops"] == m["loops"] and a["pcgIts"] == m["pcgIts"]
        assert np.array_equal(a["cHist"], m["cHist"]), tag
        assert np.array_equal(a["xPhys"], m["xPhys"]), tag
    lat_mix = max(c["tEnd"] for c in mix.values()) - min(c["tSubmit"] for c in mix.values())
    lat_alone = sum(c["tEnd"] - c["tStart"] for c in alone.values())
    print(f"  all {len(REQUESTS)} requests: mix makespan {lat_mix:.2f}s vs alone busy time {lat_alone:.2f}s; "
          f"peak admitted {peak / 2**20:.0f} MiB of budget {budget / 2**20:.0f} MiB")


PRECISION_FORM = dict(precision=2, vcycle=1, omega=0.4, coarsestControl=0)  # the shipped precision form


@pytest.mark.parametrize("s,nloop,params", [(16, 5, {}), (48, 30, PRECISION_FORM)], ids=["fp64-16", "mixed-48"])
def test_served_request_equals_direct_run(s, nloop, params):
    import top3d_xl_amd as t3x
    print(f"\n[gpu-serving] served vs direct s={s} nLoop={nloop} {params}", flush=True)
    S = t3x.Server(max_active=2, **params)
    S.submit(42, s, 2 * s, s, "GLOBAL", 0.12, nloop, math.sqrt(3), 0.0, (0.0, 0.0, -1.0), 1.0)
    got = []
    t0 = time.time()
    while not got and time.time() - t0 < 120:
        got = S.poll(wait_ms=100)
    S.close()
    P = t3x.Problem(np.ones((s, 2 * s, s), bool), **params)
    r = P.run("GLOBAL", 0.12, nloop, math.sqrt(3))
    P.close()
    assert got and got[0]["status"] == 0
    assert np.array_equal(got[0]["cHist"], r["cHist"])
    assert np.array_equal(got[0]["xPhys"], r["xPhys"])


def test_config5_sizes_alone_equals_mix():
    """The config-5 request mix at its real sizes and iteration counts in the
    precision form: one request of each s in {48, 64, 96, 128, 192} plus one
    at 32, modes and parameters from the seeded Poisson stream (GLOBAL nLoop
    30, LOCAL nLoop 60), submitted at the stream's arrival times (rate 4/s)
    under a memory budget that holds the two largest footprints but not all
    six.  Each result must be bitwise the same as the request run alone."""
    import top3d_xl_amd as t3x
    from top3d_xl_amd.serve_bench import make_stream, run_stream
    reqs = make_stream(6, 4.0, seed=9)
    for r, s in zip(reqs, (192, 48, 128, 64, 96, 32)):
        r["shape"] = (s, 2 * s, s)
    print("\n[gpu-serving] config-5 sizes:", [(r["shape"][0], r["mode"], r["nLoop"]) for r in reqs], flush=True)
    assert {r["mode"] for r in reqs} == {"GLOBAL", "LOCAL"}
    fps = sorted(t3x.footprint(*r["shape"], mode=r["mode"], **PRECISION_FORM) for r in reqs)
    budget = fps[-1] + fps[-2]
    mix, wall_mix = run_stream(reqs, 4, mem_budget=budget, timeout=300.0, **PRECISION_FORM)
    print(f"  mix: {len(mix)} done in {wall_mix:.1f}s", flush=True)
    alone = {}
    S = t3x.Server(max_active=1, **PRECISION_FORM)
    for r in reqs:
        S.submit(r["tag"], *r["shape"], r["mode"], r["V0"], r["nLoop"], math.sqrt(3), r["rHat"], r["load_dir"],
                 r["load_mag"])
        t0 = time.time()
        while r["tag"] not in alone and time.time() - t0 < 120:
            for c in S.poll(wait_ms=50):
                alone[c["tag"]] = c
    S.close()
    assert len(alone) == len(mix) == len(reqs)
    for r in reqs:
        a, m = alone[r["tag"]], mix[r["tag"]]
        print(f"  s={r['shape'][0]} {r['mode']}: loops {a['loops']} pcg {a['pcgIts']} c {a['cFinal']:.6e} "
              f"alone {a['tEnd'] - a['tStart']:.2f}s, mix latency {m['tEnd'] - m['tSubmit']:.2f}s", flush=True)
        assert a["status"] == m["status"] == 0, (a["status"], m["status"])
        assert a["loops"] == m["loops"] and a["pcgIts"] == m["pcgIts"]
        assert np.array_equal(a["cHist"], m["cHist"]), r["tag"]
        assert np.array_equal(a["xPhys"], m["xPhys"]), r["tag"]


def test_duplicate_tag_refused_until_released():
    """A tag names one request until its completion is released: a second
    submit with a live tag is refused, and accepted again after release."""
    import top3d_xl_amd as t3x
    S = t3x.Server(max_active=1)
    try:
        S.submit(7, 8, 16, 8, "GLOBAL", 0.3, 2)
        with pytest.raises(t3x.T3XError):
            S.submit(7, 8, 16, 8, "GLOBAL", 0.3, 2)
        got = []
        t0 = time.time()
        while not got and time.time() - t0 < 60:
            got = S.poll(wait_ms=100)  # poll() copies the results and releases the tag
        assert got and got[0]["tag"] == 7 and got[0]["status"] == 0
        S.submit(7, 8, 16, 8, "GLOBAL", 0.3, 2)
        got = []
        t0 = time.time()
        while not got and time.time() - t0 < 60:
            got = S.poll(wait_ms=100)
        assert got and got[0]["tag"] == 7
    finally:
        S.close()
