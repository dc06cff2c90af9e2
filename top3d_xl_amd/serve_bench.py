"""Serving workload: Poisson arrivals of topology optimisations on one GPU.

Request mix (seeded): cuboids (s, 2s, s) with s uniform over
{32, 48, 64, 96, 128, 192}; 70% GLOBAL (V0 ~ U[0.08, 0.3], nLoop 30) and 30%
LOCAL (local volume bound ~ U[0.4, 0.6], rHat ~ U{4..8}, nLoop 60); rMin =
sqrt(3); on the built-in cantilever faces a load of uniformly random
direction and magnitude ~ U[0.5, 2].

For every arrival rate the same stream runs through the serving engine
(continuous admission, `max_active` concurrent requests) and through the
one-at-a-time FP64 baseline (max_active = 1: a FIFO queue).  Reported per
run: completed requests per second, PCG solves per second (one solve per
design iteration), p50/p99 latency (arrival -> result) and p50/p99 queueing
delay.

    python -m top3d_xl_amd.serve_bench --rates 0.5 1 2 --requests 24
"""
from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np

SIZES = (32, 48, 64, 96, 128, 192)


def make_stream(n, rate, seed=0, sizes=SIZES, nloop_scale=1.0):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    reqs = []
    for i in range(n):
        s = int(rng.choice(sizes))
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        mag = float(rng.uniform(0.5, 2.0))
        if rng.random() < 0.7:
            r = dict(mode="GLOBAL", V0=float(rng.uniform(0.08, 0.3)), nLoop=max(1, int(round(30 * nloop_scale))),
                     rHat=0.0)
        else:
            r = dict(mode="LOCAL", V0=float(rng.uniform(0.4, 0.6)), nLoop=max(1, int(round(60 * nloop_scale))),
                     rHat=float(rng.integers(4, 9)))
        r.update(tag=i, shape=(s, 2 * s, s), arrival=float(t[i]), load_dir=tuple(d), load_mag=mag)
        reqs.append(r)
    return reqs


def run_stream(reqs, max_active, precision=0, mem_budget=0, timeout=3600.0, **params):
    """Submits reqs at their arrival times; params: further t3x_params fields
    (e.g. vcycle, omega) for every request."""
    import top3d_xl_amd as t3x
    S = t3x.Server(max_active=max_active, mem_budget=mem_budget, precision=precision, **params)
    done = {}
    t0 = time.perf_counter()
    i = 0
    while len(done) < len(reqs):
        now = time.perf_counter() - t0
        while i < len(reqs) and reqs[i]["arrival"] <= now:
            r = reqs[i]
            S.submit(r["tag"], *r["shape"], r["mode"], r["V0"], r["nLoop"], math.sqrt(3), r["rHat"],
                     r["load_dir"], r["load_mag"])
            i += 1
        wait = 50 if i >= len(reqs) else max(1, int(1000 * (reqs[i]["arrival"] - now)))
        for c in S.poll(wait_ms=min(wait, 50)):
            done[c["tag"]] = c
        if now > timeout:
            break
    wall = time.perf_counter() - t0
    S.close()
    return done, wall


def summarize(reqs, done, wall):
    lat = np.array([done[r["tag"]]["tEnd"] - done[r["tag"]]["tSubmit"] for r in reqs if r["tag"] in done])
    qd = np.array([done[r["tag"]]["tStart"] - done[r["tag"]]["tSubmit"] for r in reqs if r["tag"] in done])
    loops = sum(c["loops"] for c in done.values())
    bad = sum(1 for c in done.values() if c["status"] not in (0, 100))
    return dict(completed=len(done), failed=bad, wall_s=round(wall, 3),
                requests_per_s=round(len(done) / wall, 4), solves_per_s=round(loops / wall, 2),
                latency_p50_s=round(float(np.percentile(lat, 50)), 3),
                latency_p99_s=round(float(np.percentile(lat, 99)), 3),
                queue_p50_s=round(float(np.percentile(qd, 50)), 3), queue_p99_s=round(float(np.percentile(qd, 99)), 3))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rates", type=float, nargs="+", default=[1.0])
    ap.add_argument("--requests", type=int, default=24)
    ap.add_argument("--max-active", type=int, default=4)
    ap.add_argument("--precision", type=int, default=0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--sizes", type=int, nargs="+", default=list(SIZES))
    ap.add_argument("--nloop-scale", type=float, default=1.0)
    ap.add_argument("--no-baseline", action="store_true")
    args = ap.parse_args()
    for rate in args.rates:
        reqs = make_stream(args.requests, rate, args.seed, tuple(args.sizes), args.nloop_scale)
        rec = {"rate_per_s": rate, "requests": args.requests, "sizes": args.sizes, "nloop_scale": args.nloop_scale}
        done, wall = run_stream(reqs, args.max_active, args.precision)
        rec["engine"] = dict(max_active=args.max_active, precision=args.precision, **summarize(reqs, done, wall))
        if not args.no_baseline:
            done1, wall1 = run_stream(reqs, 1, 0)
            rec["one_at_a_time_fp64"] = summarize(reqs, done1, wall1)
            same = [np.array_equal(done[t]["xPhys"], done1[t]["xPhys"]) for t in done if t in done1]
            if args.precision == 0:
                rec["results_identical_to_baseline"] = bool(all(same))
        print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
