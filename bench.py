#!/usr/bin/env python
"""Benchmark of the TOP3D_XL topology-optimisation inner loop on one MI355X.

A step is one GLOBAL design iteration (TOP3D_XL.m:111-183: SIMP, Galerkin
assembly + coarsest Cholesky, MG-PCG solve, compliance, PDE filter, OC) of the
256x512x256 cuboid with the reference's built-in boundary conditions, V0=0.12,
rMin=sqrt(3).  Warm-up steps and the fully-solid reference solve run before
the timed region; the model state is resident in HBM throughout.

By default the precision form and the FP64 plain path run side by side on
the same workload.  The precision form is the reference's PCG in FP64 (K*p,
vectors, dot products) preconditioned in FP32 by the 'Standard' V(1,1) cycle
(damped Jacobi, omega 0.4, K*u on every level) around the FP64 coarsest
Cholesky; the FP32 'Adapted' cycle and the 'Standard' cycle in FP64 are run as
well.  `value` is the precision
form's ms per design iteration; the line carries the FP64 figures, the MG-PCG
speed-up and the per-iteration compliance error against the FP64 path.

Prints ONE JSON line (see the keys below).  The CPU baseline is the FP64
C++/OpenMP restatement in oracle/ (the reference itself is MATLAB and cannot
run here), timed on a bounded sample on the host cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def parse_size(s):
    ny, nx, nz = (int(v) for v in s.lower().split("x"))
    return ny, nx, nz


# mixed: the precision form ('Standard' V(1,1), omega 0.4, FP16 coarse
# Galerkin matrices with per-cell scales, hierarchy deepened where the
# coarsest band Cholesky is costly: coarsestControl 0);
# mixed_fp32_coarse: the same with FP32 coarse matrices (what FP16 storage gains);
# fp64_standard: the precision form's cycle (and omega) in FP64, which
# separates what the FP32 arithmetic gains from what the cycle gains
VARIANTS = {"fp64": dict(precision=0), "mixed": dict(precision=2, vcycle=1, omega=0.4, coarsestControl=0),
            "mixed_adapted": dict(precision=1), "fp64_standard": dict(precision=0, vcycle=1, omega=0.4),
            "mixed_fp32_coarse": dict(precision=1, vcycle=1, omega=0.4, coarsestControl=0)}


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "traffic_fp64_256x512x256.md")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` (fetched + written) from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of the 256x512x256 FP64 run
    (counter values scaled by the copy-kernel calibration stated there), or
    None when the summary has no row for it."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            for line in f:
                cols = [c.strip() for c in line.split("|")]
                if len(cols) > 5 and kernel in cols[1]:
                    return round((float(cols[4]) + float(cols[5])) * 1e6)
    except (OSError, ValueError):
        return None
    return None


def make_model(args):
    """The workload's voxels and boundary conditions (None: built-in cuboid BCs)."""
    import numpy as np
    ny, nx, nz = parse_size(args.size)
    if args.workload == "cuboid":
        return np.ones((ny, nx, nz), dtype=bool), None, 0.12
    from top3d_xl_amd.irregular import generate
    g = generate(args.seed, (ny, nx, nz))
    return g["voxels"], g, 0.4


def gpu_run(args, params, model, sample_designs=False):
    import top3d_xl_amd as t3x

    vox, bc, V0 = model
    # HIP events around the fine K*p only (the roofline kernel); every group
    # (V-cycle, coarsest solve, assembly) with --timing-groups all, at ~1% of
    # a design iteration for the extra event pairs
    timing = 1 if args.timing_groups == "all" else 2
    if bc is None:
        P = t3x.Problem(vox, timing=timing, checkEvery=args.check_every, **params)
    else:
        P = t3x.Problem(vox, (bc["fix_nodes"], bc["fix_state"]), (bc["load_nodes"], bc["load_vec"]), timing=timing,
                        checkEvery=args.check_every, **params)
    nN, nE = P.nN, P.nE
    P.opt_begin("GLOBAL", V0, args.warmup + args.steps, math.sqrt(3))
    warm = []
    for _ in range(args.warmup):
        rec, _ = P.opt_step()
        warm.append(rec)
    designs = [(len(warm), P.xphys())] if sample_designs else []
    P.kernel_time_reset()
    t3x.device_synchronize()
    t0 = time.perf_counter()
    recs = []
    for _ in range(args.steps):
        rec, fin = P.opt_step()
        recs.append(rec)
        if fin:
            break
    t3x.device_synchronize()
    t1 = time.perf_counter()
    if sample_designs:
        designs.append((len(warm) + len(recs), P.xphys()))
    kn, kms = P.kernel_time(0)
    vn, vms = P.kernel_time(1)
    cn, cms = P.kernel_time(2)
    an, ams = P.kernel_time(3)
    P.close()
    steps = len(recs)
    its = sum(r["pcgIts"] for r in recs)
    cg_s = sum(r["t"][1] for r in recs)
    return dict(steps=steps, wall_s=t1 - t0, ms_per_step=1e3 * (t1 - t0) / steps, pcg_its=its,
                pcg_dof_per_s=3 * nN * its / cg_s, cg_s=cg_s, nN=nN, nE=nE,
                c_hist=[r["c"] for r in recs], its_hist=[r["pcgIts"] for r in recs],
                c_all=[r["c"] for r in warm + recs], its_all=[r["pcgIts"] for r in warm + recs],
                t_rows=[r["t"] for r in recs], applyK=(kn, kms), vcycle=(vn, vms), coarse=(cn, cms),
                assemble=(an, ams), designs=designs)


def solve_accuracy(model, designs, params):
    """Per-solve error of the precision form against the FP64 plain path on
    identical inputs: for each sampled design of the precision form's run, the
    elasticity system (SIMP modulus, 107/116; load F_, 1315-1321) is solved
    from zero to the reference's tolerance (1e-3, 30) by both forms and to
    1e-10 by the FP64 path (the converged displacement U*).  Compliance is
    U^T K U = sum E ce (125-127); displacement errors are 2-norm relative."""
    import numpy as np

    import top3d_xl_amd as t3x
    vox, bc, _ = model

    def problem(**kw):
        if bc is None:
            return t3x.Problem(vox, **kw)
        return t3x.Problem(vox, (bc["fix_nodes"], bc["fix_state"]), (bc["load_nodes"], bc["load_vec"]), **kw)

    P64, Pm = problem(precision=0, maxIT=5000), problem(maxIT=5000, **params)
    b = P64.load_vector()
    out = []
    for loop, x in designs:
        E = 1e-6 + x ** 3 * (1.0 - 1e-6)
        c_of = {}
        for P in (P64, Pm):
            P.set_modulus(E)
            P.assemble()
        U64, it64, _ = P64.pcg(b, tol=1e-3, maxit=5000)
        Um, itm, _ = Pm.pcg(b, tol=1e-3, maxit=5000)
        Us, its_, res_s = P64.pcg(b, x0=U64, tol=1e-10, maxit=5000)
        for k, U in (("fp64", U64), ("mixed", Um), ("converged", Us)):
            c_of[k] = float(E @ P64.compliance(U))
        ns = float(np.linalg.norm(Us))
        out.append({"design_iteration": loop, "pcg_its_fp64": it64, "pcg_its_mixed": itm,
                    "compliance_rel_err_mixed_vs_fp64": abs(c_of["mixed"] - c_of["fp64"]) / c_of["fp64"],
                    "displacement_rel_err_mixed_vs_fp64": float(np.linalg.norm(Um - U64) / np.linalg.norm(U64)),
                    "compliance_rel_err_vs_converged": {"mixed": abs(c_of["mixed"] - c_of["converged"]) / c_of["converged"],
                                                        "fp64": abs(c_of["fp64"] - c_of["converged"]) / c_of["converged"]},
                    "displacement_rel_err_vs_converged": {"mixed": float(np.linalg.norm(Um - Us)) / ns,
                                                          "fp64": float(np.linalg.norm(U64 - Us)) / ns},
                    "converged_pcg_its": it64 + its_})
        del U64, Um, Us
    P64.close()
    Pm.close()
    return out


def serving_leg(args):
    """The serving workload (a seeded Poisson stream of config-5 requests:
    cuboids (s, 2s, s), s over {32, ..., 192}, 70% GLOBAL nLoop 30 / 30%
    LOCAL nLoop 60) through the engine in the precision form (4 slots,
    continuous admission, pooled device memory) and through the one-at-a-time
    FP64 baseline on the same stream: solves/s and p50/p99 latency."""
    from top3d_xl_amd.serve_bench import make_stream, run_stream, summarize
    reqs = make_stream(args.serve_requests, args.serve_rate, seed=args.seed, nloop_scale=args.serve_nloop_scale)
    done, wall = run_stream(reqs, 4, **VARIANTS["mixed"])
    print("[bench] serving: engine done", file=sys.stderr, flush=True)
    done1, wall1 = run_stream(reqs, 1, precision=0)
    print("[bench] serving: one-at-a-time FP64 done", file=sys.stderr, flush=True)
    return {"requests": len(reqs), "rate_per_s": args.serve_rate, "nloop_scale": args.serve_nloop_scale,
            "sizes_s": sorted({r["shape"][0] for r in reqs}),
            "modes": {m: sum(1 for r in reqs if r["mode"] == m) for m in ("GLOBAL", "LOCAL")},
            "engine": dict(max_active=4, precision="precision form (mixed 'Standard', omega 0.4, FP16 coarse "
                                                    "matrices, coarsestControl 0)", **summarize(reqs, done, wall)),
            "one_at_a_time_fp64": summarize(reqs, done1, wall1)}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args, fp64_its_per_iteration):
    """FP64 oracle (C++/OpenMP restatement of TOP3D_XL.m) on bounded samples:
    at 256x512x256 one assembly (919-1066) and a PCG capped at --cpu-pcg-its
    iterations (600-639) on the uniform V0 design, extrapolated to a design
    iteration with the FP64 path's mean PCG iteration count; and the
    CPU-runnable 50x100x50 configuration for --cpu-iters design iterations."""
    import numpy as np

    from oracle import Oracle, build_oracle
    build_oracle()
    ny, nx, nz = parse_size(args.size)
    o = Oracle(np.ones((ny, nx, nz), dtype=bool))
    o.set_modulus(np.full(o.nE, 1e-6 + 0.12 ** 3 * (1 - 1e-6)))  # SIMP of x = V0 (107, 116)
    t0 = time.perf_counter()
    o.assemble()
    t1 = time.perf_counter()
    b = o.get("F")
    _, its, _ = o.pcg(b, tol=1e-30, maxit=args.cpu_pcg_its)
    t2 = time.perf_counter()
    del o, b
    # the capped PCG ran its set-up (one K*u and one V-cycle, 609-612) and
    # `its` iterations (one K*p and one V-cycle each, 614-629)
    ms_pcg = 1e3 * (t2 - t1) / (its + 1)
    dof = 3 * (ny + 1) * (nx + 1) * (nz + 1)
    ms_iter = 1e3 * (t1 - t0) + ms_pcg * fp64_its_per_iteration
    small = Oracle(np.ones((50, 100, 50), dtype=bool))
    r = small.run("GLOBAL", 0.12, args.cpu_iters, math.sqrt(3))
    its1 = int(r["lssIts"].sum())
    cg1 = float(r["tHist"][:, 1].sum())
    return dict(value=round(dof / (ms_pcg * 1e-3), 1), unit="DOF/s", cores=Oracle.num_threads(), kind="port",
                cpu=cpu_model(),
                sample=f"FP64 MG-PCG ('Adapted' V-cycle) of the {args.size} cuboid on the uniform V0 design: one "
                       f"assembly ({1e3 * (t1 - t0):.0f} ms) and a PCG capped at {its} iterations "
                       f"({ms_pcg:.0f} ms per PCG iteration, set-up counted as one)",
                ms_per_pcg_iteration=round(ms_pcg, 1), assemble_ms=round(1e3 * (t1 - t0), 1),
                ms_per_design_iteration_extrapolated=round(ms_iter, 1),
                extrapolation=f"assembly + {fp64_its_per_iteration:.1f} PCG iterations (the FP64 path's mean per "
                              f"design iteration in the timed window) x ms per PCG iteration; filter and OC "
                              f"not included",
                config1_dof_per_s=round(3 * small.nN * its1 / cg1, 1),
                config1_sample=f"MG-PCG of 50x100x50, {r['loops']} GLOBAL design iterations, {its1} PCG "
                               f"iterations; ms per design iteration {1e3 * float(r['tHist'][:, 5].mean()):.0f}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", default=None, help="nely x nelx x nelz (default 256x512x256; irregular 384x768x384)")
    ap.add_argument("--workload", default="cuboid", choices=["cuboid", "irregular"],
                    help="cuboid: built-in BCs, V0 0.12; irregular: seeded ellipsoid-union domain, V0 0.4")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"])
    ap.add_argument("--check-every", type=int, default=4)
    ap.add_argument("--cpu-iters", type=int, default=2)
    ap.add_argument("--cpu-pcg-its", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the per-solve error check of the precision form")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-workload leg")
    ap.add_argument("--serve-requests", type=int, default=12)
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--serve-nloop-scale", type=float, default=1.0)
    ap.add_argument("--timing-groups", default="kp", choices=["kp", "all"],
                    help="kernel groups timed with HIP events: the fine K*p only, or every group")
    ap.add_argument("--coarsest-control", type=int, default=None,
                    help="coarsestControl of the mixed-precision variants (<= 0: depth by coarsest band cost)")
    args = ap.parse_args()
    if args.coarsest_control is not None:
        for k in ("mixed", "mixed_adapted", "mixed_fp32_coarse"):
            VARIANTS[k] = dict(VARIANTS[k], coarsestControl=args.coarsest_control)
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "one GPU only: the solver runs on one MI355X; the multi-GPU design "
                                   "(x-slab decomposition) is documented, not run",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.size is None:
        args.size = "256x512x256" if args.workload == "cuboid" else "384x768x384"
    model = make_model(args)
    runs = {}
    names = {"fp64": ["fp64"], "mixed": ["mixed"],
             "both": ["fp64", "mixed", "mixed_adapted", "fp64_standard", "mixed_fp32_coarse"]}[args.precision]
    for name in names:
        print(f"[bench] {name}: {VARIANTS[name]}", file=sys.stderr, flush=True)
        runs[name] = gpu_run(args, VARIANTS[name], model, sample_designs=(name == "mixed" and not args.no_accuracy))
    head = "mixed" if "mixed" in runs else "fp64"
    g = runs[head]
    ny, nx, nz = parse_size(args.size)

    def summary(r):
        return {"ms_per_iteration": round(r["ms_per_step"], 3),
                "mg_pcg_ms_per_design_iteration": round(1e3 * r["cg_s"] / r["steps"], 3),
                "pcg_dof_per_s": round(r["pcg_dof_per_s"], 1), "pcg_iterations": r["pcg_its"],
                "pcg_ms_per_iteration": round(1e3 * r["cg_s"] / max(r["pcg_its"], 1), 4),
                "applyK_avg_ms": round(r["applyK"][1] / max(r["applyK"][0], 1), 4),
                "vcycle_avg_ms": round(r["vcycle"][1] / r["vcycle"][0], 4) if r["vcycle"][0] else None,
                "coarse_solve_avg_ms": round(r["coarse"][1] / r["coarse"][0], 4) if r["coarse"][0] else None,
                "assemble_avg_ms": round(r["assemble"][1] / r["assemble"][0], 3) if r["assemble"][0] else None,
                "phase_s_mean": {k: round(sum(t[i] for t in r["t_rows"]) / r["steps"], 4)
                                 for i, k in enumerate(["assemble", "cg", "opt", "filter", "lvf", "total"])},
                "compliance_all_iterations": r["c_all"], "pcg_its_all_iterations": r["its_all"]}

    kn, kms = g["applyK"]
    k_avg_s = kms / max(kn, 1) / 1e3
    # sum-factorised element operator: 196 flop per cell (transforms, coupling,
    # E scaling) + 21 adds per node (gathering 8 cells' rows)
    flops = 196.0 * g["nE"] + 21.0 * g["nN"]
    # the K*p of the PCG is FP64 in every form: read u, write y, read E
    bytes_ = 48.0 * g["nN"] + 8.0 * g["nE"]
    tflops = flops / k_avg_s / 1e12
    gbs = bytes_ / k_avg_s / 1e9
    # 8 flop per byte is below the FP64 ridge (78.6 TFLOP/s / 8 TB/s = 9.8): HBM-bound
    roofline = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4),
                # measured on this workload only (rocprofv3 PMC runs, committed summary)
                "traffic": measured_traffic("k_apply_K_sf<double")
                if args.workload == "cuboid" and args.size == "256x512x256" else None,
                "traffic_source": "profiles/traffic_fp64_256x512x256.md (FETCH_SIZE + WRITE_SIZE, bytes per launch)",
                "kernel": "k_apply_K_sf<double> (fine-level K*p inside PCG, %s)" % head, "launches": kn,
                "avg_ms": round(1e3 * k_avg_s, 4), "algorithmic_flop": flops,
                "algorithmic_bytes": bytes_, "achieved_TFLOPs_algorithmic": round(tflops, 3)}
    out = {
        "metric": "ms per TO iteration + MG-PCG DOF/s at 256x512x256; solves/s and p99 latency",
        "value": round(g["ms_per_step"], 3),
        "unit": "ms/iteration",
        "n_gpus": 1,
        "steps": g["steps"],
        "warmup": args.warmup,
        "ms_per_step": round(g["ms_per_step"], 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f64+f32",
        "data": "synthetic (built-in cuboid model, no dataset)" if args.workload == "cuboid" else
                "synthetic (seeded irregular domain, no dataset)",
        "config": {"workload": (f"GLOBAL topology optimisation, cuboid {ny}x{nx}x{nz} voxels (nely x nelx x nelz), "
                                f"{3 * g['nN']} DOF, V0=0.12, rMin=sqrt(3), built-in BCs")
                   if args.workload == "cuboid" else
                   (f"GLOBAL topology optimisation, irregular domain (seed {args.seed}: 40 ellipsoids in a "
                    f"{ny}x{nx}x{nz} box intersected with a cuboid shell, largest component, "
                    f"{g['nE']} solid voxels), {3 * g['nN']} DOF, V0=0.4, rMin=sqrt(3), bottom-slab fixed, "
                    f"unit load on top-slab surface nodes"),
                   "precision": head if head == "fp64" else
                   "mixed: FP64 PCG (K*p, vectors, dots) + FP32 'Standard' V(1,1) cycle (omega 0.4; "
                   "coarse Galerkin matrices stored FP16 with per-cell power-of-two scales) around the FP64 "
                   "coarsest Cholesky"},
        "pcg_dof_per_s": round(g["pcg_dof_per_s"], 1),
        "pcg_iterations": g["pcg_its"],
        "pcg_ms_per_iteration": round(1e3 * g["cg_s"] / max(g["pcg_its"], 1), 4),
        "roofline": roofline,
    }
    for k, r in runs.items():
        out[k] = summary(r)
    # practical HBM bandwidth: the library's streaming copy of one FP64 3-vector
    import top3d_xl_amd as t3x
    nb = 24 * g["nN"]
    ms = t3x.stream_copy_ms(nb, 20)
    out["stream_copy_GBs"] = round(2 * nb / (ms * 1e-3) / 1e9, 1)
    if "fp64" in runs:
        a = runs["fp64"]
        for k, b in runs.items():
            if k == "fp64":
                continue
            out[f"mg_pcg_speedup_{k}_vs_fp64"] = round((a["cg_s"] / a["steps"]) / (b["cg_s"] / b["steps"]), 3)
            out[f"iteration_speedup_{k}_vs_fp64"] = round(a["ms_per_step"] / b["ms_per_step"], 3)
            out[f"compliance_rel_err_{k}_vs_fp64"] = [abs(x - y) / abs(y) for x, y in zip(b["c_all"], a["c_all"])]
    if "fp64_standard" in runs and "mixed" in runs:
        a, b = runs["fp64_standard"], runs["mixed"]
        out["mg_pcg_speedup_mixed_vs_fp64_standard"] = round((a["cg_s"] / a["steps"]) / (b["cg_s"] / b["steps"]), 3)
    if "mixed" in runs and runs["mixed"]["designs"]:
        print("[bench] per-solve accuracy of the precision form", file=sys.stderr, flush=True)
        acc = solve_accuracy(model, runs["mixed"]["designs"], VARIANTS["mixed"])
        out["precision_form_per_solve_error"] = acc
        out["precision_form_max_compliance_rel_err_vs_fp64"] = max(a["compliance_rel_err_mixed_vs_fp64"] for a in acc)
        out["precision_form_max_displacement_rel_err_vs_fp64"] = max(a["displacement_rel_err_mixed_vs_fp64"]
                                                                     for a in acc)
    if not args.no_serving and args.workload == "cuboid":
        print("[bench] serving workload", file=sys.stderr, flush=True)
        out["serving"] = serving_leg(args)
    if not args.no_cpu_baseline and int(os.environ.get("RANK", "0")) == 0:
        ref = runs.get("fp64", g)
        out["cpu_baseline"] = cpu_baseline(args, ref["pcg_its"] / ref["steps"])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
