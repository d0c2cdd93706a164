#!/usr/bin/env python3
"""Benchmark of the MI355X virtual-particle SPH projection.

Workload: the 3D dam break of 1,012,500 fluid particles (dx = 1/150, a
1 x 0.6 x 0.5 block in a 2 x 1 x 1 box, 3 wall layers), one GPU.  A step is
one advection (x += v dt, gravity) plus one projection as the reference
orders it: real neighbour rebuild + densities, virtual generation,
right-hand side, Laplacian diagonal, CG to 1e-6, pressure gradient.  All
state is resident in HBM before the timed region.

Prints ONE JSON line (see the fields below).  The roofline object is for
the dominant kernel (the CG SpMV), timed live with HIP events on the
library stream; cpu_baseline is the FP64 oracle in its reference-structured
mode (hash searcher, full-lattice scan, serial triplet assembly, CG on the
transposed CSR matrix) on a bounded sample of the same scene.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sand-water-coupling_amd"))

METRIC = "projection particle-steps/s + CG iters at 1e-6 residual, learned vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0

WARM = {"index_carry": 0, "zero": 1, "node_keyed": 2, "learned": 3}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--dx", type=float, default=1.0 / 150)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--warm-start", default="index_carry", choices=sorted(WARM))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-dx", type=float, default=1.0 / 60,
                    help="lattice spacing of the CPU-baseline sample")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--verbose", action="store_true")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profiles(workload):
    """HBM bytes per SpMV launch measured with rocprofv3 counters, if a
    summary for this workload is committed under profiles/."""
    path = os.path.join(ROOT, "profiles", "spmv_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        if d.get("workload") == workload:
            return d.get("bytes_per_launch")
    except (OSError, ValueError):
        pass
    return None


def cpu_baseline(args):
    """Reference-structured FP64 oracle on the same scene at a coarser
    spacing (bounded sample), median step time -> particle-steps/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    from vps import scenes, _abi
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    sc = scenes.config2(seed=args.seed, dx=args.cpu_dx)
    o = oracle_py.Oracle(sc, search=oracle_py.SEARCH_HASH, threads=threads,
                         rest_density_mode=_abi.FROM_PARAMS,
                         kappa_mode=_abi.FROM_PARAMS)
    o.initialize()
    times = []
    iters = []
    for _ in range(args.cpu_steps + 1):
        o.advect(args.dt, 9.8)
        t0 = time.perf_counter()
        st = o.projection_step()
        times.append(time.perf_counter() - t0)
        iters.append(st["cg_iterations"])
    times = times[1:]   # first step is the warm-up
    med = float(np.median(times))
    o.close()
    return {
        "value": len(sc["x"]) / med,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": "FP64 oracle, reference-structured (OR-hash searcher, full "
                  "lattice scan, serial triplet assembly, CG on A^T CSR), "
                  "same dam-break scene at dx=%.5g (%d particles), median of "
                  "%d steps after 1 warm-up, %d OpenMP threads; CG iters %s"
                  % (args.cpu_dx, len(sc["x"]), len(times), threads, iters[1:]),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU "
              "path (the sharding design is described in DESIGN.md); "
              "run with --gpus 1", file=sys.stderr)
        return 2
    from vps import Projection, scenes, _abi
    sc = scenes.config2(seed=args.seed, dx=args.dx)
    n = len(sc["x"])
    workload = "dam_break_3d_%d" % n
    proj = Projection(sc, warm_start=WARM[args.warm_start],
                      rest_density_mode=_abi.FROM_PARAMS,
                      kappa_mode=_abi.FROM_PARAMS)
    t0 = time.perf_counter()
    proj.initialize()
    log("init %.3f s, n=%d nb=%d" % (time.perf_counter() - t0, n, proj.nb))
    for k in range(args.warmup):
        proj.advect(args.dt, 9.8)
        st = proj.projection_step()
        log("warmup %d: nv=%d iters=%d" % (k, st["n_virtual"], st["cg_iterations"]))
    proj.set_kernel_timing(not args.no_kernel_timing)
    stats = []
    proj.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        proj.advect(args.dt, 9.8)
        stats.append(proj.projection_step())
    proj.synchronize()
    elapsed = time.perf_counter() - t0
    ms_step = 1e3 * elapsed / args.steps
    iters = [s["cg_iterations"] for s in stats]
    nv = [s["n_virtual"] for s in stats]
    if args.verbose:
        for s in stats:
            log(json.dumps({k: (round(v, 4) if isinstance(v, float) else v)
                            for k, v in s.items()}))
    # roofline of the CG SpMV (dominant kernel): algorithmic bytes per launch
    # = 32 B per virtual particle (p, diag, V, slot map read and M p written:
    # 20 B; the fused p = r + beta p update: 12 B)
    spmv = [s["ms_spmv_avg"] for s in stats if s["ms_spmv_avg"] > 0]
    roof = None
    if spmv:
        t_launch = float(np.mean(spmv)) * 1e-3
        bytes_launch = 32.0 * float(np.mean(nv))
        ach = bytes_launch / t_launch / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": traffic_from_profiles(workload),
                "kernel": "k_stencil3<SPMV_FUSED, TRANSPOSE>",
                "avg_launch_us": round(t_launch * 1e6, 2),
                "bytes_per_launch": bytes_launch}
    # whole-step algorithmic bytes (144N + 76N_V + 80N_B' + 8C + 56 k N_V)
    out = {
        "metric": METRIC,
        "value": round(n * args.steps / elapsed, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 sums/SpMV, f64 dots+membership",
        "data": "synthetic",
        "config": {"workload": workload, "n_real": n, "n_boundary": proj.nb,
                   "dx": args.dx, "dt": args.dt, "seed": args.seed,
                   "warm_start": args.warm_start, "tolerance": 1e-6,
                   "parallelism": "single GPU"},
        "cg_iters_mean": round(float(np.mean(iters)), 2),
        "cg_iters": iters,
        "n_virtual_mean": round(float(np.mean(nv)), 1),
        "true_rel_residual_max": max(s["cg_true_rel_residual"] for s in stats),
        "phase_ms_mean": {k: round(float(np.mean([s[k] for s in stats])), 4)
                          for k in ("ms_sort", "ms_generate", "ms_diag",
                                    "ms_cg", "ms_gradient", "ms_total")},
        "roofline": roof,
        "cpu_baseline": None,
    }
    proj.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
