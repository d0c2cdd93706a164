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

WARM = {"index_carry": 0, "zero": 1, "node_keyed": 2, "learned": 3,
        "extrapolate": 4}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--dx", type=float, default=1.0 / 150)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--warm-start", default="index_carry", choices=sorted(WARM),
                    help="warm start of the headline (plain) run")
    ap.add_argument("--train-steps", type=int, default=30,
                    help="untimed steps before --warmup in every run; the "
                         "learned run trains its network online during them")
    ap.add_argument("--train-per-solve", type=int, default=4,
                    help="Adam steps per solve while training")
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned / extrapolated warm-start runs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-dx", type=float, default=1.0 / 60,
                    help="lattice spacing of the CPU-baseline sample")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--workload", default="dam_break", choices=["dam_break", "sand_water"],
                    help="dam_break: the 3D dam break (the headline line); "
                         "sand_water: water column falling on a DEM sand bed, "
                         "grains as moving boundary particles, learned warm start")
    ap.add_argument("--sand-jitter", type=float, default=0.02,
                    help="grain packing jitter in units of dx")
    ap.add_argument("--cpu-sand-dx", type=float, default=1.0 / 40,
                    help="spacing of the CPU sample of the sand-water step")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB "
                         "(3 times) and exit: the reference kernel that "
                         "calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE")
    return ap.parse_args()


def copy_probe(mib):
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": 4 * n, "repeats": 3}), flush=True)
    return 0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# the one-launch 3D CG iteration (update of the previous iteration + SpMV)
SPMV_KERNEL = "k_cg_iter3<TRANSPOSE>"
SPMV_BYTES = 40.0


def traffic_from_profiles(workload):
    """HBM bytes per launch of the CG iteration kernel measured with
    rocprofv3 counters, if a summary for this workload and kernel is
    committed under profiles/."""
    path = os.path.join(ROOT, "profiles", "spmv_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        if d.get("workload") == workload and d.get("kernel") == SPMV_KERNEL:
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


def run(args, sc, warm, train=False, kernel_timing=False):
    """One run from the initial scene: train_steps + warmup untimed steps
    (the learned run trains online during the first train_steps), then
    `steps` timed steps, bracketed by device synchronisation."""
    from vps import Projection, _abi
    proj = Projection(sc, warm_start=WARM[warm],
                      rest_density_mode=_abi.FROM_PARAMS,
                      kappa_mode=_abi.FROM_PARAMS)
    t0 = time.perf_counter()
    proj.initialize()
    log("%s: init %.3f s, n=%d nb=%d" % (warm, time.perf_counter() - t0, proj.n, proj.nb))
    if train:
        proj.mlp_set_training(True, batch=args.batch, lr=args.lr,
                              steps_per_solve=args.train_per_solve, seed=args.seed)
    train_iters = []
    proj.synchronize()
    t0 = time.perf_counter()
    for k in range(args.train_steps):
        proj.advect(args.dt, 9.8)
        train_iters.append(proj.projection_step()["cg_iterations"])
    proj.synchronize()
    t_train = time.perf_counter() - t0
    if train:
        loss, adam = proj.mlp_training_state()
        proj.mlp_set_training(False)
        log("%s: %d training steps, %d Adam steps, last loss %.3g"
            % (warm, args.train_steps, adam, loss))
    for k in range(args.warmup):
        proj.advect(args.dt, 9.8)
        st = proj.projection_step()
        log("%s warmup %d: nv=%d iters=%d" % (warm, k, st["n_virtual"], st["cg_iterations"]))
    stats = []
    proj.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        proj.advect(args.dt, 9.8)
        stats.append(proj.projection_step())
    proj.synchronize()
    elapsed = time.perf_counter() - t0
    # per-launch HIP events on the library stream, after the timed steps
    # (the events themselves cost a little)
    timed_launch = []
    if kernel_timing:
        proj.set_kernel_timing(True)
        for k in range(5):
            proj.advect(args.dt, 9.8)
            timed_launch.append(proj.projection_step())
        proj.set_kernel_timing(False)
    nb = proj.nb
    proj.close()
    if args.verbose:
        for s in stats:
            log(json.dumps({k: (round(v, 4) if isinstance(v, float) else v)
                            for k, v in s.items()}))
    return dict(stats=stats, elapsed=elapsed, nb=nb, t_train=t_train,
                train_iters=train_iters, timed_launch=timed_launch)


def sand_water_scenes(dx, seed, jitter):
    """2M water over a 1M-grain bed at dx = 1/150: a 1 x 0.3 x 1 bed of
    grains of radius dx/2, a 0.6 deep water column 0.05 above it, in a
    1 x 1.2 x 1 box."""
    from vps import scenes
    return scenes.sand_water(3, dx, (1.0, 1.2, 1.0), bed=0.3, water=0.6, gap=0.05,
                             seed=seed, jitter=jitter)


def sand_water_step(grains, water, dt):
    """DEM step of the grains, grains into the water's boundary, advection,
    projection; returns the projection stats and the host-timed phases."""
    t0 = time.perf_counter()
    grains.advance(dt)
    t1 = time.perf_counter()
    water.couple_grains(grains)
    t2 = time.perf_counter()
    water.advect(dt, 9.8)
    st = water.projection_step()
    st["ms_dem"] = 1e3 * (t1 - t0)
    st["ms_couple"] = 1e3 * (t2 - t1)
    return st


def cpu_baseline_sand_water(args):
    """The same coupled step on the FP64 oracle (DEM step, boundary =
    walls + grains, advection, projection) at a coarser spacing."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    from vps import _abi
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    ws, gs = sand_water_scenes(args.cpu_sand_dx, args.seed, args.sand_jitter)
    og = oracle_py.Oracle(gs, threads=threads)
    w2 = dict(ws)
    w2["bx"] = np.concatenate([ws["bx"], gs["x"]])
    w2["bn"] = np.concatenate([ws["bn"], np.zeros_like(gs["x"])])
    w2["bv"] = np.concatenate([ws["bv"], gs["v"]])
    ow = oracle_py.Oracle(w2, threads=threads, warm_start=_abi.WARM_EXTRAPOLATE,
                          rest_density_mode=_abi.FROM_PARAMS, kappa_mode=_abi.FROM_PARAMS)
    ow.initialize()
    times = []
    for _ in range(args.cpu_steps + 1):
        t0 = time.perf_counter()
        og.dem_step(args.dt, 9.8)
        xg, vg = og.particles()
        ow.set_boundary(np.concatenate([ws["bx"], xg]),
                        np.concatenate([ws["bn"], np.zeros_like(xg)]),
                        np.concatenate([ws["bv"], vg]))
        ow.advect(args.dt, 9.8)
        ow.projection_step()
        times.append(time.perf_counter() - t0)
    times = times[1:]
    med = float(np.median(times))
    n = len(ws["x"])
    og.close()
    ow.close()
    return {
        "value": n / med,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": "FP64 oracle coupled step (DEM step, walls + grains boundary "
                  "refresh, advection, projection with the extrapolated warm "
                  "start) on the same scene at dx=%.5g (%d water, %d grains), "
                  "median of %d steps after 1 warm-up, %d OpenMP threads"
                  % (args.cpu_sand_dx, n, len(gs["x"]), len(times), threads),
    }


def main_sand_water(args):
    from vps import DemSand, Projection, _abi
    ws, gs = sand_water_scenes(args.dx, args.seed, args.sand_jitter)
    n, ng = len(ws["x"]), len(gs["x"])
    grains = DemSand(gs)
    water = Projection(ws, warm_start=WARM["learned"],
                       rest_density_mode=_abi.FROM_PARAMS, kappa_mode=_abi.FROM_PARAMS)
    t0 = time.perf_counter()
    water.couple_grains(grains)
    water.initialize()
    log("sand_water: init %.3f s, water %d, grains %d, boundary %d"
        % (time.perf_counter() - t0, n, ng, water.nb))
    water.mlp_set_training(True, batch=args.batch, lr=args.lr,
                           steps_per_solve=args.train_per_solve, seed=args.seed)
    train_iters = []
    for k in range(args.train_steps):
        train_iters.append(sand_water_step(grains, water, args.dt)["cg_iterations"])
    water.mlp_set_training(False)
    for k in range(args.warmup):
        st = sand_water_step(grains, water, args.dt)
        log("sand_water warmup %d: nv=%d iters=%d" % (k, st["n_virtual"], st["cg_iterations"]))
    stats = []
    water.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        stats.append(sand_water_step(grains, water, args.dt))
    water.synchronize()
    elapsed = time.perf_counter() - t0
    kt = []
    if not args.no_kernel_timing:
        water.set_kernel_timing(True)
        for k in range(5):
            kt.append(sand_water_step(grains, water, args.dt))
        water.set_kernel_timing(False)
    _, vg = grains.particles()
    grains.close()
    water.close()
    ms_step = 1e3 * elapsed / args.steps
    iters = [s["cg_iterations"] for s in stats]
    spmv = [s["ms_spmv_avg"] for s in kt if s["ms_spmv_avg"] > 0]
    roof = None
    if spmv:
        t_launch = float(np.mean(spmv)) * 1e-3
        bytes_launch = SPMV_BYTES * float(np.mean([s["n_virtual"] for s in kt]))
        ach = bytes_launch / t_launch / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": SPMV_KERNEL,
                "avg_launch_us": round(t_launch * 1e6, 2),
                "bytes_per_launch": bytes_launch}
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
        "dtype": "f32 sums/SpMV/DEM forces, f64 dots+membership+positions",
        "data": "synthetic",
        "config": {"workload": "sand_water_3d_%d_%d" % (n, ng), "n_water": n,
                   "n_grains": ng, "n_boundary": len(ws["bx"]) + ng, "dx": args.dx,
                   "dt": args.dt, "seed": args.seed, "warm_start": "learned",
                   "grain_jitter_dx": args.sand_jitter, "tolerance": 1e-6,
                   "untimed_steps": args.train_steps + args.warmup,
                   "parallelism": "single GPU"},
        "cg_iters_mean": round(float(np.mean(iters)), 2),
        "cg_iters": iters,
        "cg_iters_while_training": train_iters,
        "true_rel_residual_max": max(s["cg_true_rel_residual"] for s in stats),
        "cg_replacements": [s["cg_replacements"] for s in stats],
        "phase_ms_mean": {k: round(float(np.mean([s[k] for s in stats])), 4)
                          for k in ("ms_dem", "ms_couple", "ms_sort", "ms_generate",
                                    "ms_diag", "ms_cg", "ms_gradient", "ms_total")},
        "grain_speed_max": float(np.abs(vg).max()),
        "roofline": roof,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_sand_water(args)
    print(json.dumps(out), flush=True)
    return 0


def summary(r, n, steps):
    st = r["stats"]
    return {
        "value": round(n * steps / r["elapsed"], 1),
        "ms_per_step": round(1e3 * r["elapsed"] / steps, 4),
        "cg_iters_mean": round(float(np.mean([s["cg_iterations"] for s in st])), 2),
        "cg_ms_mean": round(float(np.mean([s["ms_cg"] for s in st])), 4),
        # the projection proper: diagonal, warm start + CG, gradient (and the
        # warm start's bookkeeping after the solve), without the advection,
        # sort, densities and virtual generation that every variant shares
        "projection_ms_mean": round(float(np.mean([s["ms_diag"] + s["ms_cg"] + s["ms_gradient"]
                                                   for s in st])), 4),
        "true_rel_residual_max": max(s["cg_true_rel_residual"] for s in st),
        "replacements_mean": round(float(np.mean([s["cg_replacements"] for s in st])), 2),
    }


def main():
    args = parse()
    if args.copy_probe_mib > 0:
        return copy_probe(args.copy_probe_mib)
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU "
              "path; run with --gpus 1", file=sys.stderr)
        return 2
    if args.workload == "sand_water":
        return main_sand_water(args)
    from vps import scenes
    sc = scenes.config2(seed=args.seed, dx=args.dx)
    n = len(sc["x"])
    workload = "dam_break_3d_%d" % n
    plain = run(args, sc, args.warm_start, kernel_timing=not args.no_kernel_timing)
    stats = plain["stats"]
    elapsed = plain["elapsed"]
    ms_step = 1e3 * elapsed / args.steps
    iters = [s["cg_iterations"] for s in stats]
    nv = [s["n_virtual"] for s in stats]
    # roofline of the CG iteration kernel (dominant kernel of the plain step):
    # algorithmic bytes per launch = 40 B per virtual particle (read p_old,
    # r_old, t_old, x, V, diag; write p, r, t, x, all FP32)
    kt = plain["timed_launch"]
    spmv = [s["ms_spmv_avg"] for s in kt if s["ms_spmv_avg"] > 0]
    roof = None
    if spmv:
        t_launch = float(np.mean(spmv)) * 1e-3
        nv_kt = float(np.mean([s["n_virtual"] for s in kt]))
        bytes_launch = SPMV_BYTES * nv_kt
        ach = bytes_launch / t_launch / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": traffic_from_profiles(workload),
                "kernel": SPMV_KERNEL,
                "avg_launch_us": round(t_launch * 1e6, 2),
                "bytes_per_launch": bytes_launch,
                "slot_nodes_per_launch": 512 * float(np.mean([s["n_active_bricks"] for s in kt]))}
    # whole-step algorithmic bytes: 144 N + 76 N_V + 8 C + 40 k N_V (cell
    # sort, densities, virtual marking and sums, diagonal, CG init, k CG
    # iterations of the 40 B one-launch update + SpMV, gradient; the
    # boundary term 80 N_B' is left out, so the fraction is a lower bound)
    h = 2.0 * sc["particle_radius"] * 2.0
    cells = float(np.prod(np.ceil((np.asarray(sc["domain_hi"]) - np.asarray(sc["domain_lo"]))[:3] / h)))
    nvm = float(np.mean(nv))
    kmean = float(np.mean(iters))
    b_step = 144.0 * n + 76.0 * nvm + 8.0 * cells + SPMV_BYTES * kmean * nvm
    step_roof = {"bytes_per_step": b_step, "achieved": round(b_step / (ms_step * 1e-3) / 1e9, 1),
                 "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(b_step / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
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
        "config": {"workload": workload, "n_real": n, "n_boundary": plain["nb"],
                   "dx": args.dx, "dt": args.dt, "seed": args.seed,
                   "warm_start": args.warm_start, "tolerance": 1e-6,
                   "untimed_steps": args.train_steps + args.warmup,
                   "parallelism": "single GPU"},
        "cg_iters_mean": round(float(np.mean(iters)), 2),
        "cg_iters": iters,
        "n_virtual_mean": round(float(np.mean(nv)), 1),
        "true_rel_residual_max": max(s["cg_true_rel_residual"] for s in stats),
        "cg_replacements": [s["cg_replacements"] for s in stats],
        "phase_ms_mean": {k: round(float(np.mean([s[k] for s in stats])), 4)
                          for k in ("ms_sort", "ms_generate", "ms_diag",
                                    "ms_cg", "ms_gradient", "ms_total")},
        "roofline": roof,
        "step_roofline": step_roof,
        "cpu_baseline": None,
    }
    if not args.no_learned:
        base = summary(plain, n, args.steps)
        lr = run(args, sc, "learned", train=True)
        ex = run(args, sc, "extrapolate")
        L = summary(lr, n, args.steps)
        E = summary(ex, n, args.steps)
        for d in (L, E):
            d["iters_ratio_vs_plain"] = round(base["cg_iters_mean"] / max(d["cg_iters_mean"], 1e-9), 3)
            d["step_speedup_vs_plain"] = round(base["ms_per_step"] / d["ms_per_step"], 3)
            d["cg_speedup_vs_plain"] = round(base["cg_ms_mean"] / max(d["cg_ms_mean"], 1e-9), 3)
            d["projection_speedup_vs_plain"] = round(
                base["projection_ms_mean"] / max(d["projection_ms_mean"], 1e-9), 3)
        L["training"] = {"steps": args.train_steps, "adam_per_solve": args.train_per_solve,
                         "batch": args.batch, "lr": args.lr,
                         "ms_per_step_while_training": round(1e3 * lr["t_train"] / max(args.train_steps, 1), 4),
                         "cg_iters_while_training": lr["train_iters"]}
        out["learned_warm_start"] = L
        out["extrapolated_warm_start"] = E
        out["plain_cg_ms_mean"] = base["cg_ms_mean"]
        out["plain_projection_ms_mean"] = base["projection_ms_mean"]
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
