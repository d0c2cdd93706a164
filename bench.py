#!/usr/bin/env python3
"""Benchmark of the MI355X virtual-particle SPH projection.

Workload: the 3D dam break of 1,012,500 fluid particles (dx = 1/150, a
1 x 0.6 x 0.5 block in a 2 x 1 x 1 box, 3 wall layers), one GPU.  A step is
DualParticleSphLiquid::advance with the demos' box collider and gravity (no
viscosity): advection and collisions, real neighbour rebuild + densities,
virtual generation, right-hand side, Laplacian diagonal, CG to 1e-6,
pressure gradient.  All state is resident in HBM before the timed region.
The timed steps of the throughput runs are queued back to back
(vps_set_async_steps: a step returns once its work is on the stream; the
host still waits inside each step for the virtual-set counts and the CG's
convergence); their phase times come from synchronous steps right after.

Prints ONE JSON line.  `roofline` is the dominant kernel (the one-launch CG
iteration), timed live with HIP events on the library stream; `config3` is
the same scene at 8,100,000 particles; `warm_start_comparison` the learned
CG warm start against the other starts on its own schedule (500 training
steps, 500 measured); `field_errors` the GPU against the FP64 oracle in
lockstep on a bounded scene; `cpu_baseline` the FP64 oracle in its
reference-structured mode (hash searcher, full-lattice scan, serial triplet
assembly, CG on the transposed CSR matrix) on a bounded sample of the same
scene, with all host threads and with the reference demos' third of them.
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
        "extrapolate": 4, "learned_online": 3}


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
                    help="untimed steps before --warmup in the headline run")
    ap.add_argument("--c4-train", type=int, default=500,
                    help="warm-start comparison (the learned CG warm start at "
                         "1M): untimed steps of every run; the learned run "
                         "trains its network online during them")
    ap.add_argument("--c4-steps", type=int, default=500,
                    help="warm-start comparison: measured steps of every run")
    ap.add_argument("--c4-modes",
                    default="learned,learned_online,extrapolate,index_carry,node_keyed,zero",
                    help="warm starts of the comparison (learned_online: the learned start "
                         "that keeps training, one Adam step per solve, during the "
                         "measured steps)")
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
    ap.add_argument("--cpu-full", action="store_true",
                    help="also time the CPU baseline at the headline size "
                         "(1,012,500 particles; several minutes)")
    ap.add_argument("--c2-steps", type=int, default=200,
                    help="steps of the config-2 record run from the initial scene "
                         "(0 skips it)")
    ap.add_argument("--no-config3", action="store_true",
                    help="skip the 8,100,000-particle run (dx = 1/300)")
    ap.add_argument("--c3-steps", type=int, default=5)
    ap.add_argument("--no-field-errors", action="store_true",
                    help="skip the GPU vs FP64 oracle lockstep")
    ap.add_argument("--fe-dx", type=float, default=1.0 / 60,
                    help="spacing of the lockstep scene of the field errors")
    ap.add_argument("--fe-steps", type=int, default=3)
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
    rocprofv3 counters (FETCH_SIZE and WRITE_SIZE, calibrated on a plain
    copy), if committed under profiles/ for this workload."""
    path = os.path.join(ROOT, "profiles", "spmv_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        if d.get("kernel") == SPMV_KERNEL:
            return d.get("workloads", {}).get(workload, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        pass
    return None


def cpu_full_from_profiles():
    """The CPU baseline at the headline size (1,012,500 particles), measured
    by `bench.py --cpu-full` on the GPU box host and committed under
    profiles/ (it takes minutes, so the default run times a bounded sample)."""
    path = os.path.join(ROOT, "profiles", "cpu_baseline_full.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def _host_threads():
    """Host threads of this process: OMP_NUM_THREADS when set (the CPU share
    of the machine), else the CPU count."""
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()


def cpu_time_steps(sc, args, threads, steps):
    """Median time of `steps` projection steps (after 1 warm-up) of the
    reference-structured FP64 oracle on scene sc with `threads` OpenMP
    threads; each step is advection (untimed) + projection_step."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    from vps import _abi
    o = oracle_py.Oracle(sc, search=oracle_py.SEARCH_HASH, threads=threads,
                         rest_density_mode=_abi.FROM_PARAMS,
                         kappa_mode=_abi.FROM_PARAMS)
    o.initialize()
    times, iters = [], []
    for k in range(steps + 1):
        o.advect(args.dt, 9.8)
        t0 = time.perf_counter()
        st = o.projection_step()
        times.append(time.perf_counter() - t0)
        iters.append(st["cg_iterations"])
        log("cpu baseline: %d particles, %d threads, step %d: %.2f s, %d CG iterations"
            % (len(sc["x"]), threads, k, times[-1], iters[-1]))
    o.close()
    return float(np.median(times[1:])), iters[1:]


def cpu_baseline(args, dx=None, steps=None):
    """Reference-structured FP64 oracle on the same scene (at dx, by default
    the bounded sample's spacing): particle-steps/s with all host threads
    and with the reference demos' max(threads / 3, 1)
    (Demos/SmthPartHydrodLiquidTest/SmthPartHydrodLiquidTest.cpp:28)."""
    from vps import scenes
    dx = dx or args.cpu_dx
    steps = steps or args.cpu_steps
    sc = scenes.config2(seed=args.seed, dx=dx)
    n = len(sc["x"])
    all_t = _host_threads()
    demo_t = max(all_t // 3, 1)
    med_all, it_all = cpu_time_steps(sc, args, all_t, steps)
    med_demo, it_demo = cpu_time_steps(sc, args, demo_t, steps)
    return {
        "value": n / med_all,
        "unit": "particle-steps/s",
        "cores": all_t,
        "kind": "port",
        "sample": "FP64 oracle, reference-structured (OR-hash searcher, full "
                  "lattice scan, serial triplet assembly, CG on A^T CSR), "
                  "same dam-break scene at dx=%.5g (%d particles), median of "
                  "%d steps after 1 warm-up, %d OpenMP threads; CG iters %s"
                  % (dx, n, steps, all_t, it_all),
        "n_particles": n,
        "ms_per_step": round(1e3 * med_all, 1),
        "demo_threads": {"cores": demo_t, "value": n / med_demo,
                         "ms_per_step": round(1e3 * med_demo, 1),
                         "cg_iters": it_demo},
        "cpu_model": _cpu_model(),
        "nproc": os.cpu_count(),
    }


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def field_errors(args):
    """The GPU path and the FP64 oracle in lockstep on the same dam-break
    scene (bounded size, dx = --fe-dx) for --fe-steps steps: per output
    field, max |gpu - oracle| / max |oracle| and RMS(gpu - oracle) /
    RMS(oracle) after each step (the worst step is reported), for the real
    density and frozen volume (SmoothedParticles.cpp:25-36), the virtual
    density, volume and divergence (VirtualParticle.cpp:18-38, 171-201), the
    CG solution (IterativeSolver.h:34) and the updated velocity
    (VirtualParticle.cpp:110-134)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    from vps import Projection, scenes, _abi
    sc = scenes.config2(seed=args.seed, dx=args.fe_dx)
    kw = dict(rest_density_mode=_abi.FROM_PARAMS, kappa_mode=_abi.FROM_PARAMS)
    g = Projection(sc, **kw)
    o = oracle_py.Oracle(sc, threads=_host_threads(), **kw)
    g.initialize()
    o.initialize()
    worst = {}
    sets_equal = True
    iters = []

    def err(name, a, b):
        a = np.asarray(a, dtype=np.float64).ravel()
        b = np.asarray(b, dtype=np.float64).ravel()
        d = a - b
        mx = float(np.abs(d).max() / max(np.abs(b).max(), 1e-300))
        rms = float(np.sqrt(np.mean(d * d)) / max(np.sqrt(np.mean(b * b)), 1e-300))
        w = worst.setdefault(name, {"max_rel": 0.0, "rms_rel": 0.0})
        w["max_rel"] = max(w["max_rel"], mx)
        w["rms_rel"] = max(w["rms_rel"], rms)

    for _ in range(args.fe_steps):
        g.advect(args.dt, 9.8)
        o.advect(args.dt, 9.8)
        sg = g.projection_step()
        so = o.projection_step()
        iters.append([sg["cg_iterations"], so["cg_iterations"]])
        rg, vg = g.real_fields()
        ro, vo = o.real_fields()
        err("rho", rg, ro)
        err("V", vg, vo)
        Vg, Vo = g.virtual(), o.virtual()
        if not np.array_equal(Vg["node"], Vo["node"]):
            sets_equal = False
            break
        err("rho_I", Vg["rho"], Vo["rho"])
        err("V_I", Vg["vol"], Vo["vol"])
        err("D_I", Vg["div"], Vo["div"])
        err("p", Vg["p"], Vo["p"])
        _, velg = g.particles()
        _, velo = o.particles()
        err("v_next", velg, velo)
    g.close()
    o.close()
    return {"scene": "dam_break_3d_%d" % len(sc["x"]), "dx": args.fe_dx,
            "steps": args.fe_steps, "virtual_sets_equal": sets_equal,
            "cg_iters_gpu_oracle": iters,
            "fields": {k: {q: float("%.3g" % x) for q, x in v.items()} for k, v in worst.items()}}


def boundary_near_fluid(x, bx, h):
    """N_B': boundary particles with a fluid particle within h."""
    from scipy.spatial import cKDTree
    d, _ = cKDTree(x).query(bx, k=1, distance_upper_bound=h)
    return int(np.count_nonzero(np.isfinite(d)))


def step(proj, args, cols):
    """One step: DualParticleSphLiquid::advance (DualParticleSphLiquid.cpp:
    107-116) with the box collider, gravity and no viscosity: advection and
    collisions, neighbour rebuild and densities, gravity, virtual generation,
    projection."""
    return proj.advance(args.dt, gravity=9.8, viscosity=0.0, colliders=cols,
                        enable_gravity=True)


def run(args, sc, warm, train=False, kernel_timing=False, untimed=None, warmup=None,
        steps=None, keep_x=False, train_measured=0, async_steps=False, phase_steps=0):
    """One run from the initial scene: `untimed` steps (the learned run trains
    its network online during them), `warmup` more untimed steps with the
    network frozen, then `steps` timed steps, bracketed by device
    synchronisation.  train_measured > 0 keeps training during the timed
    steps with that many Adam steps per solve (their cost is timed).
    async_steps: the timed steps return once queued (vps_set_async_steps),
    so consecutive steps run back to back on the GPU; their phase times are
    then taken from `phase_steps` synchronous steps after the timed ones."""
    from vps import Projection, _abi, scenes
    cols = scenes.box_colliders(sc)
    untimed = args.train_steps if untimed is None else untimed
    warmup = args.warmup if warmup is None else warmup
    steps = args.steps if steps is None else steps
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
    for k in range(untimed):
        train_iters.append(step(proj, args, cols)["cg_iterations"])
    proj.synchronize()
    t_train = time.perf_counter() - t0
    adam = 0
    if train:
        loss, adam = proj.mlp_training_state()
        if train_measured > 0:
            proj.mlp_set_training(True, batch=args.batch, lr=args.lr,
                                  steps_per_solve=train_measured, seed=args.seed + 1)
        else:
            proj.mlp_set_training(False)
        log("%s: %d training steps, %d Adam steps, last loss %.3g"
            % (warm, untimed, adam, loss))
    for k in range(warmup):
        st = step(proj, args, cols)
        log("%s warmup %d: nv=%d iters=%d" % (warm, k, st["n_virtual"], st["cg_iterations"]))
    stats = []
    proj.synchronize()
    if async_steps:
        proj.set_async_steps(True)
    t0 = time.perf_counter()
    for k in range(steps):
        stats.append(step(proj, args, cols))
    proj.synchronize()
    elapsed = time.perf_counter() - t0
    proj.set_async_steps(False)
    phase_stats = [step(proj, args, cols) for k in range(phase_steps)]
    # per-launch HIP events on the library stream, after the timed steps
    # (the events themselves cost a little)
    timed_launch = []
    if kernel_timing:
        proj.set_kernel_timing(True)
        for k in range(5):
            timed_launch.append(step(proj, args, cols))
        proj.set_kernel_timing(False)
    nb = proj.nb
    x_end = proj.particles()[0] if keep_x else None
    proj.close()
    if args.verbose:
        for s in stats:
            log(json.dumps({k: (round(v, 4) if isinstance(v, float) else v)
                            for k, v in s.items()}))
    return dict(stats=stats, elapsed=elapsed, nb=nb, t_train=t_train, steps=steps,
                train_iters=train_iters, timed_launch=timed_launch, adam=adam,
                x_end=x_end, phase_stats=phase_stats or stats)


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
    roof = cg_roofline(kt, "sand_water_3d_%d_%d" % (n, ng))
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


def summary(r, n, steps=None):
    st = r["stats"]
    steps = steps or r["steps"]
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
        "fallbacks": int(sum(s["warm_start_fallback"] for s in st)),
    }


# Algorithmic bytes per CG iteration and virtual particle: read p, diag, V
# and the node map and write A p (20 B), the x / r update with r.r (24 B),
# the p update (12 B); the one-launch iteration kernel fuses them and moves
# 40 B (read p, r, t, x, V, diag; write p, r, t, x).
CG_BYTES = 56.0


def cells_of(sc):
    h = 4.0 * sc["particle_radius"]
    ext = (np.asarray(sc["domain_hi"]) - np.asarray(sc["domain_lo"]))[:3]
    return float(np.prod(np.ceil(ext / h)))


def phase_bytes(n, nv, nbp, cells, k):
    """Algorithmic bytes of each timed phase of a step (FP32 fields):
    sort + density 48 N + 8 C + 16 N; generation (virtual marking, fused
    rho_I, V_I, vbar, D_I) 12 N + 8 N_V + 32 N + 40 N_B' + 16 N_V; diagonal
    12 N_V; CG 32 N_V + 56 k N_V; gradient 36 N + 8 N_V + 40 N_B'.  Total
    144 N + 76 N_V + 80 N_B' + 8 C + 56 k N_V."""
    return {"ms_sort": 64.0 * n + 8.0 * cells,
            "ms_generate": 44.0 * n + 24.0 * nv + 40.0 * nbp,
            "ms_diag": 12.0 * nv,
            "ms_cg": 32.0 * nv + CG_BYTES * k * nv,
            "ms_gradient": 36.0 * n + 8.0 * nv + 40.0 * nbp}


def cg_roofline(kt, workload):
    """Roofline of the CG iteration kernel from its per-launch HIP events."""
    spmv = [s["ms_spmv_avg"] for s in kt if s["ms_spmv_avg"] > 0]
    if not spmv:
        return None
    t_launch = float(np.mean(spmv)) * 1e-3
    nv_kt = float(np.mean([s["n_virtual"] for s in kt]))
    b56, b40 = CG_BYTES * nv_kt, SPMV_BYTES * nv_kt
    a56, a40 = b56 / t_launch / 1e9, b40 / t_launch / 1e9
    return {"bound": "hbm", "achieved": round(a56, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(a56 / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles(workload),
            "kernel": SPMV_KERNEL,
            "avg_launch_us": round(t_launch * 1e6, 2),
            "bytes_per_launch": b56,
            "basis": "56 B per virtual particle per iteration (SpMV 20, x/r update 24, "
                     "p update 12)",
            "achieved_fused_40B": round(a40, 1),
            "frac_fused_40B": round(a40 / HBM_PEAK_GBS, 4),
            "n_virtual_per_launch": nv_kt,
            "slot_nodes_per_launch": 512 * float(np.mean([s["n_active_bricks"] for s in kt]))}


def size_record(args, r, sc, n, workload):
    """ms/step, CG iterations, phase times with their algorithmic GB/s, the
    CG kernel's roofline and the whole-step roofline of one run."""
    stats = r["stats"]
    ms_step = 1e3 * r["elapsed"] / r["steps"]
    nvm = float(np.mean([s["n_virtual"] for s in stats]))
    kmean = float(np.mean([s["cg_iterations"] for s in stats]))
    cells = cells_of(sc)
    nbp = boundary_near_fluid(r["x_end"], sc["bx"], 4.0 * sc["particle_radius"]) \
        if r["x_end"] is not None else 0
    pb = phase_bytes(n, nvm, nbp, cells, kmean)
    phases = {}
    for k, b in pb.items():
        ms = float(np.mean([s[k] for s in r.get("phase_stats") or stats]))
        phases[k] = {"ms": round(ms, 4), "algorithmic_GBs": round(b / (ms * 1e-3) / 1e9, 1)}
    b_step = sum(pb.values())
    return {
        "workload": workload, "n_real": n, "n_boundary": r["nb"], "n_boundary_near_fluid": nbp,
        "n_virtual_mean": round(nvm, 1), "cells": cells,
        "value": round(n * r["steps"] / r["elapsed"], 1),
        "ms_per_step": round(ms_step, 4),
        "cg_iters_mean": round(kmean, 2),
        "cg_iters": [s["cg_iterations"] for s in stats],
        "true_rel_residual_max": max(s["cg_true_rel_residual"] for s in stats),
        "cg_replacements": [s["cg_replacements"] for s in stats],
        "phases": phases,
        "roofline": cg_roofline(r["timed_launch"], workload),
        "step_roofline": {"bytes_per_step": b_step,
                          "achieved": round(b_step / (ms_step * 1e-3) / 1e9, 1),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(b_step / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "basis": "144 N + 76 N_V + 80 N_B' + 8 C + 56 k N_V, step time "
                                   "incl. advection"},
    }


def warm_start_comparison(args, sc, n):
    """Learned CG warm start at 1M (its own schedule): every run starts from
    the same scene, takes --c4-train untimed steps (the learned run trains
    online on them: Adam, batch --batch, lr --lr), then --c4-steps measured
    steps with the network frozen; CG always runs to 1e-6."""
    modes = [m for m in args.c4_modes.split(",") if m]
    runs = {}
    for m in modes:
        online = m == "learned_online"
        r = run(args, sc, "learned" if online else m, train=m.startswith("learned"),
                untimed=args.c4_train, warmup=0, steps=args.c4_steps,
                train_measured=1 if online else 0)
        d = summary(r, n)
        if m.startswith("learned"):
            d["training"] = {"steps": args.c4_train, "adam_steps": r["adam"],
                             "adam_per_solve": args.train_per_solve,
                             "batch": args.batch, "lr": args.lr,
                             "ms_per_step_while_training": round(
                                 1e3 * r["t_train"] / max(args.c4_train, 1), 4),
                             "cg_iters_while_training_first_last_50": [
                                 round(float(np.mean(r["train_iters"][:50])), 2),
                                 round(float(np.mean(r["train_iters"][-50:])), 2)]}
        runs[m] = d
    out = {"schedule": {"train_steps": args.c4_train, "measured_steps": args.c4_steps,
                        "dt": args.dt, "tolerance": 1e-6},
           "runs": runs}

    def ratio(a, b):
        A, B = runs[a], runs[b]
        return {"iters": round(B["cg_iters_mean"] / max(A["cg_iters_mean"], 1e-9), 3),
                "cg_ms": round(B["cg_ms_mean"] / max(A["cg_ms_mean"], 1e-9), 3),
                "projection_ms": round(B["projection_ms_mean"] / max(A["projection_ms_mean"], 1e-9), 3),
                "step_ms": round(B["ms_per_step"] / A["ms_per_step"], 3)}
    for a in modes:
        for b in modes:
            if a.startswith("learned") and not b.startswith("learned"):
                out[a + "_speedup_vs_" + b] = ratio(a, b)
    if "extrapolate" in runs and "index_carry" in runs:
        out["extrapolate_speedup_vs_index_carry"] = ratio("extrapolate", "index_carry")
    return out


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
    plain = run(args, sc, args.warm_start, kernel_timing=not args.no_kernel_timing,
                keep_x=True, async_steps=True, phase_steps=5)
    rec = size_record(args, plain, sc, n, workload)
    out = {
        "metric": METRIC,
        "value": rec["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": rec["ms_per_step"],
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
        "cg_iters_mean": rec["cg_iters_mean"],
        "cg_iters": rec["cg_iters"],
        "n_virtual_mean": rec["n_virtual_mean"],
        "n_boundary_near_fluid": rec["n_boundary_near_fluid"],
        "true_rel_residual_max": rec["true_rel_residual_max"],
        "cg_replacements": rec["cg_replacements"],
        # phase times of 5 synchronous steps after the timed (queued) ones
        "phase_ms_mean": {k: round(float(np.mean([s[k] for s in plain["phase_stats"]])), 4)
                          for k in ("ms_sort", "ms_generate", "ms_diag",
                                    "ms_cg", "ms_gradient", "ms_total")},
        "phases": rec["phases"],
        "roofline": rec["roofline"],
        "step_roofline": rec["step_roofline"],
        "cpu_baseline": None,
    }
    # config 2 as specified: 200 steps from the initial scene, plain start
    if args.c2_steps > 0:
        r200 = run(args, sc, args.warm_start, untimed=0, warmup=0, steps=args.c2_steps,
                   async_steps=True)
        s200 = summary(r200, n)
        rec200 = {k: s200[k] for k in ("value", "ms_per_step", "cg_iters_mean",
                                       "true_rel_residual_max")}
        rec200["steps"] = args.c2_steps
        rec200["cg_iters_first_last_50"] = [
            round(float(np.mean([q["cg_iterations"] for q in r200["stats"][:50]])), 2),
            round(float(np.mean([q["cg_iterations"] for q in r200["stats"][-50:]])), 2)]
        out["config2_200_steps"] = rec200
    if not args.no_config3:
        sc3 = scenes.config3(seed=args.seed)
        n3 = len(sc3["x"])
        r3 = run(args, sc3, args.warm_start, kernel_timing=not args.no_kernel_timing,
                 untimed=3, warmup=0, steps=args.c3_steps, keep_x=True,
                 async_steps=True, phase_steps=3)
        out["config3"] = size_record(args, r3, sc3, n3, "dam_break_3d_%d" % n3)
        out["config3"]["dx"] = 1.0 / 300
        del sc3
    if not args.no_learned:
        cmp = warm_start_comparison(args, sc, n)
        out["warm_start_comparison"] = cmp
        if "learned" in cmp["runs"]:
            out["learned_warm_start"] = cmp["runs"]["learned"]
        if "extrapolate" in cmp["runs"]:
            out["extrapolated_warm_start"] = cmp["runs"]["extrapolate"]
    if not args.no_field_errors:
        out["field_errors"] = field_errors(args)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args)
        if args.cpu_full:
            out["cpu_baseline_full"] = cpu_baseline(args, dx=args.dx, steps=3)
        else:
            full = cpu_full_from_profiles()
            if full is not None:
                full["note"] = "measured by bench.py --cpu-full on the GPU box host, not in this run"
                out["cpu_baseline_full_size"] = full
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
