"""bench.py contract checks that need no GPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_multi_gpu_request_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def _bench():
    sys.path.insert(0, ROOT)
    import bench
    return bench


def _fake_stats(nv, k, nslots):
    return {"n_virtual": nv, "cg_iterations": k, "n_active_bricks": nslots,
            "cg_true_rel_residual": 9e-7, "cg_replacements": 0, "warm_start_fallback": 0,
            "ms_sort": 0.2, "ms_generate": 0.4, "ms_diag": 0.03, "ms_cg": 0.9,
            "ms_gradient": 0.3, "ms_total": 1.83, "ms_spmv_avg": 0.02}


def test_step_accounting_matches_the_survey_formula():
    """The phase bytes add up to 144 N + 76 N_V + 80 N_B' + 8 C + 56 k N_V and
    the CG roofline reports both bases from the same launch time."""
    import numpy as np
    from vps import scenes
    b = _bench()
    n, nv, nbp, cells, k = 1000.0, 1200.0, 50.0, 300.0, 10.0
    pb = b.phase_bytes(n, nv, nbp, cells, k)
    assert sum(pb.values()) == 144 * n + 76 * nv + 80 * nbp + 8 * cells + 56 * k * nv
    sc = scenes.small_3d(8, 6, 4)
    st = [_fake_stats(1500, 12, 6) for _ in range(4)]
    r = dict(stats=st, elapsed=0.01, steps=4, nb=len(sc["bx"]), x_end=sc["x"],
             timed_launch=st)
    rec = b.size_record(None, r, sc, len(sc["x"]), "w")
    roof = rec["roofline"]
    assert roof["avg_launch_us"] == 20.0
    assert abs(roof["achieved"] / roof["achieved_fused_40B"] - 56 / 40) < 1e-3
    assert 0 < rec["n_boundary_near_fluid"] < len(sc["bx"])
    # N_B' by brute force
    h = 4 * sc["particle_radius"]
    d2 = ((sc["bx"][:, None, :] - sc["x"][None, :, :]) ** 2).sum(-1)
    assert rec["n_boundary_near_fluid"] == int(np.count_nonzero((d2 < h * h).any(1)))
    assert set(rec["phases"]) == {"ms_sort", "ms_generate", "ms_diag", "ms_cg", "ms_gradient"}


def test_cpu_baseline_reports_both_thread_counts():
    import argparse
    b = _bench()
    args = argparse.Namespace(cpu_dx=1.0 / 20, cpu_steps=1, seed=7, dt=5e-4)
    d = b.cpu_baseline(args)
    assert d["kind"] == "port" and d["value"] > 0
    assert d["demo_threads"]["cores"] == max(d["cores"] // 3, 1)
    assert d["demo_threads"]["value"] > 0
