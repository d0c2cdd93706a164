"""Pins of the oracle's node-history warm starts (the extrapolation that the
learned start corrects): the initial guess is the degree-(h-1) polynomial
through the node's last h pressures, evaluated one step ahead (numpy.polyfit
as the independent routine); depth and node keying against the solutions
read back after every step."""
import numpy as np

import oracle_py as O
from vps import scenes, _abi


def _run(warm, steps=5):
    sc = scenes.small_3d(12, 8, 6)
    o = O.Oracle(sc, warm_start=warm, rest_density_mode=_abi.FROM_INIT,
                 kappa_mode=_abi.FROM_INIT)
    o.initialize()
    hist, guesses, iters = [], [], []
    for _ in range(steps):
        o.advect(1e-3, 9.8)
        st = o.projection_step()
        v = o.virtual()
        guesses.append(o.initial_guess())
        hist.append(dict(zip(v["node"].tolist(), v["p"].tolist())))
        iters.append(st["cg_iterations"])
        v["node_list"] = v["node"].tolist()
        hist[-1]["_order"] = v["node_list"]
    o.close()
    return hist, guesses, iters


def test_extrapolation_is_the_polynomial_through_the_history():
    hist, guesses, _ = _run(_abi.WARM_EXTRAPOLATE)
    checked = {1: 0, 2: 0, 3: 0}
    for k in range(1, len(hist)):
        nodes = hist[k]["_order"]
        x0 = guesses[k]
        for I, nd in enumerate(nodes):
            past = []
            for j in range(k - 1, max(k - 4, -1), -1):
                if nd not in hist[j]:
                    break
                past.append(hist[j][nd])
            if not past:
                assert x0[I] == 0.0
                continue
            vals = past[::-1]                   # oldest first, at t = -h+1..0
            t = np.arange(-len(vals) + 1, 1, dtype=float)
            coef = np.polyfit(t, vals, len(vals) - 1)
            want = np.polyval(coef, 1.0)
            scale = max(abs(v) for v in vals) + 1e-30
            assert abs(x0[I] - want) <= 1e-9 * scale, (k, nd, len(vals))
            checked[len(vals)] += 1
    assert min(checked.values()) > 50, checked


def test_node_keyed_and_depth_one_agree():
    hist_n, g_n, it_n = _run(_abi.WARM_NODE_KEYED, steps=2)
    hist_e, g_e, it_e = _run(_abi.WARM_EXTRAPOLATE, steps=2)
    # second solve: every node has at most one known pressure
    assert np.array_equal(g_n[1], g_e[1])
    assert it_n == it_e
