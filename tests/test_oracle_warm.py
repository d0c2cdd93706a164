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


def test_learned_start_is_the_galerkin_projection():
    """The learned start of an untrained network: x0 lies in the span of the
    node-keyed pressures P and the extrapolation's increments E, and its
    residual b - M x0 is orthogonal to both (the Galerkin condition
    V^T (b - M V y) = 0 with the CG's own matrix M)."""
    import scipy.sparse as sp
    sc = scenes.small_3d(12, 8, 6)
    o = O.Oracle(sc, warm_start=_abi.WARM_LEARNED, rest_density_mode=_abi.FROM_INIT,
                 kappa_mode=_abi.FROM_INIT)
    o.initialize()
    hist = []
    checked = 0
    for k in range(6):
        o.advect(1e-3, 9.8)
        st = o.projection_step()
        v = o.virtual()
        x0 = o.initial_guess()
        rp, col, val = o.matrix()
        n = len(v["node"])
        M = sp.csr_matrix((val, col, rp), shape=(n, n))
        P = np.zeros(n)
        E = np.zeros(n)
        for I, nd in enumerate(v["node"].tolist()):
            past = []
            for j in range(len(hist) - 1, max(len(hist) - 4, -1), -1):
                if nd not in hist[j]:
                    break
                past.append(hist[j][nd])
            if past:
                P[I] = past[0]
                E[I] = (2 * past[0] - 3 * past[1] + past[2] - 0.0) if len(past) == 3 \
                    else (past[0] - past[1] if len(past) == 2 else 0.0)
        hist.append(dict(zip(v["node"].tolist(), v["p"].tolist())))
        if k < 2 or not np.any(E):
            continue
        # E above is the polynomial's increment over the last value (2 dp1 -
        # dp2 with three known pressures).  A start rejected by the guard is
        # the index-carried one, outside the span: not this check's case.
        y, *_ = np.linalg.lstsq(np.stack([P, E], 1), x0, rcond=None)
        if np.linalg.norm(x0 - y[0] * P - y[1] * E) > 1e-10 * np.linalg.norm(x0):
            continue
        r0 = v["div"] - M @ x0
        for f in (P, E):
            assert abs(r0 @ f) <= 1e-8 * np.linalg.norm(r0) * np.linalg.norm(f)
        # and the weights are not the plain extrapolation's (1, 1)
        assert abs(y[0] - 1) + abs(y[1] - 1) > 1e-6
        checked += 1
    o.close()
    assert checked >= 3
