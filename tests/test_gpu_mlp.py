"""Learned CG warm start: the hand-written MFMA network vs the FP64 oracle
network, its features vs the projection's own fields, and the warm start in
use (needs an MI355X).

BF16 tolerance: weights and layer inputs are rounded to bfloat16 in both
the kernels and the oracle (bf=True); what remains is FP32 accumulation order,
the float erf/exp and bf16 rounding of the back-propagated dZ (not rounded in
the oracle): a few 1e-3 relative."""
import numpy as np
import pytest

import mlp_oracle as M
from vps import Projection, scenes, _abi

pytestmark = pytest.mark.gpu


def _ctx():
    g = Projection(scenes.small_3d(12, 8, 6))
    return g


def _features(n, seed):
    rng = np.random.default_rng(seed)
    f = rng.normal(size=(n, 8))
    f[:, 7] = rng.choice([0.0, 0.5, 1.0], n)
    return f.astype(np.float32)


def test_parameter_roundtrip_and_count():
    g = _ctx()
    p0 = g.mlp_parameters()
    assert p0.size == M.N_PARAMS == 4801
    # untrained: output layer zero, hidden layers Xavier-uniform
    assert np.all(p0[4736:] == 0)
    assert 0 < np.abs(p0[:512]).max() <= np.sqrt(6 / 72) + 1e-6
    p = M.random_params(5).astype(np.float32)
    g.mlp_set_parameters(p)
    assert np.array_equal(g.mlp_parameters(), p)
    bad = p.copy()
    bad[7] = np.inf
    with pytest.raises(Exception):
        g.mlp_set_parameters(bad)
    g.close()


@pytest.mark.parametrize("n", [1, 16, 1000, 40_003])
def test_forward_matches_oracle(n):
    g = _ctx()
    p = M.random_params(11).astype(np.float32)
    g.mlp_set_parameters(p)
    X = _features(n, 12)
    out = g.mlp_forward(X).astype(np.float64)
    ref, _ = M.forward(p.astype(np.float64), X.astype(np.float64), bf=True)
    ref64, _ = M.forward(p.astype(np.float64), X.astype(np.float64), bf=False)
    scale = max(np.sqrt(np.mean(ref ** 2)), 1e-3)
    err = np.abs(out - ref)
    print("n=%d max|d|/rms=%.3e rms=%.3e vs fp64 %.3e" % (
        n, err.max() / scale, np.sqrt(np.mean(err ** 2)) / scale,
        np.sqrt(np.mean((out - ref64) ** 2)) / scale))
    assert err.max() < 2e-2 * scale
    assert np.sqrt(np.mean(err ** 2)) < 3e-3 * scale
    # the full-precision network differs only by bf16 rounding (a few ulp of
    # bf16, 2^-8, per layer; averaged over many samples)
    tol64 = 2e-2 if n >= 1000 else 6e-2
    assert np.sqrt(np.mean((out - ref64) ** 2)) < tol64 * scale
    g.close()


BLOCKS = {"W1": slice(0, 512), "b1": slice(512, 576), "W2": slice(576, 4672),
          "b2": slice(4672, 4736), "W3": slice(4736, 4800), "b3": slice(4800, 4801)}


@pytest.mark.parametrize("n", [16, 5003, 65536])
def test_loss_and_gradient_match_oracle(n):
    g = _ctx()
    p = M.random_params(21).astype(np.float32)
    g.mlp_set_parameters(p)
    X = _features(n, 22)
    t = np.random.default_rng(23).normal(size=n).astype(np.float32)
    loss, grad = g.mlp_loss_and_grad(X, t)
    lref, gref = M.loss_and_grad(p.astype(np.float64), X.astype(np.float64),
                                 t.astype(np.float64), bf=True)
    assert loss == pytest.approx(lref, rel=5e-3)
    for name, s in BLOCKS.items():
        d = np.linalg.norm(grad[s] - gref[s]) / np.linalg.norm(gref[s])
        print("n=%d %s rel err %.3e" % (n, name, d))
        assert d < 2e-2, name
    # the gradient is the one of the loss (a transposed W2 would fail here)
    W2g = grad[BLOCKS["W2"]].reshape(64, 64)
    W2r = gref[BLOCKS["W2"]].reshape(64, 64)
    assert np.linalg.norm(W2g - W2r) < 0.2 * np.linalg.norm(W2g.T - W2r)
    g.close()


def _node_map(virt, vals):
    return dict(zip(virt["node"].tolist(), vals.tolist()))


def test_features_are_the_projection_fields():
    sc = scenes.small_3d(16, 10, 8)
    g = Projection(sc, warm_start=_abi.WARM_LEARNED)
    g.initialize()
    prev = []
    for step in range(4):
        g.advect(1e-3, 9.8)
        g.projection_step()
        virt = g.virtual()
        f = g.mlp_features()
        scal = g.scalars()
        nodes = virt["node"].tolist()
        assert f.shape == (len(nodes), 8)
        np.testing.assert_allclose(f[:, 4], virt["diag"] / scal["alpha0"] - 1,
                                   rtol=1e-5, atol=1e-5)
        vs = sc["mass"] / scal["rest_density"]
        np.testing.assert_allclose(f[:, 6], virt["vol"] / vs - 1, rtol=1e-5, atol=1e-5)
        assert np.all(f[:, 5] <= f[:, 4] + 1e-6)     # unclamped <= clamped
        if prev:
            last = prev[-1]
            ps = np.sqrt(np.mean(last["p"] ** 2))
            m1 = _node_map(last, last["p"])
            known = np.array([n in m1 for n in nodes])
            p1 = np.array([m1.get(n, 0.0) for n in nodes])
            np.testing.assert_allclose(f[known, 0] * ps, p1[known],
                                       rtol=1e-4, atol=1e-6 * np.abs(p1).max())
            assert np.all(f[~known, :3] == 0) and np.all(f[~known, 7] == 0)
            # depth feature: 0, 0.5, 1 for one, two, three known pressures
            depth = np.zeros(len(nodes))
            for I, n in enumerate(nodes):
                d = 0
                for j in range(len(prev) - 1, -1, -1):
                    if n not in prev[j]["map"]:
                        break
                    d += 1
                depth[I] = min(d, 3)
            np.testing.assert_array_equal(f[:, 7], np.select(
                [depth >= 3, depth == 2], [1.0, 0.5], 0.0))
            # the right-hand side enters scaled by one constant
            c = np.dot(f[:, 3], virt["div"]) / np.dot(f[:, 3], f[:, 3])
            np.testing.assert_allclose(f[:, 3] * c, virt["div"], rtol=1e-4,
                                       atol=1e-5 * np.abs(virt["div"]).max())
            if len(prev) >= 2:
                m2 = prev[-2]["map"]
                both_last = [n for n in last["node"].tolist() if n in m2]
                dps = np.sqrt(np.mean([(m1[n] - m2[n]) ** 2 for n in both_last]))
                both = depth >= 2
                dp1 = np.array([m1.get(n, 0.0) - m2.get(n, 0.0) for n in nodes])
                np.testing.assert_allclose(f[both, 1] * dps, dp1[both], rtol=1e-3,
                                           atol=1e-4 * np.abs(dp1).max())
                assert np.all(f[depth < 2, 1] == 0)
        virt["map"] = _node_map(virt, virt["p"])
        prev.append(virt)
    g.close()


def _lockstep(sc, modes, steps, train_steps, steps_per_solve=4, batch=65536):
    """Contexts with the given warm starts run in lockstep on identical
    inputs: every context takes the first one's particles each step.  The
    LEARNED context trains online for the first train_steps steps."""
    kw = dict(rest_density_mode=_abi.FROM_PARAMS, kappa_mode=_abi.FROM_PARAMS)
    ctx = [Projection(sc, warm_start=m, **kw) for m in modes]
    for g, m in zip(ctx, modes):
        g.initialize()
        if m == _abi.WARM_LEARNED:
            g.mlp_set_training(True, batch=batch, lr=1e-3,
                               steps_per_solve=steps_per_solve, seed=7)
    its, ms, dv, res, loss = [], [], [], [], []
    for step in range(steps):
        if step == train_steps:
            for g, m in zip(ctx, modes):
                if m == _abi.WARM_LEARNED:
                    g.mlp_set_training(False)
        ctx[0].advect(5e-4, 9.8)
        x, v = ctx[0].particles()
        for g in ctx[1:]:
            g.set_particles(x, v)
        st = [g.projection_step() for g in ctx]
        vel = [g.particles()[1] for g in ctx]
        its.append([s["cg_iterations"] for s in st])
        ms.append([s["ms_cg"] for s in st])
        res.append([s["cg_true_rel_residual"] for s in st])
        dv.append([np.linalg.norm(u - vel[0]) / np.linalg.norm(vel[0]) for u in vel])
        loss.append(max([g.mlp_training_state()[0] for g, m in zip(ctx, modes)
                         if m == _abi.WARM_LEARNED] or [0.0]))
    for g in ctx:
        g.close()
    return (np.array(its), np.array(ms), np.array(dv), np.array(res),
            np.array(loss))


def test_learned_warm_start_cuts_iterations():
    sc = scenes.config2(dx=1.0 / 40)
    modes = [_abi.WARM_INDEX_CARRY, _abi.WARM_EXTRAPOLATE, _abi.WARM_LEARNED]
    its, ms, dv, res, loss = _lockstep(sc, modes, 70, 50)
    for i in range(len(its)):
        print("step %2d iters plain/extrap/learned %3d %3d %3d  cg ms %.2f %.2f %.2f"
              "  dv %.1e %.1e  loss %.3e" % ((i,) + tuple(its[i]) + tuple(ms[i]) +
                                             tuple(dv[i, 1:]) + (loss[i],)))
    late = its[-15:].mean(0)
    trans = its[15:50].sum(0)
    print("late mean iterations plain %.2f extrapolated %.2f learned %.2f" % tuple(late))
    print("iterations over steps 15-49: plain %d extrapolated %d learned %d" % tuple(trans))
    # equal residual, so equal answers
    assert np.all(res < 1e-6)
    assert np.all(dv < 1e-4)
    # the loss is in units of the extrapolation error (1 for a zero
    # correction): the trained network removes a share of that error ...
    assert loss[20:50].mean() < 0.8
    # ... which shows while the flow is still changing
    assert trans[2] <= 0.8 * trans[1]
    # against the reference's warm start (index carry)
    assert late[2] * 2 <= late[0]
    # after training stops: the frozen network's Galerkin start against the
    # extrapolation (runs are deterministic, so the comparison is stable)
    assert late[2] < late[1]


def test_overflowing_network_falls_back_to_the_plain_start():
    """Finite parameters whose output overflows FP32 give a non-finite x0 at
    every node with a history.  Each such solve must run again from the
    reference's index-carried start and match a plain context step for step
    (same iterations, same velocities); training on the overflowing network
    must skip every Adam step (non-finite loss) and leave it unchanged."""
    sc = scenes.small_3d(16, 10, 8)
    plain = Projection(sc, warm_start=_abi.WARM_INDEX_CARRY)
    learned = Projection(sc, warm_start=_abi.WARM_LEARNED)
    p = M.random_params(5).astype(np.float32)
    p[4736:4801] = 3e38   # W3 and b3: the output sum overflows
    learned.mlp_set_parameters(p)
    learned.mlp_set_training(True, batch=4096, lr=1e-3, steps_per_solve=2, seed=3)
    plain.initialize()
    learned.initialize()
    fallbacks = 0
    for step in range(5):
        plain.advect(1e-3, 9.8)
        learned.advect(1e-3, 9.8)
        sp = plain.projection_step()
        sl = learned.projection_step()
        fallbacks += sl["warm_start_fallback"]
        assert sp["warm_start_fallback"] == 0
        assert sl["cg_iterations"] == sp["cg_iterations"], step
        assert sl["cg_true_rel_residual"] < 1e-6
        _, vp = plain.particles()
        _, vl = learned.particles()
        assert np.array_equal(vl, vp), step
    # the first solve has no history (x0 = 0 in both); every later one fell back
    assert fallbacks == 4
    assert np.array_equal(learned.mlp_parameters(), p)
    _, steps = learned.mlp_training_state()
    assert steps == 0
    plain.close()
    learned.close()
