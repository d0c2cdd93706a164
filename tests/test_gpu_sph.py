"""The rest of the SPH step on the GPU against the FP64 oracle, through the
C ABI (needs an MI355X): advection with colliders (every kind of surface,
restitution and friction), gravity, the divergence-free viscosity and the
whole DualParticleSphLiquid::advance.  Single passes start the oracle from
the GPU's state (FP32 velocities), so positions compare to 1e-12 and
velocities to FP32 rounding; whole steps follow the projection's bars."""
import numpy as np
import pytest

import oracle_py as O
from vps import Projection, VpsError, scenes, _abi

pytestmark = pytest.mark.gpu

INIT = dict(rest_density_mode=_abi.FROM_INIT, kappa_mode=_abi.FROM_INIT)
S = _abi


def colliders(dim, box_hi):
    return [
        S.collider(S.SURFACE_BOX, [0, 0, 0], box_hi, complement=True, restitution=0.3,
                   friction=0.2),
        S.collider(S.SURFACE_SPHERE, [0.3, 0.1, 0.1][:dim] + [0] * (3 - dim), [0.06, 0, 0],
                   restitution=0.5, friction=0.1),
        S.collider(S.SURFACE_PLANE, [0, 0.02, 0], [0.1, 1, 0.05], restitution=0.0,
                   friction=0.5),
    ]


def scene(dim):
    sc = scenes.small_3d(12, 8, 6) if dim == 3 else scenes.small_2d(20)
    rng = np.random.default_rng(11)
    sc = dict(sc, v=rng.normal(scale=0.8, size=sc["x"].shape))
    box_hi = list(np.asarray(sc["domain_hi"]) * 0 + 1.0)
    return sc, box_hi


@pytest.mark.parametrize("dim", [2, 3])
def test_move_collide_and_viscosity_passes(dim):
    sc, _ = scene(dim)
    x0 = sc["x"]
    lo, hi = x0.min(axis=0), x0.max(axis=0)
    box = [hi[a] - lo[a] for a in range(dim)] + [0] * (3 - dim)
    cols = [S.collider(S.SURFACE_BOX, list(lo) + [0] * (3 - dim), box, complement=True,
                       restitution=0.3, friction=0.2),
            S.collider(S.SURFACE_SPHERE, list((lo + hi) / 2) + [0] * (3 - dim),
                       [0.1, 0, 0], restitution=0.5, friction=0.1)]
    g = Projection(sc, **INIT)
    g.initialize()
    o = O.Oracle(sc, **INIT)
    o.initialize()
    dt = 4e-3
    hits = 0
    for step in range(3):
        xs, vs = g.particles()
        o.set_particles(xs, vs)
        g.move_particles(dt, cols)
        o.move_particles(dt, cols)
        xg, vg = g.particles()
        xo, vo = o.particles()
        assert np.abs(xg - xo).max() < 1e-12, step
        assert np.abs(vg - vo).max() <= 2e-7 * np.abs(vo).max(), step
        hits += int(np.sum(np.any(vg != vs, axis=1)))
        rg, _ = g.real_fields()
        ro = o.real_fields()[0]
        assert np.max(np.abs(rg - ro) / ro) < 2e-6, step
        o.set_particles(xg, vg)
        o.move_particles(0.0)            # same positions: densities only
        g.apply_viscosity(dt, 0.01)
        o.apply_viscosity(dt, 0.01)
        _, vg2 = g.particles()
        _, vo2 = o.particles()
        dv = np.abs(vo2 - vg).max()
        assert dv > 0, step
        assert np.abs(vg2 - vo2).max() <= 1e-4 * dv + 2e-7 * np.abs(vo2).max(), step
        g.apply_gravity(dt, 9.8)
        o.apply_gravity(dt, 9.8)
    assert hits > 10   # the colliders acted
    g.close()
    o.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_advance_steps(dim):
    sc, _ = scene(dim)
    sc = dict(sc, v=np.zeros_like(sc["x"]))
    box = [1.0, 0.5, 0.45] if dim == 3 else [1.0, 1.0, 0]
    cols = [S.collider(S.SURFACE_BOX, [0, 0, 0], box, complement=True, restitution=0.2,
                       friction=0.1)]
    g = Projection(sc, **INIT)
    g.initialize()
    o = O.Oracle(sc, **INIT)
    o.initialize()
    for step in range(4):
        sg = g.advance(1e-3, 9.8, 0.01, cols, enable_gravity=True)
        so = o.advance(1e-3, 9.8, 0.01, cols, enable_gravity=True)
        assert abs(sg["cg_iterations"] - so["cg_iterations"]) <= 1, step
        assert sg["cg_true_rel_residual"] < 1e-6, step
        vgv, vov = g.virtual(), o.virtual()
        assert np.array_equal(vgv["node"], vov["node"]), step
        xg, vg = g.particles()
        xo, vo = o.particles()
        assert np.linalg.norm(vg - vo) / np.linalg.norm(vo) < 1e-4, step
        assert np.abs(xg - xo).max() < 1e-9, step
    bad = [S.collider(9, [0, 0, 0], [1, 1, 1])]
    with pytest.raises(VpsError):
        g.advance(1e-3, colliders=bad)
    with pytest.raises(VpsError):
        g.move_particles(1e-3, [cols[0]] * 9)
    g.close()
    o.close()
