"""GPU DEM grain step and grains as moving boundary particles of the water
projection, against the FP64 oracle through the C ABI (needs an MI355X).

DEM step: each step starts the oracle from the GPU's state (the grains are
chaotic, and the GPU keeps FP32 velocities), then compares positions
bit-exact (x + v dt on the same velocities, both rounded twice) and
velocities within 1e-4 relative (norm over all grains, and per grain against
the largest velocity change of the step).

Coupling: the water context's boundary is its walls followed by the grains
(zero normals); the oracle water gets the same boundary from the GPU's grain
state every step.  Bars as for the plain projection: virtual set bit-exact,
boundary fields within 2e-6, CG to 1e-6 true residual, velocities within
1e-4 relative."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from vps import DemSand, Projection, scenes, _abi

pytestmark = pytest.mark.gpu

YOUNG, POISSON, PHI = 1e5, 0.3, 0.5
INIT = dict(rest_density_mode=_abi.FROM_INIT, kappa_mode=_abi.FROM_INIT)


def small(dim):
    if dim == 3:
        return scenes.sand_water(3, 1.0 / 40, (0.5, 0.6, 0.25), bed=0.15, water=0.3,
                                 jitter=0.1, grain_speed=0.2)
    return scenes.sand_water(2, 1.0 / 60, (1.0, 1.0), bed=0.2, water=0.4,
                             jitter=0.1, grain_speed=0.2)


@pytest.mark.parametrize("dim", [2, 3])
def test_dem_steps(dim):
    _, gs = small(dim)
    g = DemSand(gs, young=YOUNG, poisson=POISSON, friction_angle=PHI)
    o = O.Oracle(gs)
    dt = 2e-4
    for step in range(6):
        x0, v0 = g.particles()
        o.set_particles(x0, v0)
        g.advance(dt)
        o.dem_step(dt, 9.8, YOUNG, POISSON, PHI)
        xg, vg = g.particles()
        xo, vo = o.particles()
        assert np.array_equal(xg, xo), step
        dv = np.abs(vo - v0).max()
        assert dv > 1e-3, step                       # forces acted
        assert np.linalg.norm(vg - vo) / np.linalg.norm(vo) < 1e-4, step
        assert np.abs(vg - vo).max() / dv < 1e-4, step
    g.close()
    o.close()


def test_dem_errors():
    _, gs = small(3)
    g = DemSand(gs)
    assert g.L.vps_dem_step(g.h, 1e-4, 9.8, None) == _abi.VPS_ERR_INVALID_ARG
    for dt in (float("nan"), -1.0):
        assert g.L.vps_dem_step(g.h, dt, 9.8, C.byref(g.dem)) == _abi.VPS_ERR_INVALID_ARG
    bad = _abi.dem_params(young=0.0)
    assert g.L.vps_dem_step(g.h, 1e-4, 9.8, C.byref(bad)) == _abi.VPS_ERR_INVALID_ARG
    g.advance(1e-4)                 # the context still works
    w2, _ = scenes.sand_water(2, 1.0 / 20, (1.0, 1.0), bed=0.2, water=0.4)
    water2 = Projection(w2)
    assert g.L.vps_couple_grains(water2.h, g.h) == _abi.VPS_ERR_INVALID_ARG
    g.close()
    water2.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_coupled_projection(dim):
    ws, gs = small(dim)
    grains = DemSand(gs, young=YOUNG, poisson=POISSON, friction_angle=PHI)
    water = Projection(ws, **INIT)
    water.couple_grains(grains)
    water.initialize()
    nw = len(ws["bx"])

    def boundary(x, v):
        bx = np.concatenate([ws["bx"], x])
        bn = np.concatenate([ws["bn"], np.zeros_like(x)])
        bv = np.concatenate([ws["bv"], v])
        return bx, bn, bv

    xg, vg = grains.particles()
    s2 = dict(ws)
    s2["bx"], s2["bn"], s2["bv"] = boundary(xg, vg)
    o = O.Oracle(s2, **INIT)
    o.initialize()
    rg, vg_ = water.boundary_fields()
    ro, vo_ = o.boundary_fields()
    assert len(rg) == nw + len(xg)
    assert np.max(np.abs(rg - ro) / ro) < 2e-6
    assert np.max(np.abs(vg_ - vo_) / vo_) < 2e-6
    dt = 1e-3 if dim == 3 else 5e-4
    for step in range(4):
        grains.advance(dt)
        water.couple_grains(grains)
        xg, vg = grains.particles()
        o.set_boundary(*boundary(xg, vg))
        rg, _ = water.boundary_fields()
        ro, _ = o.boundary_fields()
        assert np.max(np.abs(rg - ro) / ro) < 2e-6, step
        water.advect(dt, 9.8)
        o.advect(dt, 9.8)
        sg = water.projection_step()
        so = o.projection_step()
        vgv, vov = water.virtual(), o.virtual()
        assert np.array_equal(vgv["node"], vov["node"]), step
        assert abs(sg["cg_iterations"] - so["cg_iterations"]) <= 1, step
        assert sg["cg_true_rel_residual"] < 1e-6, step
        assert np.linalg.norm(vgv["p"] - vov["p"]) / np.linalg.norm(vov["p"]) < 1e-4, step
        xw, v_w = water.particles()
        xo, v_o = o.particles()
        assert np.linalg.norm(v_w - v_o) / np.linalg.norm(v_o) < 1e-4, step
        assert np.abs(xw - xo).max() < 1e-9, step
    grains.close()
    water.close()
    o.close()
