"""Pins of the oracle's rest of the SPH step (SmthParticleHydrodLiquid.cpp:
59-98, Collider.cpp:13-63, ParticlesBasedVectorField.cpp:56-69):

* a static plane collider: position put back at the surface + r, normal
  velocity reversed with the restitution, tangential velocity reduced by
  friction, clamped at zero (closed forms);
* the reference's box-complement collider keeps a particle inside the box;
* the divergence-free Laplacian is zero for a linear velocity field on a
  regular lattice (odd kernel gradient, symmetric neighbours) and tends to
  the continuum Laplacian (2, 0, 0) of v = (y^2, 0, 0) as the kernel covers
  more neighbours;
* advance = move + colliders, gravity, viscosity, generation, projection."""
import numpy as np
import pytest

import oracle_py as O
from vps import scenes, _abi

S = _abi


def lattice_scene(n, s, r, v_of):
    ax = (np.arange(n) + 0.5) * s
    g = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    z = np.zeros((0, 3))
    return g, scenes.grain_scene(g, v_of(g), r, z, z, z, np.full(3, -1.0),
                                 np.full(3, n * s + 1.0))


def single(x, v, r, dim=2):
    z = np.zeros((0, dim))
    sc = scenes.grain_scene(np.asarray([x], float), np.asarray([v], float), r, z, z, z,
                            np.full(3, -3.0), np.full(3, 3.0))
    return O.Oracle(sc)


@pytest.mark.parametrize("a,e,mu", [(0.4, 0.5, 0.2), (0.05, 0.0, 0.9), (0.0, 1.0, 0.3)])
def test_plane_collider_closed_form(a, e, mu):
    r, dt, b = 0.01, 1e-3, 2.0
    o = single([0.3, 0.5 * r], [a, -b], r)
    floor = S.collider(S.SURFACE_PLANE, [0, 0, 0], [0, 1, 0], restitution=e, friction=mu)
    o.move_particles(dt, [floor])
    x, v = o.particles()
    x_move = 0.3 + a * dt
    assert x[0, 0] == pytest.approx(x_move, rel=1e-15)
    assert x[0, 1] == pytest.approx(r, rel=1e-14)
    f = max(1 - mu * (e + 1) * b / abs(a), 0) if a != 0 else 1.0
    assert v[0, 0] == pytest.approx(a * f, rel=1e-14, abs=1e-300)
    assert v[0, 1] == pytest.approx(e * b, rel=1e-14, abs=1e-300)
    # moving away from the surface: only the position is corrected
    o2 = single([0.3, 0.5 * r], [a, b], r)
    o2.move_particles(dt, [floor])
    x2, v2 = o2.particles()
    assert np.array_equal(v2[0], [a, b]) and x2[0, 1] == pytest.approx(r, rel=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_box_complement_keeps_particles_inside(dim):
    r = 0.02
    box = S.collider(S.SURFACE_BOX, [-1, -1, -1], [2, 2, 2], complement=True)
    x0 = np.array([0.99, 0.3, -0.2])[:dim]
    v0 = np.array([3.0, -0.5, 0.25])[:dim]
    o = single(x0, v0, r, dim)
    o.move_particles(0.01, [box])
    x, v = o.particles()
    assert x[0, 0] == pytest.approx(1 - r, rel=1e-14)
    assert np.all(np.abs(x[0]) <= 1 - r + 1e-15)
    assert v[0, 0] == 0.0                       # restitution 0: normal velocity gone
    assert np.array_equal(v[0, 1:], v0[1:])      # friction 0


def _laplacian(n, s, r, v_of):
    """Laplacian at the particles whose neighbours (within h = 4r) all have
    complete neighbourhoods (uniform densities)."""
    g, sc = lattice_scene(n, s, r, v_of)
    o = O.Oracle(sc)
    o.move_particles(0.0)          # neighbour rebuild and densities
    o.apply_viscosity(1.0, 1.0)    # dt = nu = 1: v gains the Laplacian
    _, vn = o.particles()
    o.close()
    m = 2 * 4 * r + s
    inner = np.all((g > m) & (g < n * s - m), axis=1)
    assert inner.sum() >= 8
    return (vn - v_of(g))[inner]


def test_viscosity_zero_on_a_linear_field():
    lin = lambda g: np.stack([0.3 * g[:, 1] + 0.1 * g[:, 2], -0.2 * g[:, 0],
                              0.5 * g[:, 0] - 0.1 * g[:, 1]], 1)
    lap = _laplacian(14, 0.1, 0.05, lin)
    assert np.abs(lap).max() < 1e-12


def test_viscosity_tends_to_the_laplacian():
    quad = lambda g: np.stack([g[:, 1] ** 2, 0 * g[:, 0], 0 * g[:, 0]], 1)
    errs = []
    for r in (0.05, 0.075, 0.1):   # kernel radius 2, 3 and 4 lattice spacings
        lap = _laplacian(22, 0.1, r, quad)
        assert np.abs(lap[:, 1:]).max() < 1e-10
        errs.append(abs(lap[:, 0].mean() - 2.0))
    assert errs[0] > errs[1] > errs[2] and errs[2] < 0.03


def test_advance_is_the_sequence_of_its_parts():
    sc = scenes.small_3d(10, 6, 5)
    box = S.collider(S.SURFACE_BOX, [0, 0, 0], [1.0, 0.5, 0.45], complement=True,
                     restitution=0.2, friction=0.1)
    kw = dict(rest_density_mode=_abi.FROM_INIT, kappa_mode=_abi.FROM_INIT)
    a, b = O.Oracle(sc, **kw), O.Oracle(sc, **kw)
    a.initialize()
    b.initialize()
    for _ in range(2):
        sa = a.advance(1e-3, 9.8, 0.01, [box], enable_gravity=True)
        b.move_particles(1e-3, [box])
        b.apply_gravity(1e-3, 9.8)
        b.apply_viscosity(1e-3, 0.01)
        b.generate()
        sb = b.project()
        assert sa["cg_iterations"] == sb["cg_iterations"]
        xa, va = a.particles()
        xb, vb = b.particles()
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
