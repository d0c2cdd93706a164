"""Pins of the oracle's DEM grain step (DEMParticleSand::advance,
Cores/Physics/DEMParticleSand.cpp:90-152, with DEMParticle::ComputeDemForces,
Cores/Structures/DEMParticle.cpp:69-96) and of the grain normal of the
pressure-gradient boundary term, against closed forms:

* a head-on pair: penetration after the move, normal spring only;
* a sheared pair: Coulomb cap tan(phi) |F_n| on the tangential spring;
* one grain over one wall particle: V_b = 1 / W(0), the velocity-only wall
  force without dt, its friction cap and its one-sidedness;
* a gap inside the capillary range but beyond 2r: no force;
* pairwise antisymmetry: momentum conserved exactly up to rounding;
* hash searcher vs brute force;
* a grain (zero normal) acts on the fluid like a wall particle whose normal
  is the unit vector from the grain to the fluid particle, and pushes the
  fluid away from itself.
The cubic spline below is the textbook one (Monaghan 1992 / the reference's
SmoothedParticles.h:60-66 normalisation), written independently."""
import math

import numpy as np
import pytest

import oracle_py as O
from vps import scenes, _abi

YOUNG, POISSON, PHI = 1e5, 0.3, 0.5
G = 9.8


def spline(dim, r, h):
    q = r / h
    sigma = 40 / (7 * math.pi * h * h) if dim == 2 else 8 / (math.pi * h ** 3)
    if q < 0.5:
        return sigma * (6 * q ** 3 - 6 * q * q + 1)
    if q < 1:
        return sigma * 2 * (1 - q) ** 3
    return 0.0


def dem_oracle(x, v, r, bx=None, bn=None, bv=None, search=O.SEARCH_HASH):
    dim = x.shape[1]
    if bx is None:
        bx = np.zeros((0, dim))
        bn = np.zeros((0, dim))
        bv = np.zeros((0, dim))
    pts = np.concatenate([x, bx]) if len(bx) else x
    lo = pts.min(axis=0) - 10 * r
    hi = pts.max(axis=0) + 10 * r
    sc = scenes.grain_scene(x, v, r, bx, bn, bv, np.pad(lo, (0, 3 - dim)),
                            np.pad(hi, (0, 3 - dim)))
    return O.Oracle(sc, search=search)


@pytest.mark.parametrize("dim", [2, 3])
def test_head_on_pair(dim):
    r, dt, u, delta = 0.01, 1e-4, 0.01, 0.1 * 0.01
    x = np.zeros((2, dim))
    x[1, 0] = 2 * r - delta
    v = np.zeros((2, dim))
    v[0, 0], v[1, 0] = u, -u
    o = dem_oracle(x, v, r)
    o.dem_step(dt, G, YOUNG, POISSON, PHI)
    xn, vn = o.particles()
    assert xn[0, 0] == u * dt and xn[1, 0] == (2 * r - delta) - u * dt
    pen = 2 * r - (xn[1, 0] - xn[0, 0])
    kn = YOUNG * r
    exp = np.zeros((2, dim))
    exp[0, 0] = u - kn * pen * dt
    exp[1, 0] = -u + kn * pen * dt
    exp[:, 1] -= G * dt
    assert np.allclose(vn, exp, rtol=1e-12, atol=1e-15)
    o.close()


@pytest.mark.parametrize("w", [1e-4, 10.0])
def test_sheared_pair_friction_cap(w):
    """Relative tangential velocity w along z: below the cap the shear is
    -K_t v_t, above it tan(phi) |F_n| along -v_t (grain i is dragged along
    with j)."""
    r, dt, delta = 0.01, 1e-6, 0.001
    x = np.array([[0.0, 0.0, 0.0], [2 * r - delta, 0.0, 0.0]])
    v = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, w]])
    o = dem_oracle(x, v, r)
    o.dem_step(dt, 0.0, YOUNG, POISSON, PHI)
    _, vn = o.particles()
    d = (x[1] + v[1] * dt) - x[0]
    dist = np.linalg.norm(d)
    n = d / dist
    pen = 2 * r - dist
    kn, kt = YOUNG * r, YOUNG * r * POISSON
    vt = v[1] - np.dot(v[1], n) * n
    fn = kn * pen
    shear = -kt * vt
    cap = fn * math.tan(PHI)
    capped = np.linalg.norm(shear) > cap
    assert capped == (w > 1)
    if capped:
        shear = shear * cap / np.linalg.norm(shear)
    f_i = -fn * n - shear
    assert np.allclose(vn[0], f_i * dt, rtol=1e-9, atol=1e-18)
    assert np.allclose(vn[1], v[1] - f_i * dt, rtol=1e-9, atol=1e-18)
    # i is dragged towards +z, j slowed
    assert vn[0, 2] > 0 and vn[1, 2] < w
    o.close()


@pytest.mark.parametrize("case", ["into", "away", "normal_only"])
def test_single_wall_particle(case):
    r, dt = 0.01, 1e-5
    h = 4 * r
    a, b = {"into": (0.5, 0.3), "away": (0.5, -0.3), "normal_only": (0.0, 0.3)}[case]
    x = np.array([[0.0, 1.5 * r, 0.0]])
    v = np.array([[a, -b, 0.0]])
    bx = np.zeros((1, 3))
    bn = np.array([[0.0, 1.0, 0.0]])
    bv = np.zeros((1, 3))
    o = dem_oracle(x, v, r, bx, bn, bv)
    o.dem_step(dt, G, YOUNG, POISSON, PHI)
    xn, vn = o.particles()
    vb = 1.0 / spline(3, 0.0, h)            # V_b = m / (m W(0))
    w = spline(3, np.linalg.norm(xn[0]), h)
    fn = vb * max(b, 0.0) * w              # along +y
    exp = np.array([a, -b, 0.0])
    exp[1] += fn
    if a != 0.0:
        kt = YOUNG * r * POISSON
        ft = kt * a                          # |F_tang|, capped by tan(phi) |F_norm|
        exp[0] -= min(ft, fn * math.tan(PHI))
    exp[1] -= G * dt
    assert np.allclose(vn[0], exp, rtol=1e-12, atol=1e-15)
    o.close()


def test_capillary_range_is_beyond_the_search():
    r = 0.01
    x = np.array([[0.0, 0.0, 0.0], [2 * r * (1 + 1e-3), 0.0, 0.0]])
    v = np.zeros((2, 3))
    o = dem_oracle(x, v, r)
    o.dem_step(1e-4, 0.0, YOUNG, POISSON, PHI)
    _, vn = o.particles()
    assert np.array_equal(vn, v)
    o.close()


def _cluster(n=6, r=0.01, seed=3):
    rng = np.random.default_rng(seed)
    ax = (np.arange(n) + 0.5) * 2 * r
    g = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    x = g + rng.uniform(-0.15 * r, 0.15 * r, g.shape)
    v = rng.normal(scale=0.05, size=g.shape)
    return x, v


def test_momentum_conserved_without_walls_and_gravity():
    r = 0.01
    x, v = _cluster(r=r)
    o = dem_oracle(x, v, r)
    o.dem_step(1e-4, 0.0, YOUNG, POISSON, PHI)
    _, vn = o.particles()
    dv = vn - v
    assert np.abs(dv).sum() > 1e-3                 # contacts acted
    assert np.abs(dv.sum(axis=0)).max() < 1e-12 * np.abs(dv).sum()
    o.close()


def test_hash_matches_brute_force():
    r = 0.01
    x, v = _cluster(r=r, seed=5)
    out = []
    for mode in (O.SEARCH_HASH, O.SEARCH_BRUTE):
        o = dem_oracle(x, v, r, search=mode)
        o.dem_step(1e-4, G, YOUNG, POISSON, PHI)
        out.append(o.particles())
        o.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.allclose(out[0][1], out[1][1], rtol=1e-13, atol=1e-15)


def test_grain_normal_points_from_grain_to_fluid():
    """One grain near the corner of a fluid block, reached by exactly one
    fluid particle: with a zero normal it gives the same velocity update as
    a wall particle whose normal is the unit vector grain -> fluid, and it
    pushes that particle away from itself."""
    sc = scenes.small_3d(6, 4, 4, jitter=0.0)
    dx = sc["dx"]
    h = 2 * dx
    corner = sc["x"][np.argmax(sc["x"].sum(axis=1))]
    grain = corner + 1.0 * dx * np.ones(3)
    d = np.linalg.norm(sc["x"] - grain, axis=1)
    assert (d < h).sum() == 1
    radial = (corner - grain) / np.linalg.norm(corner - grain)
    vgrain = 0.5 * radial              # moving towards the fluid particle
    outs = []
    for nrm in (np.zeros(3), radial):
        s2 = dict(sc)
        s2["bx"] = np.concatenate([sc["bx"], grain[None]])
        s2["bn"] = np.concatenate([sc["bn"], nrm[None]])
        s2["bv"] = np.concatenate([sc["bv"], vgrain[None]])
        o = O.Oracle(s2, rest_density_mode=_abi.FROM_INIT, kappa_mode=_abi.FROM_INIT)
        o.initialize()
        o.set_pressure(np.zeros(o.virtual()["node"].shape[0]))
        x0, v0 = o.particles()
        o.pressure_gradient()
        outs.append(o.particles()[1] - v0)
        o.close()
    assert np.array_equal(outs[0], outs[1])
    k = int(np.argmin(np.linalg.norm(sc["x"] - corner, axis=1)))
    assert np.dot(outs[0][k], corner - grain) > 0
    moved = np.nonzero(np.abs(outs[0]).sum(axis=1))[0]
    assert moved.tolist() == [k]
