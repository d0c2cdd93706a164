"""Pins of the sampling oracle (BoundaryParticles::addSurface and
Shapes::generateBox restated) against geometry it must satisfy:

* 2D: the hexagonal rows hold the closed-form point counts, neighbours in a
  row are one spacing (2r) apart, rows sqrt(3) r apart, the block is
  centred on `center`;
* 3D: face + node centres of a staggered grid are an FCC packing: the
  closed-form count, every point's nearest neighbour at 2r;
* surfaces: the box-complement normals equal the dam-break scenes' own
  implementation (written independently), inside tests equal brute-force
  formulas for sphere, plane and ellipsoid, sphere normals are radial;
* the reference's shell lies outside the box [-1, 1]^Dim."""
import math

import numpy as np
import pytest

import sampling_oracle as S
from vps import scenes


def test_2d_rows_closed_form():
    r = 0.05
    x = S.generate_box(2, r, [0.25, -0.5], [1.0, 0.6])
    sp = 2 * r
    xs = int(math.floor(2.0 / sp + 0.5))
    ys = int(math.floor((1.2 / sp - 1) * 2 / math.sqrt(3) + 1 + 0.5))
    assert len(x) == sum(xs - (j & 1) for j in range(ys))
    rows = np.unique(np.round(x[:, 1], 9))
    assert len(rows) == ys
    assert np.allclose(np.diff(rows), math.sqrt(3) * r, rtol=1e-12)
    for y in rows[:3]:
        xr = np.sort(x[np.isclose(x[:, 1], y), 0])
        assert np.allclose(np.diff(xr), sp, rtol=1e-12)
    assert np.allclose(x.mean(axis=0), [0.25, -0.5], atol=sp)


def test_3d_fcc_closed_form():
    r = 0.05
    half = np.array([0.4, 0.3, 0.35])
    x = S.generate_box(3, r, [0.0, 0.1, -0.2], half)
    sp = 2 * math.sqrt(2) * r
    res = [int(math.floor((h - r) * 2 / sp + 0.5)) for h in half]
    faces = sum(np.prod([res[c] + (c == a) for c in range(3)]) for a in range(3))
    assert len(x) == faces + np.prod([q + 1 for q in res])
    d = np.linalg.norm(x[:, None, :] - x[None, :, :], axis=2)
    np.fill_diagonal(d, np.inf)
    assert np.allclose(d.min(axis=1), 2 * r, rtol=1e-12)
    assert np.allclose(x.mean(axis=0), [0.0, 0.1, -0.2], atol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_box_complement_normals_match_the_scenes(dim):
    r = 0.04
    box_lo, box_len = -np.ones(3), 2 * np.ones(3)
    s = S.Surface(S.BOX, box_lo, box_len, complement=True, dim=dim)
    x, n = S.add_surface(dim, r, s, np.zeros(dim), np.full(dim, 1.4))
    assert len(x) > 0
    assert np.all(np.any(np.abs(x) >= 1.0, axis=1))          # outside or on the box
    ref = scenes.box_complement_normals(x, -np.ones(dim), np.ones(dim))
    assert np.allclose(n, ref, atol=1e-15)
    lat = S.lattice(dim, r, np.zeros(dim), np.full(dim, 1.4))
    outside = np.any(np.abs(lat) >= 1.0, axis=1)
    assert len(x) == outside.sum()


@pytest.mark.parametrize("dim", [2, 3])
def test_sphere_plane_ellipsoid_inside_sets(dim):
    r = 0.05
    lat = S.lattice(dim, r, np.zeros(dim), np.full(dim, 0.8))
    c = np.array([0.1, -0.05, 0.02])[:dim]
    s = S.Surface(S.SPHERE, c, [0.5, 0, 0], dim=dim)
    x, n = S.add_surface(dim, r, s, np.zeros(dim), np.full(dim, 0.8))
    dist = np.sqrt(((lat - c) ** 2).sum(axis=1))
    assert len(x) == (dist - 0.5 <= 0).sum()
    rad = (x - c) / np.linalg.norm(x - c, axis=1, keepdims=True)
    assert np.allclose(n, rad, atol=1e-15)
    xc, nc = S.add_surface(dim, r, S.Surface(S.SPHERE, c, [0.5, 0, 0], True, dim),
                           np.zeros(dim), np.full(dim, 0.8))
    assert len(xc) + len(x) == len(lat) + np.isclose(dist, 0.5, rtol=0, atol=0).sum()
    assert np.allclose(nc, -(xc - c) / np.linalg.norm(xc - c, axis=1, keepdims=True))
    pos, dirn = np.array([0, 0.2, 0])[:dim], np.array([0.3, 1.0, -0.2])[:dim]
    xp, npl = S.add_surface(dim, r, S.Surface(S.PLANE, pos, dirn, dim=dim), np.zeros(dim),
                            np.full(dim, 0.8))
    u = dirn / np.linalg.norm(dirn)
    assert len(xp) == (((lat - pos) * u).sum(axis=1) <= 1e-15).sum()
    assert np.allclose(npl, u)
    semi = np.array([0.6, 0.4, 0.5])[:dim]
    xe, _ = S.add_surface(dim, r, S.Surface(S.ELLIPSOID, 0 * semi, semi, dim=dim),
                          np.zeros(dim), np.full(dim, 0.8))
    assert len(xe) == (((lat / semi) ** 2).sum(axis=1) <= 1).sum()


@pytest.mark.parametrize("dim", [2, 3])
def test_reference_shell(dim):
    r = 0.1
    x, n = S.reference_shell(dim, r)
    assert np.all(np.any(np.abs(x) >= 1.0, axis=1))
    assert np.all(np.abs(x) <= 2.4 + 1e-12)
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0)
    # normals point into the box (towards the fluid)
    assert np.all((n * x).sum(axis=1) < 0)
