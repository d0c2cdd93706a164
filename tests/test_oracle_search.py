"""Pins of the oracle's hash-grid searcher (ParticlesNearbySearcher.h:37-74,
.cpp:10-35): hand-computed keys, and the neighbour sets of the hashed
walk equal the brute-force definition (x_j - q).squaredNorm() < h^2 with
strict inequality (exact ties excluded)."""
import numpy as np
import pytest

import oracle_py as O
from vps import scenes


def _key(coord, nb):
    m = 2 ** 64
    k = (73856093 * (coord[0] % m)) % m | (19349663 * (coord[1] % m)) % m
    if len(coord) == 3:
        k |= (83492791 * (coord[2] % m)) % m
    return k % nb


def test_hash_keys_by_hand():
    assert O.hash_key(3, [0, 0, 0], 10) == 0
    assert O.hash_key(3, [1, 2, 3], 10) == (73856093 | 38699326 | 250478373) % 10
    assert O.hash_key(2, [1, 2], 10) == (73856093 | 38699326) % 10
    # negative coordinates wrap as ullong(int) does
    for c in ([-1, 0, 2], [5, -3, -7], [-2, -2]):
        for nb in (2, 17, 1000, 2 * 12345):
            assert O.hash_key(len(c), c, nb) == _key(c, nb)


def _brute(x, q, r2):
    d = x - q
    if x.shape[1] == 2:
        s = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]
    else:
        s = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    return np.nonzero(s < r2)[0]


@pytest.mark.parametrize("dim", [2, 3])
def test_hash_equals_brute_force(dim):
    sc = scenes.small_2d(16) if dim == 2 else scenes.small_3d(10, 6, 5)
    o = O.Oracle(sc, search=O.SEARCH_HASH)
    o.initialize()
    x = sc["x"]
    h = 4 * sc["particle_radius"]
    rng = np.random.default_rng(3)
    q = np.concatenate([x[:200], rng.uniform(x.min(0) - h, x.max(0) + h, size=(200, dim))])
    offs, idx = o.query(0, q)
    for k in range(len(q)):
        assert np.array_equal(idx[offs[k]:offs[k + 1]], _brute(x, q[k], h * h))
    # boundary set too
    offs, idx = o.query(1, q)
    for k in range(len(q)):
        assert np.array_equal(idx[offs[k]:offs[k + 1]], _brute(sc["bx"], q[k], h * h))


def test_exact_ties_excluded():
    # dyadic lattice: squared distances are exact, ties at exactly h^2
    sc = scenes.small_2d(8)
    sc = dict(sc)
    r = 1.0 / 64            # h = 1/16
    h = 4 * r
    g = np.arange(0, 8) / 16.0
    xx, yy = np.meshgrid(g, g, indexing="ij")
    sc["x"] = np.stack([xx.ravel(), yy.ravel()], axis=1) + 0.25
    sc["v"] = np.zeros_like(sc["x"])
    sc["particle_radius"] = r
    sc["lattice_spacing"] = r * 2
    o = O.Oracle(sc, search=O.SEARCH_HASH)
    o.initialize()
    q = sc["x"][27:28]
    offs, idx = o.query(0, q, ordered=False)
    d2 = ((sc["x"] - q) ** 2).sum(1)
    assert np.any(d2 == h * h)            # the lattice has exact ties
    assert np.array_equal(np.sort(idx), np.nonzero(d2 < h * h)[0])
    # brute-force mode agrees
    ob = O.Oracle(sc, search=O.SEARCH_BRUTE)
    offs2, idx2 = ob.query(0, q)
    assert np.array_equal(np.sort(idx), idx2)
