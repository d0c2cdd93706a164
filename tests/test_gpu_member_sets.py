"""The sets the FP32 neighbour walks actually sum over (needs an MI355X).

The density walks (real and boundary, SmoothedParticles.cpp:25-31), the
viscosity walk and the pressure gradient's walk over the virtual particles
(VirtualParticle.cpp:118-124) decide membership on FP32 cell-relative distances, where the
reference tests the FP64 squared distance against h^2 (H^2).  Their terms
vanish at the cutoff, so the fields cannot show a differing pair; here the
number of terms each walk enters per particle (vps_debug_member_counts) is
checked against the exact FP64 sets (vps_query_neighbors, itself bit-exact
against brute force): every pair clearly inside the cutoff must be counted,
no pair clearly outside, and only pairs within 1e-5 of the cutoff radius
may differ (so a query with no such pair has exactly the reference's set)."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from vps import Projection, _abi, scenes

pytestmark = pytest.mark.gpu

INIT = dict(rest_density_mode=_abi.FROM_INIT, kappa_mode=_abi.FROM_INIT)
BAND = 1e-5   # relative band of the radius where a pair may go either way


def node_positions(sc, lin):
    dims = sc["lattice_dims"]
    o = sc["lattice_origin"]
    s = sc["lattice_spacing"]
    cx = lin % dims[0]
    cy = (lin // dims[0]) % dims[1]
    cz = lin // dims[0] // dims[1]
    return np.stack([o[0] + cx * s, o[1] + cy * s, o[2] + cz * s], axis=1)


def check_counts(counts, exact, pts, q, radius):
    """counts, exact: terms per query; the cKDTree bands bracket both."""
    tree = cKDTree(pts)
    inner = tree.query_ball_point(q, radius * (1 - BAND), return_length=True)
    outer = tree.query_ball_point(q, radius * (1 + BAND), return_length=True)
    assert np.all(inner <= exact) and np.all(exact <= outer)   # the bands themselves
    # where no pair lies within the band (inner == outer) the count is exact;
    # elsewhere only band pairs may go either way (wall particles on a
    # lattice have many pairs exactly at h, whose term is zero)
    assert np.all(inner <= counts), np.argmin(counts - inner)
    assert np.all(counts <= outer), np.argmax(counts - outer)
    return int(np.sum(counts != exact))


@pytest.mark.parametrize("name", ["small", "dx40"])
def test_walk_member_counts(name):
    sc = scenes.small_3d(12, 8, 6) if name == "small" else scenes.config2(dx=1.0 / 40)
    g = Projection(sc, **INIT)
    g.initialize()
    for _ in range(2):
        g.advect(1e-3, 9.8)
        g.projection_step()
    x, _ = g.particles()
    h = 4 * sc["particle_radius"]
    H = 4 * sc["lattice_spacing"]

    # real density walk: real particles within h, the particle itself included
    offs, _ = g.query(0, x)
    check_counts(g.member_counts(0), np.diff(offs), x, x, h)

    # viscosity walk: the same pairs without the particle itself (the
    # scene holds no coincident particles)
    check_counts(g.member_counts(3) + 1, np.diff(offs), x, x, h)

    # boundary density walk: boundary particles within h of each boundary particle
    bx = g.boundary()[0]
    offs, _ = g.query(1, bx)
    check_counts(g.member_counts(1), np.diff(offs), bx, bx, h)

    # gradient: virtual particles within H of each real particle
    counts = g.member_counts(2)
    assert np.all(counts >= 0)   # every span on the staged path
    offs, _ = g.query(2, x)
    nodes = node_positions(sc, g.virtual()["node"])
    check_counts(counts, np.diff(offs), nodes, x, H)
    g.close()
