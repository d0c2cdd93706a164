"""GPU wall and block sampling against the sampling oracle, bit-exact
(positions and normals), and a projection on the reference's own scene
(its hard-coded wall shell and a packed block) against the FP64 oracle
(needs an MI355X)."""
import numpy as np
import pytest

import oracle_py as O
import sampling_oracle as S
from vps import Projection, VpsError, scenes, _abi

pytestmark = pytest.mark.gpu

INIT = dict(rest_density_mode=_abi.FROM_INIT, kappa_mode=_abi.FROM_INIT)


def empty_scene(dim, r, lo=-2.6, hi=2.6):
    x = np.zeros((0, dim))
    sc = scenes.grain_scene(np.zeros((1, dim)), np.zeros((1, dim)), r, x, x, x,
                            np.full(3, lo), np.full(3, hi))
    return sc


SURFACES = [
    (S.BOX, [-1, -1, -1], [2, 2, 2], True),
    (S.BOX, [-0.3, -0.2, -0.1], [0.9, 0.5, 0.7], False),
    (S.SPHERE, [0.1, -0.05, 0.02], [0.55, 0, 0], False),
    (S.SPHERE, [0.1, -0.05, 0.02], [0.55, 0, 0], True),
    (S.PLANE, [0, 0.2, 0], [0.3, 1.0, -0.2], False),
    (S.ELLIPSOID, [0, 0, 0], [0.6, 0.4, 0.5], False),
    (S.ELLIPSOID, [0, 0, 0], [0.6, 0.4, 0.5], True),
]


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("si", range(len(SURFACES)))
def test_add_surface_bit_exact(dim, si):
    kind, a, b, comp = SURFACES[si]
    r = 0.03
    center, half = np.array([0.05, -0.1, 0.0])[:dim], np.array([1.3, 1.2, 1.25])[:dim]
    g = Projection(empty_scene(dim, r))
    g.add_surface(kind, a, b, center, half, complement=comp)
    x, n, v = g.boundary()
    xo, no = S.add_surface(dim, r, S.Surface(kind, a, b, comp, dim), center, half)
    assert len(xo) > 0
    assert np.array_equal(x, xo)
    assert np.array_equal(n, no)
    assert np.all(v == 0)
    # appending keeps the earlier walls first
    g.add_surface(S.SPHERE, [0, 0, 0], [0.2, 0, 0], np.zeros(dim), np.full(dim, 0.3))
    x2, n2, _ = g.boundary()
    xs, ns = S.add_surface(dim, r, S.Surface(S.SPHERE, [0, 0, 0], [0.2, 0, 0], False, dim),
                           np.zeros(dim), np.full(dim, 0.3))
    assert np.array_equal(x2, np.concatenate([xo, xs]))
    assert np.array_equal(n2, np.concatenate([no, ns]))
    g.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_generate_box_bit_exact(dim):
    r = 0.02
    g = Projection(empty_scene(dim, r))
    c, h = np.array([-0.6, -0.55, -0.5])[:dim], np.array([0.4, 0.33, 0.25])[:dim]
    assert np.array_equal(g.generate_box(c, h), S.generate_box(dim, r, c, h))
    with pytest.raises(VpsError):
        g.generate_box(c, -h)
    bad = _abi.surface(9, [0, 0, 0], [1, 1, 1])
    import ctypes as C
    cc = np.ascontiguousarray(c)
    hh = np.ascontiguousarray(h)
    assert g.L.vps_add_surface(g.h, C.byref(bad), cc.ctypes.data_as(C.POINTER(C.c_double)),
                               hh.ctypes.data_as(C.POINTER(C.c_double))) == \
        _abi.VPS_ERR_INVALID_ARG
    g.close()


def reference_scene(dim, r):
    """DualParticleSphLiquid::initialize's walls with a packed block of fluid
    in the lower corner of the box [-1, 1]^Dim; lattice of spacing 2r over
    the box (StaggeredGrid node grid, boundary width 2)."""
    walls, nrm = S.reference_shell(dim, r)
    fluid = S.generate_box(dim, r, np.full(dim, -0.6), np.full(dim, 0.38))
    dx = 2 * r
    origin, dims = scenes.staggered_node_grid(-np.ones(dim), np.ones(dim), dx, dim, 2)
    lo, hi = np.zeros(3), np.zeros(3)
    lo[:dim], hi[:dim] = -2.5, 2.5
    rho0 = 1000.0
    mass = rho0 * dx ** dim
    return {
        "name": "reference_shell_%dd" % dim, "dim": dim, "dx": dx,
        "x": fluid, "v": np.zeros_like(fluid), "bx": walls, "bn": nrm,
        "bv": np.zeros_like(walls), "particle_radius": r, "lattice_spacing": dx,
        "lattice_origin": origin, "lattice_dims": dims, "mass": mass,
        "rest_density": rho0, "kappa": mass / rho0, "domain_lo": lo, "domain_hi": hi,
    }


@pytest.mark.parametrize("dim", [2, 3])
def test_projection_on_the_reference_scene(dim):
    r = 0.025 if dim == 2 else 0.05
    sc = reference_scene(dim, r)
    g = Projection(dict(sc, bx=np.zeros((0, dim)), bn=np.zeros((0, dim)),
                        bv=np.zeros((0, dim))), **INIT)
    g.add_surface(S.BOX, [-1, -1, -1], [2, 2, 2], np.zeros(dim), np.full(dim, 2 * 1.2),
                  complement=True)
    xb, nb, _ = g.boundary()
    assert np.array_equal(xb, sc["bx"]) and np.array_equal(nb, sc["bn"])
    assert np.array_equal(g.generate_box(np.full(dim, -0.6), np.full(dim, 0.38)), sc["x"])
    g.initialize()
    o = O.Oracle(sc, **INIT)
    o.initialize()
    for step in range(3):
        g.advect(1e-3, 9.8)
        o.advect(1e-3, 9.8)
        sg = g.projection_step()
        so = o.projection_step()
        assert abs(sg["cg_iterations"] - so["cg_iterations"]) <= 1, step
        assert sg["cg_true_rel_residual"] < 1e-6, step
        _, vg = g.particles()
        _, vo = o.particles()
        assert np.linalg.norm(vg - vo) / np.linalg.norm(vo) < 1e-4, step
    g.close()
    o.close()
