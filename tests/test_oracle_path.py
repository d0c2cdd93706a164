"""Pins of the oracle's Laplacian assembly, operator choice and pressure
gradient (VirtualParticle.cpp:64-73, 110-134, 149-169; IterativeSolver.h:12)."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_py as O
from vps import scenes, _abi
from test_oracle_fields import node_positions, W, sqd


def dense(o):
    rp, col, val = o.matrix()
    n = len(rp) - 1
    return sp.csr_matrix((val, col, rp), shape=(n, n))


@pytest.mark.parametrize("dim", [2, 3])
def test_transpose_operator_and_alpha0(dim):
    sc = scenes.small_2d(12) if dim == 2 else scenes.small_3d(8, 6, 4)
    kw = dict(rest_density_mode=_abi.FROM_INIT, kappa_mode=_abi.FROM_INIT)
    oT = O.Oracle(sc, op_mode=_abi.OP_TRANSPOSE, **kw)
    oR = O.Oracle(sc, op_mode=_abi.OP_ROW, **kw)
    oT.initialize()
    oR.initialize()
    MT, MR = dense(oT), dense(oR)
    # the CG operator of the reference is the transpose of the assembled rows
    assert abs(MT - MR.T).max() == 0.0
    # A is not symmetric (row I carries V_J)
    assert abs(MR - MR.T).max() > 1e-3 * abs(MR).max()
    # at initialize (alpha0 = -1e6) the diagonal is s_I = -sum_J A_IJ and
    # alpha0 = max_I s_I
    d = MR.diagonal()
    off = np.asarray(MR.sum(1)).ravel() - d
    np.testing.assert_allclose(d, -off, rtol=1e-12)
    assert oR.scalars()["alpha0"] == d.max()
    # entries: A_IJ = 2 V_J W'_H(r) / (r + 1e-6) <= 0
    virt = oR.virtual()
    nodes = node_positions(sc)[virt["node"]]
    H = 4 * sc["lattice_spacing"]
    coo = MR.tocoo()
    offd = coo.row != coo.col
    r = np.sqrt(sqd(nodes[coo.row[offd]], nodes[coo.col[offd]]))
    kernel_d1 = np.vectorize(lambda s: O.kernel_d1(dim, sc["lattice_spacing"], s))
    expect = 2 * virt["vol"][coo.col[offd]] * kernel_d1(r) / (r + 1e-6)
    np.testing.assert_allclose(coo.data[offd], expect, rtol=1e-12)
    assert np.all(coo.data[offd] <= 0)


def test_clamped_diagonal_after_a_step():
    sc = scenes.small_3d(8, 6, 4)
    o = O.Oracle(sc, op_mode=_abi.OP_ROW, rest_density_mode=_abi.FROM_INIT,
                 kappa_mode=_abi.FROM_INIT)
    o.initialize()
    a0 = o.scalars()["alpha0"]
    o.advect(1e-3, 9.8)
    o.projection_step()
    M = dense(o)
    d = M.diagonal()
    s = -(np.asarray(M.sum(1)).ravel() - d)
    np.testing.assert_allclose(d, np.maximum(s, a0), rtol=1e-12)


def test_linear_pressure_is_in_the_interior_null_space():
    # perfect lattice (no jitter): interior virtual volumes are uniform, the
    # interior rows have zero row sum and a symmetric stencil, so A p = 0
    sc = scenes.small_3d(20, 18, 18, jitter=0.0)
    o = O.Oracle(sc, op_mode=_abi.OP_ROW, rest_density_mode=_abi.FROM_INIT,
                 kappa_mode=_abi.FROM_INIT)
    o.initialize()
    virt = o.virtual()
    nodes = node_positions(sc)[virt["node"]]
    M = dense(o)
    g = np.array([0.3, -1.1, 0.6])
    p = nodes @ g
    Ap = M @ p
    dx = sc["dx"]
    lo = sc["x"].min(0) + 6.5 * dx
    hi = sc["x"].max(0) - 6.5 * dx
    deep = np.all((nodes > lo) & (nodes < hi), axis=1)
    assert deep.sum() > 0
    scale = np.abs(M).sum(1).A.ravel()[deep] * np.abs(p).max()
    assert np.max(np.abs(Ap[deep]) / scale) < 1e-12
    # and a linear pressure is NOT in the null space near the free surface
    assert np.max(np.abs(Ap[~deep])) > 1e-3 * np.abs(Ap[deep]).max() + 1e-9


def test_constant_pressure_gradient_vanishes_inside():
    sc = scenes.small_3d(20, 18, 18, jitter=0.0)
    o = O.Oracle(sc, rest_density_mode=_abi.FROM_INIT, kappa_mode=_abi.FROM_INIT)
    o.initialize()
    o.set_pressure(np.full(len(o.virtual()["node"]), 2.5))
    o.pressure_gradient()
    x, v = o.particles()
    dx = sc["dx"]
    lo = x.min(0) + 6 * dx
    hi = x.max(0) - 6 * dx
    deep = np.all((x > lo) & (x < hi), axis=1)
    assert deep.sum() > 0
    assert np.abs(v[deep]).max() < 1e-9
    assert np.abs(v[~deep]).max() > 1e-3   # the free surface is pushed


def test_boundary_term_only_resists_motion_into_the_wall():
    sc = dict(scenes.small_2d(10))
    dim = 2
    h = 4 * sc["particle_radius"]
    o = O.Oracle(sc, rest_density_mode=_abi.FROM_INIT, kappa_mode=_abi.FROM_INIT)
    o.initialize()
    nvirt = len(o.virtual()["node"])
    # bottom-row particles move down (into the wall) with zero pressure
    x0, _ = o.particles()
    v = np.zeros_like(x0)
    bottom = x0[:, 1] < 1.5 * sc["dx"]
    v[bottom, 1] = -1.0
    o2 = O.Oracle(dict(sc, v=v), rest_density_mode=_abi.FROM_INIT, kappa_mode=_abi.FROM_INIT)
    o2.initialize()
    o2.set_pressure(np.zeros(nvirt))
    o2.pressure_gradient()
    _, v2 = o2.particles()
    dv = v2 - v
    _, vb = o2.boundary_fields()
    bx, bn = sc["bx"], sc["bn"]
    for i in np.nonzero(bottom)[0][:5]:
        d2 = sqd(x0[i][None, :], bx)
        near = d2 < h * h
        dn = (bn[near] * (v[i] - 0.0)).sum(1)
        expect = -(vb[near] * np.minimum(dn, 0.0))[:, None] * bn[near] * W(dim, h, np.sqrt(d2[near]))[:, None]
        np.testing.assert_allclose(dv[i], expect.sum(0), rtol=1e-10, atol=1e-14)
        assert dv[i, 1] > 0   # pushed back out of the wall
    assert np.all(dv[~bottom] == 0)


def test_index_carry_warm_start():
    # ParticlesBasedData::resize keeps p[0..min) and zero-fills the rest
    sc = scenes.small_2d(12)
    o = O.Oracle(sc, rest_density_mode=_abi.FROM_INIT, kappa_mode=_abi.FROM_INIT,
                 max_iterations=0)
    o.initialize()
    o.advect(2e-3, 9.8)
    o.refresh_real()
    o.generate()
    n1 = len(o.virtual()["node"])
    prev = np.arange(1, n1 + 1, dtype=float)
    o.set_pressure(prev)
    for k in range(3):
        o.advect(5e-3, 9.8)
        o.refresh_real()
        o.generate()
        st = o.project(allow_fail=True)     # zero iterations: p = x0
        p = o.virtual()["p"]
        keep = min(len(prev), len(p))
        assert np.array_equal(p[:keep], prev[:keep])
        assert np.all(p[keep:] == 0)
        prev = p


# ---- closed forms of the pressure gradient and the Laplacian --------------
# On a jitter-free lattice the interior virtual volumes are uniform, V_I ~ s^Dim
# (s the lattice spacing), so the lattice sums are quadratures of integrals with
# known values:
#   sum_J V_J p_J gradK_H(x_i - x_J)  ->  -grad p      (gradK = -grad W; the
#       identity  int (g.y) grad W dV = g  for a normalised radial W)
#   sum_J alpha_IJ (p_J - p_I)        ->  -Laplacian p (alpha_IJ = 2 V_J W'/r;
#       int r W'(r) dV = -Dim)
# A wrong sign, a missing V_J, the factor 2 or the wrong kernel radius moves
# these by O(1), far beyond the quadrature error (< 1 %).

def _lattice_scene(dim):
    return scenes.small_3d(20, 18, 18, jitter=0.0) if dim == 3 else scenes.small_2d(40, jitter=0.0)


def _deep(points, x, margin):
    lo = x.min(0) + margin
    hi = x.max(0) - margin
    return np.all((points > lo) & (points < hi), axis=1)


def linear_pressure_error(dim):
    """max |dv_i + g| / |g| over deep real particles for p = g.x."""
    sc = _lattice_scene(dim)
    o = O.Oracle(sc, rest_density_mode=_abi.FROM_INIT, kappa_mode=_abi.FROM_INIT)
    o.initialize()
    nodes = node_positions(sc)[o.virtual()["node"]]
    g = np.array([0.3, -1.1, 0.6])[:dim]
    o.set_pressure(nodes @ g)
    x, v0 = o.particles()
    o.pressure_gradient()
    _, v1 = o.particles()
    deep = _deep(x, x, 6 * sc["dx"])
    assert deep.sum() > 50
    o.close()
    return np.abs(v1[deep] - v0[deep] + g).max() / np.abs(g).max()


def single_node_pressure_error(dim):
    """p = 1 at one deep node J, 0 elsewhere: every real particle i with
    |x_i - x_J| < H moves by V_J (-W'_H(r)) (x_i - x_J)/r (outward), the
    others do not move.  Returns the max deviation relative to the largest
    push, and the fraction of particles within H that moved."""
    sc = _lattice_scene(dim)
    o = O.Oracle(sc, rest_density_mode=_abi.FROM_INIT, kappa_mode=_abi.FROM_INIT)
    o.initialize()
    virt = o.virtual()
    nodes = node_positions(sc)[virt["node"]]
    x, v0 = o.particles()
    J = int(np.argmin(((nodes - x.mean(0)) ** 2).sum(1)))
    p = np.zeros(len(nodes))
    p[J] = 1.0
    o.set_pressure(p)
    o.pressure_gradient()
    _, v1 = o.particles()
    dv = v1 - v0
    s = sc["lattice_spacing"]
    H = 4 * s
    d = x - nodes[J]
    r = np.sqrt((d * d).sum(1))
    # cubic spline derivative in closed form (SmoothedParticles.cpp:19-22,
    # SmoothedParticles.h:67-72), radius H
    c0 = 40 / (7 * np.pi * H * H) if dim == 2 else 8 / (np.pi * H ** 3)
    c1 = 6 * c0 / H
    q = r / H
    dW = np.where(q < 0.5, c1 * q * (3 * q - 2), np.where(q < 1, -c1 * (1 - q) ** 2, 0.0))
    expect = (virt["vol"][J] * -dW / np.where(r > 0, r, 1))[:, None] * d
    o.close()
    inside = r < H * (1 - 1e-9)
    moved = np.abs(dv[inside]).max(1) > 0
    return np.abs(dv - expect).max() / np.abs(expect).max(), moved.mean()


def quadratic_pressure_error(dim):
    """p = |x|^2 / 2 at the virtual nodes: interior (A p)_I = -Dim V_I / s^Dim.
    Also checks that deep rows hold 250 (3D) / 44 (2D) off-diagonal entries
    below -1e-9 of the row's largest (the |offset|^2 = 16 ties at exactly H
    carry W'(H) = 0 up to rounding).  Returns max |(A p)_I / (-Dim V_I / s^Dim) - 1|."""
    sc = _lattice_scene(dim)
    o = O.Oracle(sc, op_mode=_abi.OP_ROW, rest_density_mode=_abi.FROM_INIT,
                 kappa_mode=_abi.FROM_INIT)
    o.initialize()
    virt = o.virtual()
    nodes = node_positions(sc)[virt["node"]]
    M = dense(o)
    o.close()
    s = sc["lattice_spacing"]
    x = sc["x"]
    deep = _deep(nodes, x, 6.5 * sc["dx"])
    assert deep.sum() > 40
    rows = M[np.nonzero(deep)[0]].tocoo()
    offd = rows.col != np.nonzero(deep)[0][rows.row]
    big = -1e-9 * np.abs(rows.data[offd]).max()
    per_row = np.bincount(rows.row[offd & (rows.data < big)], minlength=deep.sum())
    if not np.all(per_row == (250 if dim == 3 else 44)):
        return np.inf
    Ap = M @ (0.5 * (nodes ** 2).sum(1))
    expect = -dim * virt["vol"][deep] / s ** dim
    return np.abs(Ap[deep] / expect - 1).max()


def projected_divergence_ratio():
    """||D|| after one projection of a uniformly diverging flow (kappa = 0),
    over ||D|| before.  The alpha0 clamp of the diagonal makes the solve a
    screened Poisson problem, so the interior divergence is only partly
    removed; it must still drop."""
    sc = dict(scenes.small_3d(20, 20, 20, jitter=0.0))
    sc["kappa"] = 0.0
    x = sc["x"]
    sc["v"] = 0.4 * (x - x.mean(0))
    kw = dict(rest_density_mode=_abi.FROM_PARAMS, kappa_mode=_abi.FROM_PARAMS)
    o = O.Oracle(sc, **kw)
    o.initialize()
    o.project()
    d0 = o.virtual()["div"]
    _, v1 = o.particles()
    o.close()
    o2 = O.Oracle(dict(sc, v=v1), max_iterations=0, **kw)
    o2.initialize()
    o2.project(allow_fail=True)
    d1 = o2.virtual()["div"]
    o2.close()
    return np.linalg.norm(d1) / np.linalg.norm(d0)


@pytest.mark.parametrize("dim", [2, 3])
def test_linear_pressure_gives_minus_its_gradient(dim):
    assert linear_pressure_error(dim) < 0.02


@pytest.mark.parametrize("dim", [2, 3])
def test_single_node_pressure_pushes_out_to_H(dim):
    err, moved = single_node_pressure_error(dim)
    assert err < 1e-9
    assert moved == 1.0


@pytest.mark.parametrize("dim", [2, 3])
def test_quadratic_pressure_laplacian_scale(dim):
    assert quadratic_pressure_error(dim) < 0.01


def test_projection_reduces_divergence():
    assert projected_divergence_ratio() < 0.6


# each planted mistake must fail at least the checks listed for it
_PLANTED = [
    (O.MUT_GRAD_SIGN, ["linear", "single", "divergence"]),
    (O.MUT_GRAD_NO_VOL, ["linear", "single"]),
    (O.MUT_GRAD_KERNEL_H, ["single"]),
    (O.MUT_COEF_HALF, ["quadratic"]),
    (O.MUT_COEF_SIGN, ["quadratic"]),
    (O.MUT_COEF_KERNEL_H, ["quadratic"]),
]


@pytest.mark.parametrize("mut,checks", _PLANTED)
def test_planted_mistakes_fail_the_closed_forms(mut, checks):
    with O.mutation(mut):
        for dim in (2, 3):
            if "linear" in checks:
                assert not linear_pressure_error(dim) < 0.02
            if "single" in checks:
                err, moved = single_node_pressure_error(dim)
                assert not (err < 1e-9 and moved == 1.0)
            if "quadratic" in checks:
                assert not quadratic_pressure_error(dim) < 0.01
        if "divergence" in checks:
            assert not projected_divergence_ratio() < 0.6
