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
