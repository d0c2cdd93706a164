"""Full-size (1,012,500 particles) checks of the HIP path through the C ABI.

The oracle cannot run at this size in seconds, so the checks are
size-independent: FP64 numpy brute force over ALL particles for sampled
queries (neighbour sets, rho, virtual rho/V/D, the velocity update), and
the CG true relative residual."""
import math

import numpy as np
import pytest

from vps import Projection, scenes, _abi

pytestmark = pytest.mark.gpu


def W(dim, kr, d, deriv=False):
    c0 = 40.0 / (7 * math.pi * kr * kr) if dim == 2 else 8.0 / (math.pi * kr ** 3)
    x = d / kr
    if not deriv:
        return np.where(x < 0.5, c0 * (6 * x * x * (x - 1) + 1),
                        np.where(x < 1, 2 * c0 * (1 - x) ** 3, 0.0))
    c1 = 6 * c0 / kr
    return np.where(x < 0.5, c1 * x * (3 * x - 2), np.where(x < 1, -c1 * (1 - x) ** 2, 0.0))


def sqd(a, b):
    d = a - b
    return (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]


@pytest.fixture(scope="module")
def run():
    sc = scenes.config2()
    g = Projection(sc, rest_density_mode=_abi.FROM_PARAMS, kappa_mode=_abi.FROM_PARAMS)
    g.initialize()
    for _ in range(2):
        g.advect(5e-4, 9.8)
        g.projection_step()
    g.advect(5e-4, 9.8)
    g.refresh_real()
    g.generate_particles()
    x, v = g.particles()
    rho, vol = g.real_fields()
    brho, bvol = g.boundary_fields()
    virt = g.virtual()
    st = g.project()
    x2, v2 = g.particles()
    virt2 = g.virtual()
    out = dict(sc=sc, x=x, v=v, rho=rho, vol=vol, brho=brho, bvol=bvol,
               virt=virt, st=st, v2=v2, p=virt2["p"], g=g,
               scal=g.scalars())
    yield out
    g.close()


def node_pos(sc, lin):
    dims, o, s = sc["lattice_dims"], sc["lattice_origin"], sc["lattice_spacing"]
    cx = lin % dims[0]
    cy = (lin // dims[0]) % dims[1]
    cz = lin // dims[0] // dims[1]
    return np.stack([o[0] + cx * s, o[1] + cy * s, o[2] + cz * s], axis=1)


def test_cg_reaches_true_residual(run):
    st = run["st"]
    assert st["cg_iterations"] > 0
    assert st["cg_true_rel_residual"] < 1e-6
    assert st["n_virtual"] > 1_000_000


def test_sampled_neighbour_sets_and_density(run):
    sc, x, g = run["sc"], run["x"], run["g"]
    h = 4 * sc["particle_radius"]
    rng = np.random.default_rng(11)
    idx = rng.choice(len(x), 48, replace=False)
    offs, nb = g.query(0, x[idx])
    for k, i in enumerate(idx):
        d2 = sqd(x, x[i])
        ref = np.nonzero(d2 < h * h)[0]
        assert np.array_equal(nb[offs[k]:offs[k + 1]], ref)
        # the density of the step is that of the refreshed positions
        r = run["sc"]["mass"] * W(3, h, np.sqrt(d2[ref])).sum()
        assert run["rho"][i] == pytest.approx(r, rel=1e-5)


def test_sampled_virtual_fields(run):
    sc, x, v, virt = run["sc"], run["x"], run["v"], run["virt"]
    h = 4 * sc["particle_radius"]
    H = 4 * sc["lattice_spacing"]
    m = sc["mass"]
    rho, vol = run["rho"], run["vol"]
    bx, bv, brho, bvol = sc["bx"], sc["bv"], run["brho"], run["bvol"]
    scal = run["scal"]
    rng = np.random.default_rng(12)
    pick = rng.choice(len(virt["node"]), 64, replace=False)
    q = node_pos(sc, virt["node"][pick])
    dscale = np.abs(virt["div"]).max()
    for k, I in enumerate(pick):
        d2 = sqd(x, q[k])
        j = np.nonzero(d2 < h * h)[0]
        r = np.sqrt(d2[j])
        w = W(3, h, r)
        assert virt["rho"][I] == pytest.approx(m * W(3, H, r).sum(), rel=1e-5)
        assert virt["vol"][I] == pytest.approx((vol[j] ** 2 * w).sum(), rel=1e-5)
        d2b = sqd(bx, q[k])
        jb = np.nonzero(d2b < h * h)[0]
        rb = np.sqrt(d2b[jb])
        wb = W(3, h, rb)
        num = m * (v[j] * (w / rho[j])[:, None]).sum(0) + m * (bv[jb] * (wb / brho[jb])[:, None]).sum(0)
        den = (vol[j] * w).sum() + (bvol[jb] * wb).sum() + 1e-6
        vbar = num / den

        def grad(dpos, rr, wd):
            with np.errstate(invalid="ignore", divide="ignore"):
                s = np.where(rr > 0, -wd / rr, 0.0)
            return dpos * s[:, None]
        g_r = grad(q[k] - x[j], r, W(3, h, r, True))
        g_b = grad(q[k] - bx[jb], rb, W(3, h, rb, True))
        D = ((v[j] * g_r).sum(1) * vol[j]).sum() - vbar @ (vol[j][:, None] * g_r).sum(0) \
            + ((bv[jb] * g_b).sum(1) * bvol[jb]).sum() - vbar @ (bvol[jb][:, None] * g_b).sum(0) \
            - scal["kappa"] * max(virt["rho"][I] - scal["rest_density"], 0) / scal["rest_density"]
        assert abs(virt["div"][I] - D) < 1e-4 * dscale


def test_sampled_velocity_update(run):
    sc, x, v, v2, virt, p = run["sc"], run["x"], run["v"], run["v2"], run["virt"], run["p"]
    h = 4 * sc["particle_radius"]
    H = 4 * sc["lattice_spacing"]
    nodes = node_pos(sc, virt["node"])
    bx, bn, bv, bvol = sc["bx"], sc["bn"], sc["bv"], run["bvol"]
    rng = np.random.default_rng(13)
    pick = rng.choice(len(x), 48, replace=False)
    dv_all = np.linalg.norm(v2 - v, axis=1)
    scale = np.percentile(dv_all, 99)
    for i in pick:
        d2 = sqd(nodes, x[i])
        J = np.nonzero(d2 < H * H)[0]
        r = np.sqrt(d2[J])
        with np.errstate(invalid="ignore", divide="ignore"):
            s = np.where(r > 0, -W(3, H, r, True) / r, 0.0)
        f = ((virt["vol"][J] * p[J] * s)[:, None] * (x[i] - nodes[J])).sum(0)
        d2b = sqd(bx, x[i])
        jb = np.nonzero(d2b < h * h)[0]
        dn = (bn[jb] * (v[i] - bv[jb])).sum(1)
        f -= ((bvol[jb] * np.minimum(dn, 0) * W(3, h, np.sqrt(d2b[jb])))[:, None] * bn[jb]).sum(0)
        assert np.linalg.norm((v2[i] - v[i]) - f) < 1e-4 * max(np.linalg.norm(f), scale)
