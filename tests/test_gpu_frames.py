"""Frames and checkpoints through the C ABI (needs an MI355X).

* the .mesh/.sav files written for a frame hold exactly the resident state
  (positions and |v| as FP32, the virtual set of the last generation with
  V_I, all boundary particles) and the virtual set matches the oracle's
  node for node;
* a run saved at a frame and resumed from it in a new context continues
  bit-identically (positions, velocities, CG iterations) for the warm
  starts that carry pressures, including the learned one while training;
* without projection.sav the load re-initialises as the reference does
  (volumes recomputed from the loaded positions)."""
import os
import struct

import numpy as np
import pytest
import yaml

import oracle_py as O
from vps import Projection, VpsError, frames, scenes, _abi
from vps.core import write_description

pytestmark = pytest.mark.gpu

INIT = dict(rest_density_mode=_abi.FROM_INIT, kappa_mode=_abi.FROM_INIT)


def test_frame_files_hold_the_state(tmp_path):
    sc = scenes.small_3d(12, 8, 6)
    g = Projection(sc, **INIT)
    g.initialize()
    o = O.Oracle(sc, **INIT)
    o.initialize()
    for _ in range(3):
        g.advect(1e-3, 9.8)
        g.projection_step()
        o.advect(1e-3, 9.8)
        o.projection_step()
    out = str(tmp_path / "run")
    write_description(3, out)
    g.write_and_save_frame(out, 3)
    fd = os.path.join(out, "3")
    assert frames.read_end_frame(out) == 4
    tree = yaml.safe_load(open(os.path.join(out, "description.yaml")))
    assert [ob["name"] for ob in tree["objects"]] == [
        ob["name"] for ob in frames.description(3)["objects"]]
    x, v = g.particles()
    pos, nrm, speed = frames.read_mesh(os.path.join(fd, "particles.mesh"), 3)
    assert np.array_equal(pos, x.astype(np.float32))
    assert np.array_equal(nrm, np.tile(np.float32([0, 0, 1]), (len(x), 1)))
    assert np.array_equal(speed, np.linalg.norm(v, axis=1).astype(np.float32))
    assert np.array_equal(frames.read_sav(os.path.join(fd, "particles.sav"), 3), x)
    assert np.array_equal(frames.read_sav(os.path.join(fd, "velocities.sav"), 3), v)
    vg = g.virtual()
    vpos, _, vvol = frames.read_mesh(os.path.join(fd, "virtual_particles.mesh"), 3)
    assert np.array_equal(vpos, frames.node_positions(sc, vg["node"]).astype(np.float32))
    assert np.array_equal(vvol, vg["vol"].astype(np.float32))
    vo = o.virtual()   # the oracle's virtual set, written by the Python writer
    ref = str(tmp_path / "virtual_oracle.mesh")
    frames.write_mesh(ref, frames.node_positions(sc, vo["node"]), vo["vol"])
    rpos, _, rvol = frames.read_mesh(ref, 3)
    assert np.array_equal(rpos, vpos)
    assert np.max(np.abs(rvol - vvol) / np.abs(rvol)) < 5e-6
    bpos, _, _ = frames.read_mesh(os.path.join(fd, "boundary_particles.mesh"), 3, scalar=False)
    assert np.array_equal(bpos, sc["bx"].astype(np.float32))
    g.close()
    o.close()


def _run(g, steps, dt=1e-3):
    out = []
    for _ in range(steps):
        g.advect(dt, 9.8)
        st = g.projection_step()
        x, v = g.particles()
        out.append((x, v, st["cg_iterations"]))
    return out


@pytest.mark.parametrize("warm", [_abi.WARM_INDEX_CARRY, _abi.WARM_NODE_KEYED,
                                  _abi.WARM_EXTRAPOLATE, _abi.WARM_LEARNED])
def test_resume_is_exact(tmp_path, warm):
    sc = scenes.small_3d(12, 8, 6)
    train = warm == _abi.WARM_LEARNED
    a = Projection(sc, warm_start=warm, **INIT)
    a.initialize()
    if train:
        a.mlp_set_training(True, batch=4096, lr=1e-3, steps_per_solve=2, seed=7)
    _run(a, 4)
    fd = str(tmp_path / "f4")
    os.makedirs(fd)
    a.save_frame(fd)
    ra = _run(a, 4)
    b = Projection(sc, warm_start=warm, **INIT)
    if train:
        b.mlp_set_training(True, batch=4096, lr=1e-3, steps_per_solve=2, seed=7)
    b.load_frame(fd)
    rb = _run(b, 4)
    for k, ((xa, va, ia), (xb, vb, ib)) in enumerate(zip(ra, rb)):
        assert ia == ib, k
        assert np.array_equal(xa, xb), k
        assert np.array_equal(va, vb), k
    if train:
        assert np.array_equal(a.mlp_parameters(), b.mlp_parameters())
    a.close()
    b.close()


def _saved_drift(fd):
    """The 3D CG drift stored in projection.sav (version 2 header: magic,
    dim, warm start, initialized, n, rho0, kappa, alpha0, drift)."""
    with open(os.path.join(fd, "projection.sav"), "rb") as f:
        head = f.read(60)
    assert head[:8] == b"VPSPROJ2"
    return struct.unpack_from("<d", head, 8 + 4 + 4 + 4 + 8 + 24)[0]


def test_resume_keeps_the_cg_drift(tmp_path):
    """The recursion's stopping threshold depends on the FP32 drift measured
    at the last solve: the checkpoint carries it, so the resumed run stops at
    the same iterations with bit-equal x and v."""
    sc = scenes.config2(dx=1.0 / 40)
    a = Projection(sc)
    a.initialize()
    _run(a, 4, dt=5e-4)
    fd = str(tmp_path / "f4")
    os.makedirs(fd)
    a.save_frame(fd)
    drift = _saved_drift(fd)
    assert drift != 1.0 and np.isfinite(drift)
    ra = _run(a, 4, dt=5e-4)
    b = Projection(sc)
    b.load_frame(fd)
    rb = _run(b, 4, dt=5e-4)
    for k, ((xa, va, ia), (xb, vb, ib)) in enumerate(zip(ra, rb)):
        assert ia == ib, k
        assert np.array_equal(xa, xb), k
        assert np.array_equal(va, vb), k
    a.close()
    b.close()


def test_load_without_checkpoint_reinitialises(tmp_path):
    sc = scenes.small_3d(12, 8, 6)
    a = Projection(sc, **INIT)
    a.initialize()
    _run(a, 3)
    out = str(tmp_path / "run")
    a.write_and_save_frame(out, 3)
    os.remove(os.path.join(out, "3", "projection.sav"))
    b = Projection(sc, **INIT)
    b.load_frame_number(out, 3)
    x, v = a.particles()
    xb, vb = b.particles()
    assert np.array_equal(x, xb) and np.array_equal(v, vb)
    # volumes recomputed from the loaded positions, as a fresh initialize
    c = Projection(dict(sc, x=x, v=v), **INIT)
    c.initialize()
    assert np.array_equal(b.real_fields()[1], c.real_fields()[1])
    with pytest.raises(VpsError):
        b.load_frame_number(out, 9)   # no such frame
    for p in (a, b, c):
        p.close()
