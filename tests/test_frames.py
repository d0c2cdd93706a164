"""Frame file formats on the CPU: the byte layout of .mesh and .sav files as
the reference writes them (pinned by hand-packed bytes), the description
tree, and the C ABI's description/end_frame writers (host-only code of the
library, no GPU call)."""
import ctypes as C
import os
import struct

import numpy as np
import pytest
import yaml

from vps import frames, _abi


def test_mesh_bytes_3d(tmp_path):
    pos = np.array([[1.0, 2.0, 3.0], [-0.5, 0.25, 8.0]])
    speed = np.array([0.75, 1.5])
    p = str(tmp_path / "particles.mesh")
    frames.write_mesh(p, pos, speed)
    expect = (struct.pack("<I", 2) + struct.pack("<6f", 1, 2, 3, -0.5, 0.25, 8) +
              struct.pack("<6f", 0, 0, 1, 0, 0, 1) + struct.pack("<2f", 0.75, 1.5))
    assert open(p, "rb").read() == expect
    rp, rn, rs = frames.read_mesh(p, 3)
    assert np.array_equal(rp, pos.astype(np.float32))
    assert np.array_equal(rn, np.tile([0, 0, 1], (2, 1)))
    assert np.array_equal(rs, speed.astype(np.float32))


def test_mesh_bytes_2d_without_scalar(tmp_path):
    pos = np.array([[0.1, 0.2], [0.3, 0.4], [0.5, 0.6]])
    p = str(tmp_path / "boundary_particles.mesh")
    frames.write_mesh(p, pos)
    expect = struct.pack("<I", 3) + np.asarray(pos, dtype="<f4").tobytes()
    assert open(p, "rb").read() == expect
    rp, rn, rs = frames.read_mesh(p, 2, scalar=False)
    assert rn is None and rs is None and rp.shape == (3, 2)
    with pytest.raises(ValueError):
        frames.read_mesh(p, 2, scalar=True)   # size check


def test_sav_is_raw_fp64(tmp_path):
    x = np.random.default_rng(1).normal(size=(5, 3))
    p = str(tmp_path / "particles.sav")
    frames.write_sav(p, x)
    assert os.path.getsize(p) == 5 * 3 * 8
    assert open(p, "rb").read() == struct.pack("<15d", *x.reshape(-1))
    assert np.array_equal(frames.read_sav(p, 3), x)


def test_description_tree():
    d = frames.description(3)
    assert d["dimension"] == 3
    assert [o["name"] for o in d["objects"]] == [
        "particles", "virtual_particles", "boundary_particles"]
    assert [o["color_map"]["enabled"] for o in d["objects"]] == [True, True, False]
    for o in d["objects"]:
        assert o["data_mode"] == "dynamic" and o["primitive_type"] == "point_list"
        assert o["indexed"] is False
        assert np.allclose(o["material"]["diffuse_albedo"],
                           np.array([52, 108, 156, 255]) / 255)


def _lib():
    from vps import core
    return core.load_library()


@pytest.mark.parametrize("dim", [2, 3])
def test_library_description_and_end_frame(tmp_path, dim):
    L = _lib()
    out = str(tmp_path / "out")
    assert L.vps_write_description(dim, out.encode()) == 0
    tree = yaml.safe_load(open(os.path.join(out, "description.yaml")))
    ref = frames.description(dim)
    assert tree["dimension"] == dim
    assert len(tree["objects"]) == 3
    for a, b in zip(tree["objects"], ref["objects"]):
        assert a["name"] == b["name"] and a["color_map"] == b["color_map"]
        assert a["indexed"] is False and a["primitive_type"] == "point_list"
        assert np.allclose(a["material"]["diffuse_albedo"], b["material"]["diffuse_albedo"],
                           rtol=1e-7)
    assert L.vps_write_end_frame(out.encode(), 7) == 0
    assert frames.read_end_frame(out) == 8
    assert L.vps_write_description(4, out.encode()) == _abi.VPS_ERR_INVALID_ARG
    assert L.vps_write_end_frame(str(tmp_path / "missing" / "x").encode(), 1) == _abi.VPS_ERR_IO
