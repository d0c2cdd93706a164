"""Readers and writers of the reference's frame files (pure numpy).

* ``<frame>/particles.mesh``, ``virtual_particles.mesh``,
  ``boundary_particles.mesh`` (DualParticleSphLiquid::writeFrame,
  Cores/Physics/DualParticleSphLiquid.cpp:41-76): little-endian uint32 count,
  count float32[Dim] positions, in 3D count float32[3] normals (0, 0, 1),
  then count float32 scalars (|v| for the fluid, V_I for the virtual
  particles, none for the boundary);
* ``<frame>/particles.sav``, ``velocities.sav`` (ParticlesAttribute::save,
  Cores/Structures/ParticlesAttribute.h:89-90): the raw float64 Dim-vectors,
  no header;
* ``<out>/description.yaml`` (Simulator.cpp:33-39 with
  DualParticleSphLiquid::writeDescription, DualParticleSphLiquid.cpp:8-39);
* ``<out>/end_frame.txt``: the frame after the last one written
  (Simulator.cpp:86-88).
"""
import os
import struct

import numpy as np

ALBEDO = [52 / 255, 108 / 255, 156 / 255, 255 / 255]


def write_mesh(path, pos, scalar=None):
    pos = np.asarray(pos, dtype=np.float32)
    n, dim = pos.shape
    with open(path, "wb") as f:
        f.write(struct.pack("<I", n))
        f.write(pos.astype("<f4").tobytes())
        if dim == 3:
            f.write(np.tile(np.array([0, 0, 1], dtype="<f4"), n).tobytes())
        if scalar is not None:
            f.write(np.asarray(scalar, dtype="<f4").reshape(n).tobytes())


def read_mesh(path, dim, scalar=True):
    """Returns (positions (n, dim), normals (n, 3) or None, scalar (n,) or
    None) and checks that the file holds exactly that."""
    data = open(path, "rb").read()
    (n,) = struct.unpack_from("<I", data, 0)
    off = 4
    pos = np.frombuffer(data, dtype="<f4", count=n * dim, offset=off).reshape(n, dim)
    off += 4 * n * dim
    nrm = None
    if dim == 3:
        nrm = np.frombuffer(data, dtype="<f4", count=n * 3, offset=off).reshape(n, 3)
        off += 12 * n
    sc = None
    if scalar:
        sc = np.frombuffer(data, dtype="<f4", count=n, offset=off)
        off += 4 * n
    if off != len(data):
        raise ValueError("%s: %d bytes, expected %d" % (path, len(data), off))
    return pos, nrm, sc


def write_sav(path, arr):
    np.ascontiguousarray(arr, dtype="<f8").tofile(path)


def read_sav(path, dim):
    a = np.fromfile(path, dtype="<f8")
    if a.size % dim:
        raise ValueError("%s: %d doubles is not a multiple of %d" % (path, a.size, dim))
    return a.reshape(-1, dim)


def description(dim):
    """The description.yaml tree of the dual-particle liquid."""
    objs = []
    for name, cmap in (("particles", True), ("virtual_particles", True),
                       ("boundary_particles", False)):
        objs.append({
            "name": name,
            "data_mode": "dynamic",
            "primitive_type": "point_list",
            "material": {"diffuse_albedo": list(ALBEDO)},
            "indexed": False,
            "color_map": {"enabled": cmap},
        })
    return {"dimension": dim, "objects": objs}


def read_end_frame(output_dir):
    with open(os.path.join(output_dir, "end_frame.txt")) as f:
        return int(f.read().split()[0])


def node_positions(scene, nodes):
    """Lattice positions origin + coord * spacing of linear node indices
    (x fastest), as Grid::dataPosition forms them (Grid.h:51-58)."""
    nodes = np.asarray(nodes, dtype=np.int64)
    dims = np.asarray(scene["lattice_dims"], dtype=np.int64)
    dim = scene["dim"]
    cc = np.stack([nodes % dims[0], (nodes // dims[0]) % dims[1],
                   nodes // dims[0] // dims[1]], axis=1)
    org = np.asarray(scene["lattice_origin"], dtype=np.float64)
    return (org[None, :dim] + cc[:, :dim].astype(np.float64) * scene["lattice_spacing"])
