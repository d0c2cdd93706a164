"""The synthetic scenes: sizes, wall normals, lattice geometry."""
import numpy as np

from vps import scenes


def test_config_sizes():
    assert len(scenes.config1()["x"]) == 10000
    s2 = scenes.config2()
    assert len(s2["x"]) == 150 * 90 * 75 == 1012500
    assert tuple(s2["lattice_dims"]) == (305, 155, 155)
    assert len(scenes.config2(dx=0.01)["x"]) == 300000


def test_seeded_and_jitter_bounded():
    a, b = scenes.config1(seed=5), scenes.config1(seed=5)
    assert np.array_equal(a["x"], b["x"])
    c = scenes.config1(seed=6)
    assert not np.array_equal(a["x"], c["x"])
    dx = a["dx"]
    base = (np.floor(a["x"] / dx) + 0.5) * dx
    assert np.abs(a["x"] - base).max() <= 0.1 * dx + 1e-15


def test_wall_normals_point_into_the_box():
    s = scenes.small_3d(6, 5, 4)
    n = s["bn"]
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, rtol=1e-12)
    # moving a wall particle along its normal brings it toward the box
    box_hi = np.array([12, 9, 8]) * s["dx"]
    dist = lambda p: np.linalg.norm(np.maximum(np.maximum(-p, p - box_hi), 0), axis=1)
    assert np.all(dist(s["bx"] + 0.5 * s["dx"] * n) < dist(s["bx"]))


def test_lattice_is_staggered_node_grid():
    # StaggeredGrid(bw = 2, dx, box / dx, centre): origin = c - (res + 4) dx / 2
    s = scenes.config1()
    np.testing.assert_allclose(s["lattice_origin"][:2], 0.5 - 204 * 0.005 / 2)
    assert tuple(s["lattice_dims"][:2]) == (205, 205)
