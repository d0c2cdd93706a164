"""The staged 3D gradient's static candidate bounds (csrc/vps_gradient.hip,
k_gradient3s; the sum of VirtualParticle.cpp:118-124 over the nodes within
H = 4 s of a particle): candidate c of row (cy, cz) is node f - 3 + c per
axis (f = floor((x - origin) / s)), rows with d(cy)^2 + d(cz)^2 >= 16 are
skipped and the others keep c in [cut, 7 - cut] (cut 0 for R2 >= 11, 1 for
R2 >= 6, else 2; R2 = 16 - d(cy)^2 - d(cz)^2, d = 3, 2, 1, 0, 0, 1, 2, 3).
Checked by brute force on the geometry: every node within H of any position
in the node cell is a candidate, and the rule keeps 408 candidates."""
import numpy as np


def _d(c):
    return 3 - c if c <= 3 else c - 4


def candidates():
    keep = set()
    for cz in range(8):
        for cy in range(8):
            r2 = 16 - _d(cy) ** 2 - _d(cz) ** 2
            if r2 <= 0:
                continue
            cut = 0 if r2 >= 11 else (1 if r2 >= 6 else 2)
            for c in range(cut, 8 - cut):
                keep.add((c, cy, cz))
    return keep


def test_candidate_count():
    assert len(candidates()) == 408


def test_every_node_within_h_is_a_candidate():
    keep = candidates()
    rng = np.random.default_rng(5)
    eps = 1e-12
    # positions in the node cell (units of s): random, the corners and faces
    phis = list(rng.uniform(0.0, 1.0, size=(3000, 3)))
    for corner in np.ndindex(2, 2, 2):
        phis.append(np.array(corner, dtype=float) * (1 - eps))
    offs = np.array(list(np.ndindex(12, 12, 12))) - 5   # node - f in [-5, 6]
    needed = set()
    for phi in phis:
        d = offs - phi                                   # node - x, units of s
        inside = offs[np.einsum("ij,ij->i", d, d) < 16.0]
        for n in inside:
            needed.add((n[0] + 3, n[1] + 3, n[2] + 3))
    assert needed <= keep, sorted(needed - keep)[:5]
    # and the bounds are tight: every kept candidate is within H of some
    # position in the cell (its nearest point to the cell lies within H)
    for c, cy, cz in keep:
        near = np.array([_d(c), _d(cy), _d(cz)], dtype=float)
        assert near @ near < 16.0
