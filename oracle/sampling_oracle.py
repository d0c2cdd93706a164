"""Oracle (test infrastructure only) of the reference's particle sampling:
BoundaryParticles::addSurface (Cores/Structures/SmoothedParticles.cpp:
111-155), Shapes::generateBox without Poisson sampling (Cores/Utilities/
Shapes.h:32-58) and the implicit surfaces they test against
(Cores/Geometries/ImplicitSurface.h:27-113, Surface.h:25 and 50-69).

FP64 numpy, elementwise, in the reference's order of operations (Eigen
evaluates these expressions per coefficient, left to right; the squared
norm of a 2- or 3-vector is ((x^2 + y^2) + z^2)), so positions and normals
are bit-exact restatements.  Lattices:

* 2D: rows of a hexagonal lattice of spacing 2r (odd rows shifted by half a
  spacing and one point shorter);
* 3D: the face centres of the three face grids, then the node centres, of
  StaggeredGrid(0, 2 sqrt(2) r, round((half - r) 2 / spacing), center)
  (StaggeredGrid.cpp:20-34), each grid x-fastest (Grid.cpp:231-243): an FCC
  packing with nearest distance 2r.
"""
import math

import numpy as np

SQRT2 = 1.4142135623730951        # std::numbers::sqrt2
SQRT3 = 1.7320508075688772        # std::numbers::sqrt3
INV_SQRT3 = 0.5773502691896257    # std::numbers::inv_sqrt3

BOX, SPHERE, PLANE, ELLIPSOID = 0, 1, 2, 3


def std_round(x):
    """std::round: halfway cases away from zero."""
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def _grid_positions(origin, spacing, size):
    """Grid::dataPosition origin + coord * spacing over Grid::forEach."""
    dim = len(size)
    axes = [origin[a] + np.arange(size[a], dtype=np.float64) * spacing for a in range(dim)]
    mesh = np.meshgrid(*axes, indexing="ij")
    return np.stack([m.transpose().reshape(-1) for m in mesh], axis=1)


def lattice(dim, radius, center, half):
    """Candidate points of addSurface / generateBox in the reference's order."""
    center = [float(c) for c in center[:dim]]
    half = [float(h) for h in half[:dim]]
    if dim == 2:
        spacing = radius * 2
        xs = int(std_round(half[0] * 2 / spacing))
        ys = int(std_round((half[1] * 2 / spacing - 1) * 2 * INV_SQRT3 + 1))
        ox = center[0] - (float(xs - 1) * spacing) / 2
        oy = center[1] - (((ys - 1.0) * SQRT3 / 2) * spacing) / 2
        pts = []
        for j in range(ys):
            yox = ox + (0.5 if j & 1 else 0.0) * spacing
            yoy = oy + (j * SQRT3 / 2) * spacing
            n = xs - (j & 1)
            if n <= 0:
                continue
            xi = yox + np.arange(n, dtype=np.float64) * spacing
            pts.append(np.stack([xi, np.full(n, yoy + 0.0)], axis=1))
        return np.concatenate(pts) if pts else np.zeros((0, 2))
    spacing = 2 * SQRT2 * radius
    res = [int(std_round((half[a] - radius) * 2 / spacing)) for a in range(3)]
    origin = [center[a] - (float(res[a]) * spacing) / 2 for a in range(3)]
    parts = []
    for axis in range(3):
        o = [origin[c] + ((0.0 if c == axis else 1.0) * spacing) / 2 for c in range(3)]
        size = [res[c] + (1 if c == axis else 0) for c in range(3)]
        parts.append(_grid_positions(o, spacing, size))
    parts.append(_grid_positions(origin, spacing, [r + 1 for r in res]))
    return np.concatenate(parts)


def _sqnorm(v):
    s = v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]
    if v.shape[1] == 3:
        s = s + v[:, 2] * v[:, 2]
    return s


def _normalized(v):
    z = _sqnorm(v)
    out = v.copy()
    nz = z > 0
    out[nz] = v[nz] / np.sqrt(z[nz])[:, None]
    return out


class Surface:
    """kind: BOX (a = min corner, b = lengths), SPHERE (a = center,
    b[0] = radius), PLANE (a = position, b = direction), ELLIPSOID (a unused,
    b = semi-axes); complement: ComplementarySurface of it."""

    def __init__(self, kind, a, b, complement=False, dim=3):
        self.kind = kind
        self.dim = dim
        self.a = np.asarray(a, dtype=np.float64)[:dim]
        self.b = np.asarray(b, dtype=np.float64)[:dim]
        self.complement = complement

    def _inner(self, p):
        """(signed distance, closest normal) of the surface itself."""
        dim = self.dim
        if self.kind == BOX:
            c = self.a + self.b / 2
            h = self.b / 2
            d = p - c
            phi = np.abs(d) - h
            inside = np.all(phi <= 0, axis=1)
            sd = np.where(inside, phi.max(axis=1), np.sqrt(_sqnorm(np.maximum(phi, 0))))
            n = np.maximum(phi, 0.0)
            ax = np.argmax(phi, axis=1)   # first maximum, as Eigen's maxCoeff
            unit = np.zeros_like(p)
            unit[np.arange(len(p)), ax] = 1.0
            n = np.where(inside[:, None], unit, n)
            n = _normalized(n * np.sign(d))
            return sd, n
        if self.kind == SPHERE:
            d = p - self.a
            return np.sqrt(_sqnorm(d)) - self.b[0], _normalized(d)
        if self.kind == PLANE:
            nrm = _normalized(self.b[None, :])[0]
            d = p - self.a
            sd = d[:, 0] * nrm[0] + d[:, 1] * nrm[1]
            if dim == 3:
                sd = sd + d[:, 2] * nrm[2]
            return sd, np.tile(nrm, (len(p), 1))
        if self.kind == ELLIPSOID:
            q = p / self.b
            s2 = _sqnorm(q)
            inside = (s2 - 1) <= 0
            closest = p / np.sqrt(s2)[:, None]
            dv = p - closest
            sign = np.where(inside, -1.0, 1.0)
            return np.sqrt(_sqnorm(dv)) * sign, _normalized(dv) * sign[:, None]
        raise ValueError(self.kind)

    def evaluate(self, p):
        """(is_inside, closest normal) as Surface::isInside / closestNormal."""
        if self.kind == ELLIPSOID and not self.complement:
            q = p / self.b
            inside = (_sqnorm(q) - 1) <= 0
            return inside, self._inner(p)[1]
        sd, n = self._inner(p)
        if self.complement:
            sd, n = -sd, -n
        return sd <= 0, n


def add_surface(dim, radius, surface, center, half):
    """BoundaryParticles::addSurface: (positions, normals) of the lattice
    points inside the surface, in lattice order."""
    p = lattice(dim, radius, center, half)
    inside, n = surface.evaluate(p)
    return p[inside], n[inside]


def generate_box(dim, radius, center, half):
    """Shapes::generateBox (no Poisson sampling): every lattice point."""
    return lattice(dim, radius, center, half)


def reference_shell(dim, radius):
    """DualParticleSphLiquid::initialize's walls (DualParticleSphLiquid.cpp:
    79-82): the complement of the box [-1, 1]^Dim sampled in 2.4 (= 2 * 1.2)
    half-lengths around the origin."""
    s = Surface(BOX, -np.ones(3), 2 * np.ones(3), complement=True, dim=dim)
    return add_surface(dim, radius, s, np.zeros(dim), np.full(dim, 2 * 1.2))
