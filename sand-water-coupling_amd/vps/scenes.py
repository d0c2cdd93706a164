"""Seeded synthetic scenes for the projection path (dam breaks in 2D and 3D).

The reference builds its scenes in C++ builders with non-deterministic
samplers (Cores/Utilities/PoissonDiskSampler.h:28-32) and a hard-coded
boundary shell (Cores/Physics/DualParticleSphLiquid.cpp:79-82).  Here every
scene is a pure function of its parameters and a seed:

* fluid particles on a lattice of spacing dx, cell-centred ((i + 0.5) dx),
  each coordinate jittered by a uniform value in [-0.1 dx, 0.1 dx];
* solid walls: the same cell-centred lattice, ``wall_layers`` layers thick,
  outside the box, zero velocity;
* wall normals: the normal of the complement of the box, as
  ComplementarySurface<ImplicitBox>::closestNormal computes it
  (Cores/Geometries/Surface.h:65, ImplicitSurface.h:56-66), pointing into the
  fluid;
* the virtual-particle lattice is the node grid of
  StaggeredGrid(boundaryWidth, spacing = dx, resolution = box / dx,
  center = box centre) (Cores/Structures/StaggeredGrid.cpp:20-34), with the
  boundary width 2 the reference's SPH builder uses
  (Demos/SmthPartHydrodLiquidTest/SmthPartHydrodLiquidBuilder.h:32-33);
* particle radius r = dx / 2, so h = 4 r = 2 dx and H = 4 dx;
* mass m = rho0 dx^dim, kappa = m / rho0, rho0 = 1000.
"""
import numpy as np

DEFAULT_SEED = 20261202


def staggered_node_grid(box_lo, box_hi, dx, dim, boundary_width=2):
    """Node grid of StaggeredGrid<Dim>(bw, dx, resolution, center)."""
    box_lo = np.asarray(box_lo, dtype=np.float64)
    box_hi = np.asarray(box_hi, dtype=np.float64)
    res = np.rint((box_hi - box_lo) / dx).astype(np.int64)
    center = (box_lo + box_hi) / 2
    full = res + 2 * boundary_width
    origin = np.zeros(3)
    dims = np.ones(3, dtype=np.int64)
    for a in range(dim):
        # _origin = center - _resolution.cast<real>() * _spacing / 2
        origin[a] = center[a] - float(full[a]) * dx / 2
        dims[a] = full[a] + 1
    return origin, dims


def box_complement_normals(x, box_lo, box_hi):
    """-ImplicitBox::closestNormal for points outside the box."""
    dim = x.shape[1]
    c = (np.asarray(box_lo[:dim]) + np.asarray(box_hi[:dim])) / 2
    half = (np.asarray(box_hi[:dim]) - np.asarray(box_lo[:dim])) / 2
    rel = x - c
    phi = np.abs(rel) - half
    inside = np.all(phi <= 0, axis=1)
    n = np.where(phi > 0, phi, 0.0)
    if np.any(inside):
        ax = np.argmax(phi[inside], axis=1)
        e = np.zeros((inside.sum(), dim))
        e[np.arange(inside.sum()), ax] = 1
        n[inside] = e
    n = n * np.sign(rel)
    norm = np.linalg.norm(n, axis=1, keepdims=True)
    norm[norm == 0] = 1
    return -(n / norm)


def dam_break(dim, dx, box, block, seed=DEFAULT_SEED, wall_layers=3,
              jitter=0.1, rho0=1000.0, boundary_width=2, name=None):
    """Fluid block [0, block] in the corner of the box [0, box]."""
    box = np.asarray(box, dtype=np.float64)[:dim]
    block = np.asarray(block, dtype=np.float64)[:dim]
    rng = np.random.Generator(np.random.PCG64(seed))
    nb = np.rint(block / dx).astype(np.int64)
    axes = [(np.arange(nb[a]) + 0.5) * dx for a in range(dim)]
    grid = np.meshgrid(*axes, indexing="ij")
    # order x fastest for locality (the reference adds particles in its
    # builders' order; any fixed order is fine)
    x = np.stack([g.transpose().reshape(-1) for g in grid], axis=1)
    x = x + rng.uniform(-jitter * dx, jitter * dx, size=x.shape)
    v = np.zeros_like(x)
    # walls: cell-centred lattice in [-L dx, box + L dx] minus the box
    nbox = np.rint(box / dx).astype(np.int64)
    waxes = [(np.arange(-wall_layers, nbox[a] + wall_layers) + 0.5) * dx
             for a in range(dim)]
    wgrid = np.meshgrid(*waxes, indexing="ij")
    w = np.stack([g.transpose().reshape(-1) for g in wgrid], axis=1)
    outside = np.any((w < 0) | (w > box), axis=1)
    w = w[outside]
    wn = box_complement_normals(w, np.zeros(dim), box)
    wv = np.zeros_like(w)
    origin, dims = staggered_node_grid(np.zeros(dim), box, dx, dim,
                                       boundary_width)
    mass = rho0 * dx ** dim
    h = 2 * dx
    lo = np.zeros(3)
    hi = np.zeros(3)
    lo[:dim] = -(wall_layers + 1) * dx - h
    hi[:dim] = box + (wall_layers + 1) * dx + h
    return {
        "name": name or "dam_break_%dd" % dim,
        "dim": dim,
        "dx": dx,
        "x": np.ascontiguousarray(x),
        "v": np.ascontiguousarray(v),
        "bx": np.ascontiguousarray(w),
        "bn": np.ascontiguousarray(wn),
        "bv": np.ascontiguousarray(wv),
        "particle_radius": dx / 2,
        "lattice_spacing": dx,
        "lattice_origin": origin,
        "lattice_dims": dims,
        "mass": mass,
        "rest_density": rho0,
        "kappa": mass / rho0,
        "domain_lo": lo,
        "domain_hi": hi,
        "box_hi": np.concatenate([box, np.zeros(3 - dim)]),
        "seed": seed,
    }


def box_colliders(scene):
    """The static collider of the reference's liquid demos: the complement
    of the box [0, box] with restitution and friction 0
    (Demos/SmthPartHydrodLiquidTest/SmthPartHydrodLiquidBuilder.h:47-49)."""
    from . import _abi
    return [_abi.collider(_abi.SURFACE_BOX, [0.0, 0.0, 0.0], list(scene["box_hi"]),
                          complement=True)]


def config1(seed=DEFAULT_SEED):
    """2D dam break: 100 x 100 particles at dx = 0.005 in the 0.5 x 0.5
    lower-left block of a unit box, 3 wall layers."""
    return dam_break(2, 0.005, (1.0, 1.0), (0.5, 0.5), seed, name="cfg1_2d_10k")


def config2(seed=DEFAULT_SEED, dx=1.0 / 150):
    """3D dam break: a 1 x 0.6 x 0.5 block in a 2 x 1 x 1 box.  dx = 1/150
    gives 150 x 90 x 75 = 1,012,500 particles (dx = 0.01 gives 300,000)."""
    return dam_break(3, dx, (2.0, 1.0, 1.0), (1.0, 0.6, 0.5), seed,
                     name="cfg2_3d_%dk" % round(
                         np.prod(np.rint(np.array([1.0, 0.6, 0.5]) / dx)) / 1000))


def config3(seed=DEFAULT_SEED, dx=1.0 / 300):
    """Config 2's scene at dx = 1/300: 8,100,000 particles."""
    return config2(seed, dx)


def small_2d(n=20, seed=DEFAULT_SEED, jitter=0.1):
    """Small 2D dam break for oracle parity (n x n fluid particles)."""
    dx = 0.5 / n
    return dam_break(2, dx, (1.0, 1.0), (0.5, 0.5), seed, jitter=jitter,
                     name="small_2d_%d" % (n * n))


def small_3d(nx=12, ny=8, nz=6, seed=DEFAULT_SEED, jitter=0.1):
    """Small 3D dam break: nx x ny x nz fluid particles at dx = 1/(2 nx) in a
    box twice the block's x length."""
    dx = 1.0 / (2 * nx)
    block = (nx * dx, ny * dx, nz * dx)
    box = (2 * nx * dx, (ny + 4) * dx, (nz + 4) * dx)
    return dam_break(3, dx, box, block, seed, jitter=jitter,
                     name="small_3d_%d" % (nx * ny * nz))


def _box_walls(dim, dx, box, wall_layers):
    """Cell-centred wall lattice around the box [0, box] with the normals of
    the box's complement (pointing inwards) and zero velocity."""
    nbox = np.rint(np.asarray(box[:dim]) / dx).astype(np.int64)
    waxes = [(np.arange(-wall_layers, nbox[a] + wall_layers) + 0.5) * dx
             for a in range(dim)]
    wgrid = np.meshgrid(*waxes, indexing="ij")
    w = np.stack([g.transpose().reshape(-1) for g in wgrid], axis=1)
    outside = np.any((w < 0) | (w > np.asarray(box[:dim])), axis=1)
    w = np.ascontiguousarray(w[outside])
    return w, box_complement_normals(w, np.zeros(dim), np.asarray(box[:dim])), np.zeros_like(w)


def grain_scene(x, v, particle_radius, bx, bn, bv, domain_lo, domain_hi,
                mass=1.0, name="grains"):
    """Scene of a DEM context: the grains are the particles, the walls the
    boundary particles.  The virtual lattice is a single node (a DEM context
    never generates virtual particles).  Grain-grain contacts reach 2r; the
    context's cells are sized by its kernel radius 4r."""
    dim = x.shape[1]
    lo = np.zeros(3)
    hi = np.zeros(3)
    lo[:dim] = domain_lo[:dim]
    hi[:dim] = domain_hi[:dim]
    return {
        "name": name,
        "dim": dim,
        "dx": 2 * particle_radius,
        "x": np.ascontiguousarray(x, dtype=np.float64),
        "v": np.ascontiguousarray(v, dtype=np.float64),
        "bx": np.ascontiguousarray(bx, dtype=np.float64).reshape(-1, dim),
        "bn": np.ascontiguousarray(bn, dtype=np.float64).reshape(-1, dim),
        "bv": np.ascontiguousarray(bv, dtype=np.float64).reshape(-1, dim),
        "particle_radius": particle_radius,
        "lattice_spacing": 2 * particle_radius,
        "lattice_origin": lo.copy(),
        "lattice_dims": np.ones(3, dtype=np.int64),
        "mass": mass,
        "rest_density": 1000.0,
        "kappa": mass / 1000.0,
        "domain_lo": lo,
        "domain_hi": hi,
    }


def sand_water(dim=3, dx=1.0 / 150, box=(1.0, 1.2, 1.0), bed=0.3, water=0.6,
               seed=DEFAULT_SEED, jitter=0.0, wall_layers=3, rho0=1000.0,
               grain_speed=0.0, gap=0.0):
    """Sand bed under a water column, both on the cell-centred lattice of
    spacing dx, in the box [0, box]: grains of radius r = dx / 2 (spacing 2r)
    fill 0 <= y < bed, water fills bed + gap <= y < bed + gap + water over
    the whole footprint (a gap > 0 makes it a falling column).  dx = 1/150 in the 1 x 1.2 x 1 box gives 150 x 45 x 150 =
    1,012,500 grains and 150 x 90 x 150 = 2,025,000 water particles.

    Returns (water scene, grain scene): the water's boundary is the walls (the
    grains are added each step by coupling), the grains' boundary the same
    walls.  ``jitter`` displaces grains by a uniform value in
    [-jitter dx, jitter dx] per coordinate (overlapping grains push apart);
    ``grain_speed`` gives every grain a random velocity of that size."""
    box = np.asarray(box, dtype=np.float64)[:dim]
    rng = np.random.Generator(np.random.PCG64(seed))
    n = np.rint(box / dx).astype(np.int64)
    nbed = int(round(bed / dx))
    nwat = int(round(water / dx))
    ngap = int(round(gap / dx))

    def layer_block(j0, j1):
        axes = [(np.arange(n[a]) + 0.5) * dx for a in range(dim)]
        axes[1] = (np.arange(j0, j1) + 0.5) * dx
        grid = np.meshgrid(*axes, indexing="ij")
        return np.stack([g.transpose().reshape(-1) for g in grid], axis=1)

    g = layer_block(0, nbed)
    if jitter:
        g = g + rng.uniform(-jitter * dx, jitter * dx, size=g.shape)
    gv = np.zeros_like(g)
    if grain_speed:
        d = rng.normal(size=g.shape)
        gv = grain_speed * d / np.linalg.norm(d, axis=1, keepdims=True)
    x = layer_block(nbed + ngap, nbed + ngap + nwat)
    w, wn, wv = _box_walls(dim, dx, box, wall_layers)
    origin, dims = staggered_node_grid(np.zeros(dim), box, dx, dim, 2)
    mass = rho0 * dx ** dim
    h = 2 * dx
    lo = np.zeros(3)
    hi = np.zeros(3)
    lo[:dim] = -(wall_layers + 1) * dx - h
    hi[:dim] = box + (wall_layers + 1) * dx + h
    water_scene = {
        "name": "sand_water_%dd" % dim,
        "dim": dim,
        "dx": dx,
        "x": np.ascontiguousarray(x),
        "v": np.zeros_like(x),
        "bx": w,
        "bn": wn,
        "bv": wv,
        "particle_radius": dx / 2,
        "lattice_spacing": dx,
        "lattice_origin": origin,
        "lattice_dims": dims,
        "mass": mass,
        "rest_density": rho0,
        "kappa": mass / rho0,
        "domain_lo": lo,
        "domain_hi": hi,
        "seed": seed,
    }
    grains = grain_scene(g, gv, dx / 2, w, wn, wv, lo, hi, mass=mass,
                         name="sand_bed_%dd" % dim)
    return water_scene, grains
