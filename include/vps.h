/*
 * vps.h -- C ABI of the MI355X virtual-particle SPH pressure projection.
 *
 * The library replaces the dual-particle ("virtual particle") projection of
 * the reference solver:
 *
 *   DualParticleSphLiquid<Dim>::initialize / advance / applyPressureForce
 *       (Cores/Physics/DualParticleSphLiquid.cpp:78-126)
 *   VirtualParticle<Dim>::generateParticles / setAlpha / setKappa / project
 *       (Cores/Structures/VirtualParticle.h:76-87, VirtualParticle.cpp:45-201)
 *   SmoothedParticles<Dim>::resetNearbySearcher / computeDensities /
 *   computeVolumes  (Cores/Structures/SmoothedParticles.h:53-55,110;
 *                    SmoothedParticles.cpp:25-41)
 *   HashGridSearcher<Dim>::reset / forEach
 *       (Cores/Structures/ParticlesNearbySearcher.cpp:10-35)
 *   IterativeSolver::solve  (Cores/Solvers/IterativeSolver.h:19-40)
 *
 * The reference exposes that path only as C++ template virtuals; a caller
 * binds this ABI instead (plain pointers, sizes and status codes, no torch or
 * Eigen types).  All host arrays use the reference memory layout: FP64,
 * array-of-structures with stride Dim (an Eigen Vector<Dim,double> is Dim
 * contiguous doubles).  The context owns every device buffer; calls on one
 * context must be serialised by the caller, separate contexts are
 * independent.  Errors are returned as status codes (the reference calls
 * std::exit, IterativeSolver.h:28-38); vps_last_error() gives the message.
 */
#ifndef VPS_H
#define VPS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VPS_ABI_VERSION 1

/* Status codes returned by every entry point. */
enum vps_status {
    VPS_OK = 0,
    VPS_ERR_INVALID_ARG = 1,   /* bad size, null pointer, empty set (the
                                  reference divides by zero there,
                                  ParticlesNearbySearcher.h:71 with .cpp:13) */
    VPS_ERR_NO_CONVERGENCE = 2,/* CG hit max_iterations (reference: exit(-1)) */
    VPS_ERR_NONFINITE = 3,     /* NaN/Inf in the right-hand side or solution
                                  (a bad learned or extrapolated warm start is
                                  not an error: the solve falls back to the
                                  plain start) */
    VPS_ERR_OOM = 4,           /* device allocation failed */
    VPS_ERR_HIP = 5,           /* any other HIP runtime error */
    VPS_ERR_STATE = 6,         /* call order violated (e.g. project before
                                  initialize) */
    VPS_ERR_IO = 7             /* a frame or checkpoint file could not be
                                  written or read (reference: std::exit) */
};

/* Linear operator handed to CG.  Eigen's ConjugateGradient with
 * UpLo = Lower|Upper on a column-major SparseMatrix multiplies by the
 * transpose of the assembled matrix (IterativeSolver.h:12, Types.h:21); the
 * assembled rows (VirtualParticle.cpp:149-169) are not symmetric because row
 * I carries the volume V_J of its neighbour, so the two choices differ. */
enum vps_operator {
    VPS_OP_TRANSPOSE = 0,      /* M = A^T: what the reference actually runs */
    VPS_OP_ROW = 1             /* M = A: the rows as written */
};

/* CG initial guess. */
enum vps_warm_start {
    VPS_WARM_INDEX_CARRY = 0,  /* x0[k] = p_prev[k] by position in the virtual
                                  list, zero past its end: what
                                  ParticlesBasedData::resize does to the
                                  pressures (ParticlesBasedData.h:28 via
                                  VirtualParticle.cpp:95) */
    VPS_WARM_ZERO = 1,         /* x0 = 0 */
    VPS_WARM_NODE_KEYED = 2,   /* x0 = previous pressure at the same lattice
                                  node, zero for nodes that are new */
    VPS_WARM_LEARNED = 3,      /* x0 = y0 p1 + y1 e + y2 c with c = tau *
                                  MLP(features), the learned correction to
                                  the extrapolation below (tau the RMS error
                                  of that extrapolation at the last solve),
                                  and y the Galerkin weights (V^T M V) y =
                                  V^T b, V = [p1 e c], formed on the device
                                  per solve (no reference counterpart).
                                  This start and the next are guarded: one
                                  whose initial residual is not finite or
                                  exceeds ||b|| is replaced by the
                                  index-carried start */
    VPS_WARM_EXTRAPOLATE = 4   /* x0 = p1 + e: polynomial extrapolation of the
                                  node's last three pressures, e = dp1 with two
                                  known, 2 dp1 - dp2 with three (no reference
                                  counterpart) */
};

/* Where rho0 and kappa come from. */
enum vps_rest_mode {
    VPS_FROM_INIT = 0,         /* reference: rho0 = max real density at
                                  initialize (DualParticleSphLiquid.cpp:91),
                                  kappa = 1/packed kernel sum (:98) */
    VPS_FROM_PARAMS = 1        /* use vps_params.rest_density / .kappa */
};

typedef struct vps_params {
    int32_t dim;               /* 2 or 3 */
    int32_t op_mode;           /* enum vps_operator */
    int32_t warm_start;        /* enum vps_warm_start */
    int32_t rest_density_mode; /* enum vps_rest_mode */
    int32_t kappa_mode;        /* enum vps_rest_mode */
    int32_t reserved0;
    double particle_radius;    /* real and boundary kernel radius h = 4 r
                                  (SmoothedParticles.h:45, kernel_rate 4) */
    double lattice_spacing;    /* virtual kernel radius H = 4 s
                                  (DualParticleSphLiquid.h:28) */
    double lattice_origin[3];  /* node (0,0,0) position (Grid.h:51) */
    int32_t lattice_dims[3];   /* node counts per axis; dims[2] = 1 in 2D */
    int32_t reserved1;
    double mass;               /* real particle mass; boundary mass = mass
                                  (DualParticleSphLiquid.cpp:84) */
    double rest_density;       /* used when rest_density_mode = FROM_PARAMS */
    double kappa;              /* used when kappa_mode = FROM_PARAMS */
    double tolerance;          /* CG relative tolerance, reference 1e-6 */
    int64_t max_iterations;    /* < 0: 2 * N_V, Eigen's default */
    double domain_lo[3];       /* axis-aligned box that real and boundary
                                  particles stay in; particles outside are
                                  clamped to the border cells (still correct,
                                  only slower) */
    double domain_hi[3];
} vps_params;

/* Per-call statistics.  Times are device times (HIP events on the library
 * stream) in milliseconds. */
typedef struct vps_stats {
    int64_t n_real;
    int64_t n_boundary;
    int64_t n_virtual;
    int64_t n_active_bricks;
    int64_t cg_iterations;     /* Eigen convention: the converging iteration is
                                  not counted (ConjugateGradient.h) */
    double cg_rel_residual;    /* sqrt(r.r / b.b) of the recursive residual */
    double cg_true_rel_residual; /* ||b - M x|| / ||b||, FP64 on the device,
                                  computed only when requested (< 0 if not) */
    double rhs_norm;           /* ||b|| = ||D|| */
    double alpha0;
    double kappa;
    double rest_density;
    double ms_sort;            /* real counting sort + density */
    double ms_generate;        /* virtual marking + fused virtual sums */
    double ms_diag;            /* Laplacian diagonal / alpha0 */
    double ms_cg;              /* warm start + CG */
    double ms_gradient;        /* pressure gradient + velocity update */
    double ms_total;
    double ms_spmv_avg;        /* average duration of one CG SpMV launch */
    double ms_update_avg;      /* average duration of one CG update launch */
    int64_t spmv_launches;
    int64_t reserved[4];       /* [0] residual replacements of the last solve,
                                  [1] host microseconds spent launching its CG
                                  kernels, [2] 1 when a learned or
                                  extrapolated start was rejected (initial
                                  residual non-finite or above ||b||, or the
                                  iteration from it overflowed) and the solve
                                  ran again from the plain, index-carried
                                  start */
} vps_stats;

typedef struct vps_context vps_context;

/* ---- lifetime --------------------------------------------------------- */

/* Size checks for foreign-language bindings. */
int32_t vps_abi_version(void);
int64_t vps_sizeof_params(void);
int64_t vps_sizeof_stats(void);

/* Creates a context on the current HIP device.  Replaces the constructors
 * DualParticleSphLiquid(grid, particleRadius) and
 * VirtualParticle(grid, grid.spacing()) (DualParticleSphLiquid.h:27-29).
 * VPS_ERR_INVALID_ARG for a domain of more than 2^30 cells, a lattice of
 * more than 2^31 - 1024 nodes, or (3D) more than 2^21 - 1 bricks of 8^3
 * nodes or 2^24 or more cells (of size h) along x. */
int vps_create(const vps_params* params, vps_context** out);
void vps_destroy(vps_context* ctx);

/* Message of the last failed call on ctx (or of the last failed vps_create
 * when ctx is NULL).  Never NULL. */
const char* vps_last_error(const vps_context* ctx);

/* ---- state ------------------------------------------------------------ */

/* Boundary (solid wall) particles: positions, inward unit normals and
 * velocities, each n*dim doubles.  A zero normal marks a grain (a moving
 * sphere, e.g. a sand grain): its normal in the pressure-gradient boundary
 * term is the unit vector from the grain to the fluid particle (the
 * reference defines no grain normal).  Replaces BoundaryParticles::addSurface
 * and the boundary velocity field (DualParticleSphLiquid.cpp:79-84): the
 * caller samples the walls.  May be called again to move the walls (their
 * densities are recomputed at the next vps_initialize or
 * vps_refresh_boundary). */
int vps_set_boundary(vps_context* ctx, int64_t n, const double* x,
                     const double* normals, const double* v);

/* Recomputes boundary neighbour lists, densities and volumes without a full
 * re-initialisation (moving boundaries). */
int vps_refresh_boundary(vps_context* ctx);

/* Implicit surfaces of Cores/Geometries/ImplicitSurface.h:27-113. */
enum vps_surface_kind {
    VPS_SURFACE_BOX = 0,       /* ImplicitBox(min corner a, lengths b) */
    VPS_SURFACE_SPHERE = 1,    /* ImplicitSphere(center a, radius b[0]) */
    VPS_SURFACE_PLANE = 2,     /* ImplicitPlane(position a, direction b) */
    VPS_SURFACE_ELLIPSOID = 3  /* ImplicitEllipsoid(semi-axes b; the
                                  reference's inside test and closest point
                                  ignore its center, kept as is) */
};
typedef struct vps_surface {
    int32_t kind;         /* vps_surface_kind */
    int32_t complement;   /* 1: ComplementarySurface of it (Surface.h:50-69) */
    double a[3];
    double b[3];
} vps_surface;

/* BoundaryParticles::addSurface (Cores/Structures/SmoothedParticles.cpp:
 * 111-155): appends the lattice points of the region center +- half_lengths
 * (Dim doubles each) that lie inside the surface (signed distance <= 0) to
 * the walls, with the surface's closest normal and zero velocity, then
 * refreshes the resident boundary as vps_set_boundary does.  The lattice
 * is the reference's: hexagonal rows of spacing 2r in 2D, the face and node
 * centres of a staggered grid of spacing 2 sqrt(2) r in 3D (r the particle
 * radius), in its order.  Sampled on the GPU. */
int vps_add_surface(vps_context* ctx, const vps_surface* surface, const double* center,
                    const double* half_lengths);

/* Shapes::generateBox without Poisson sampling (Cores/Utilities/Shapes.h:
 * 32-58): the same lattice over center +- half_lengths, every point (a block
 * of fluid particles).  *count is always written; x (count * dim doubles)
 * is filled when not NULL and capacity >= count. */
int vps_generate_box(vps_context* ctx, const double* center, const double* half_lengths,
                     int64_t capacity, double* x, int64_t* count);

/* The walls as given (vps_set_boundary, vps_add_surface): *count and, for
 * non-NULL pointers, positions, normals and velocities (count * dim each). */
int vps_get_boundary(vps_context* ctx, int64_t* count, double* x, double* normals,
                     double* v);

/* Real (fluid) particle positions and velocities, n*dim doubles each. */
int vps_set_particles(vps_context* ctx, int64_t n, const double* x,
                      const double* v);

/* Copies the real particle state back, in the caller's original order. */
int vps_get_particles(vps_context* ctx, double* x, double* v);

/* ---- the path --------------------------------------------------------- */

/* DualParticleSphLiquid::initialize minus the boundary sampling
 * (DualParticleSphLiquid.cpp:78-110): real rho and frozen V, rho0, boundary
 * rho and V, virtual generation, alpha0 = max diagonal, kappa. */
int vps_initialize(vps_context* ctx, vps_stats* stats);

/* The two lines of SmthParticleHydrodLiquid::moveParticles that precede the
 * path (SmthParticleHydrodLiquid.cpp:80-81): neighbour structure rebuild and
 * real densities. */
int vps_refresh_real(vps_context* ctx);

/* VirtualParticle::generateParticles (VirtualParticle.cpp:45-62). */
int vps_generate_particles(vps_context* ctx);

/* VirtualParticle::project (VirtualParticle.cpp:76-86): right-hand side,
 * Laplacian, CG, pressure gradient and velocity update, in place on the
 * resident velocities.  stats may be NULL. */
int vps_project(vps_context* ctx, vps_stats* stats);

/* One timed projection step as the reference orders it:
 * vps_refresh_real + vps_generate_particles + vps_project. */
int vps_projection_step(vps_context* ctx, vps_stats* stats);

/* x += v dt, then v.y -= g dt (SmthParticleHydrodLiquid.cpp:73, 84-88),
 * on the resident state.  Not part of the projection. */
int vps_advect(vps_context* ctx, double dt, double gravity);

/* Requests that the next vps_project also computes the FP64 true residual
 * ||b - M x||/||b|| into stats.cg_true_rel_residual. */
int vps_set_check_true_residual(vps_context* ctx, int32_t enable);

/* Stamps the start and end of every fused CG SpMV and CG update launch of
 * the next solves (events of hipExtLaunchKernelGGL: the kernels' own
 * execution), reported as stats.ms_spmv_avg / ms_update_avg. */
int vps_set_kernel_timing(vps_context* ctx, int32_t enable);

/* Asynchronous steps (no counterpart in the reference, whose steps run on
 * the host): with enable != 0, vps_project, vps_projection_step and
 * vps_advance return once the step's work is queued on the context's
 * stream instead of waiting for its end, so the next step's kernels queue
 * behind it with no idle gap.  Their stats then hold zero phase times
 * (ms_*); the counts, iterations and residuals are known on the host before
 * the step's last kernels.  Every call that returns data (vps_get_*,
 * frames) and vps_synchronize wait for the queued work; a device error of a
 * queued step is reported by a later call.  Default 0 (each step waits for
 * its end). */
int vps_set_async_steps(vps_context* ctx, int32_t enable);

/* Blocks until all work on the context's stream is done. */
int vps_synchronize(vps_context* ctx);

/* ---- sand grains (DEM) --------------------------------------------- */

/* DEMParticle<Dim> material (Cores/Structures/DEMParticle.cpp:13-28):
 * K_norm = young * r, K_tang = K_norm * poisson, Coulomb limit
 * tan(friction_angle).  Reference defaults: 1e9, 0.3, 0.5 rad. */
typedef struct vps_dem_params {
    double young;
    double poisson;
    double friction_angle;
    double reserved;
} vps_dem_params;

/* One DEMParticleSand<Dim>::advance (DEMParticleSand.cpp:89-152) on a context
 * whose real particles are the grains (radius = vps_params.particle_radius)
 * and whose boundary particles are the walls: x += v dt; v += dt * sum of
 * grain contacts within 2r + wall normal / friction forces within 4r;
 * v.y -= gravity * dt.  The lattice of such a context is unused. */
int vps_dem_step(vps_context* grains, double dt, double gravity,
                 const vps_dem_params* params);

/* Sets the boundary of the water context to its walls (vps_set_boundary)
 * followed by the current grains of a DEM context (zero normals, the grains'
 * velocities), then recomputes the boundary densities and volumes: the
 * grains act on the water as moving boundary particles.  Both contexts must
 * be on the same device. */
int vps_couple_grains(vps_context* water, vps_context* grains);

/* ---- the rest of the SPH step ------------------------------------------ */

/* A static collider (StaticCollider, Cores/Geometries/Collider.h:58-78):
 * a surface with restitution and friction coefficients (Collider.cpp:36-63;
 * the reference's default is 0 and 0). */
typedef struct vps_collider {
    vps_surface surface;
    double restitution;
    double friction;
} vps_collider;

/* SmthParticleHydrodLiquid::moveParticles (Cores/Physics/
 * SmthParticleHydrodLiquid.cpp:72-82): x += v dt; each collider resolves
 * every particle in turn (signed distance - r <= 0: velocity reflected with
 * restitution and friction, position moved to the closest surface point + r
 * along the normal; Collider.cpp:13-63); then the neighbour rebuild and the
 * real densities (vps_refresh_real). */
int vps_move_particles(vps_context* ctx, double dt, const vps_collider* colliders,
                       int32_t n_colliders);

/* applyExternalForces (SmthParticleHydrodLiquid.cpp:84-88): v.y -= g dt. */
int vps_apply_gravity(vps_context* ctx, double dt, double gravity);

/* applyViscosityForce (SmthParticleHydrodLiquid.cpp:90-98) with
 * divFreeLaplacianAtDataPoint (ParticlesBasedVectorField.cpp:56-69):
 * v_i += nu dt 2 (Dim + 2) m sum_j ((v_j - v_i).d / |d|^2) gradK(d) / rho_j,
 * d = x_j - x_i over the real neighbours within h (d = 0 skipped), all from
 * the velocities before the pass.  Needs the current densities (after
 * vps_move_particles or vps_refresh_real). */
int vps_apply_viscosity(vps_context* ctx, double dt, double coefficient);

/* DualParticleSphLiquid::advance (DualParticleSphLiquid.cpp:107-116):
 * vps_move_particles, gravity when enabled (the reference's default is off,
 * SmthParticleHydrodLiquid.h:25), viscosity (0.01 there, :28), virtual
 * generation and the projection. */
typedef struct vps_step_params {
    double dt;
    double gravity;                 /* 9.8 (Utilities/Constants.h:7) */
    double viscosity;               /* nu; 0 skips the pass */
    const vps_collider* colliders;  /* may be NULL when n_colliders = 0 */
    int32_t n_colliders;
    int32_t enable_gravity;
} vps_step_params;
int vps_advance(vps_context* ctx, const vps_step_params* step, vps_stats* stats);

/* ---- frames and checkpoints ------------------------------------------ */

/* DualParticleSphLiquid::writeFrame (Cores/Physics/DualParticleSphLiquid.cpp:
 * 41-76): particles.mesh, virtual_particles.mesh and boundary_particles.mesh
 * in frame_dir (an existing directory).  Each file: uint32 count, count
 * float[Dim] positions, in 3D count float[3] normals (0, 0, 1), then count
 * floats: |v| for the fluid, V_I for the virtual particles of the last
 * generation, none for the boundary (walls and coupled grains, input order). */
int vps_write_frame(vps_context* ctx, const char* frame_dir);

/* SmthParticleHydrodLiquid::saveFrame (SmthParticleHydrodLiquid.cpp:34-43):
 * particles.sav and velocities.sav, the raw FP64 Dim-vectors in caller
 * order (ParticlesAttribute::save, ParticlesAttribute.h:90), plus
 * projection.sav: the state the reference does not keep (frozen volumes,
 * rho0, kappa, alpha0, the last solve's pressures and layout, the learned
 * warm start's parameters, moments and history).  The last solve is kept
 * when the call follows a projection (no generation in between). */
int vps_save_frame(vps_context* ctx, const char* frame_dir);

/* SmthParticleHydrodLiquid::loadFrame (SmthParticleHydrodLiquid.cpp:45-55):
 * positions and velocities from the .sav files (the particle count is their
 * size).  With projection.sav the context continues exactly where
 * vps_save_frame left it (next call: vps_projection_step or refresh,
 * generate, project); without it, the load re-initialises from the loaded
 * positions as the reference does (vps_initialize).  The boundary is the
 * context's own (vps_set_boundary). */
int vps_load_frame(vps_context* ctx, const char* frame_dir);

/* description.yaml of an output directory (Simulator.cpp:33-39 with
 * DualParticleSphLiquid::writeDescription, DualParticleSphLiquid.cpp:8-39):
 * dimension and the three point-list objects of the viewer. */
int vps_write_description(int32_t dim, const char* output_dir);

/* end_frame.txt = frame + 1 (Simulator.cpp:86-88). */
int vps_write_end_frame(const char* output_dir, int32_t frame);

/* Simulator::writeAndSaveToFrameDirectory (Simulator.cpp:79-89): creates
 * <output_dir>/<frame>, writes and saves the frame there, updates
 * end_frame.txt. */
int vps_write_and_save_frame(vps_context* ctx, const char* output_dir, int32_t frame);

/* Simulator::loadFromFrameDirectory (Simulator.cpp:91-100): vps_load_frame
 * of <output_dir>/<frame>; VPS_ERR_IO when the directory is missing. */
int vps_load_frame_number(vps_context* ctx, const char* output_dir, int32_t frame);

/* ---- read-back (tests, output) --------------------------------------- */

int vps_get_virtual_count(vps_context* ctx, int64_t* n_virtual);

/* Virtual particle fields in the reference's order (lattice-linear, x
 * fastest; VirtualParticle.cpp:55-57).  Any pointer may be NULL.
 * node: linear node index; rho: rho_I; vol: V_I; div: D_I (CG right-hand
 * side); diag: Laplacian diagonal; pressure: CG solution p_I. */
int vps_get_virtual(vps_context* ctx, int64_t* node, double* rho, double* vol,
                    double* div, double* diag, double* pressure);

/* Real particle densities (current) and frozen volumes, caller order. */
int vps_get_real_fields(vps_context* ctx, double* rho, double* vol);

/* Boundary densities and volumes in the order given to vps_set_boundary. */
int vps_get_boundary_fields(vps_context* ctx, double* rho, double* vol);

/* Scalars fixed at initialize. */
int vps_get_scalars(vps_context* ctx, double* rest_density, double* kappa,
                    double* alpha0);

/* Neighbour query against the resident real (which = 0), boundary (1) or
 * virtual (2) set, with the reference's membership rule
 * (x_j - q).squaredNorm() < radius^2 in FP64 (ParticlesNearbySearcher.h:30-34,
 * .cpp:32), radius = h for sets 0/1 and H for set 2.  Writes a CSR:
 * offsets[nq+1] and, per query, the caller-order indices (reference virtual
 * order for set 2) sorted ascending.  With indices == NULL only offsets and
 * *needed are written (size query).  Returns VPS_ERR_INVALID_ARG with
 * *needed set when capacity is too small. */
int vps_query_neighbors(vps_context* ctx, int32_t which, int64_t nq,
                        const double* q, int64_t* offsets, int64_t* indices,
                        int64_t capacity, int64_t* needed);

/* Applies the CG operator M of the current system (VPS_OP_TRANSPOSE: A^T,
 * VPS_OP_ROW: A, with the assembled diagonal) to p, both in the reference
 * virtual order, with the same FP32 stencil CG uses.  Replaces the
 * SparseMatrix * vector product inside Eigen's CG (IterativeSolver.h:35). */
int vps_debug_apply_operator(vps_context* ctx, const double* p, double* out);

/* Test support, no counterpart in the reference (3D): the number of terms
 * that each FP32 neighbour walk sums over, per particle in caller order.
 * which = 0: the real density walk (real particles within h of each real
 * particle, itself included; SmoothedParticles.cpp:25-31), 1: the boundary
 * density walk (boundary particles within h of each boundary particle),
 * 2: the pressure gradient's walk over the virtual particles within H of
 * each real particle (VirtualParticle.cpp:118-124; -1 for a particle whose
 * span is beyond the staged path), 3: the viscosity walk (real particles
 * within h of each real particle, itself excluded).  The walks decide
 * membership on FP32
 * distances: only a pair within rounding of the cutoff, whose term is zero
 * to FP32 resolution, can differ from the exact sets of
 * vps_query_neighbors.  Call it on the current generation (after a
 * projection, before advection). */
int vps_debug_member_counts(vps_context* ctx, int32_t which, int32_t* counts);

/* ---- learned warm start (no reference counterpart) ------------------- */

/* Online training of the warm-start MLP on the (features, solution) pairs
 * of the converged solves: after each vps_project with training enabled
 * (and two earlier solves with history), steps_per_solve Adam steps
 * (beta 0.9 / 0.999, eps 1e-8) run on the GPU, each on `batch` virtual
 * particles sampled uniformly with replacement.  Training keeps the node
 * history even when the warm start in use is a plain one. */
int vps_mlp_set_training(vps_context* ctx, int32_t enable, int32_t batch,
                         double learning_rate, int32_t steps_per_solve,
                         uint64_t seed);

/* Parameter I/O for checks and checkpoints: 4,801 FP32 values in the order
 * W1[64x8], b1[64], W2[64x64], b2[64], W3[1x64], b3[1] (row-major). */
int64_t vps_mlp_parameter_count(void);
int vps_mlp_get_parameters(vps_context* ctx, float* out);
int vps_mlp_set_parameters(vps_context* ctx, const float* in);

/* One forward + backward pass over given features (n x 8, FP32) and targets
 * (n, FP32): writes the loss (mean squared error) and the gradient of every
 * parameter (same order as above).  Used to check the hand-written MFMA
 * kernels against the FP64 oracle network. */
int vps_mlp_loss_and_grad(vps_context* ctx, int64_t n, const float* features,
                          const float* targets, double* loss, float* grad);

/* vps_mlp_loss_and_grad on device memory, allocating nothing (no
 * counterpart in the reference: for callers that cannot allocate device
 * memory, such as code under HIP graph capture, tight serving or training
 * loops and harnesses that forbid hipMalloc).  d_features (n x 8 FP32),
 * d_targets (n FP32), d_loss (1 double) and d_grad (4,801 FP32) are device
 * pointers on the context's device; n <= 67,108,864.  The per-block partial
 * sums live in a fixed 4.92 MB scratch array of the library on each device
 * (set up at vps_create), so calls on one device that overlap in time
 * (other contexts, other threads) race on it.  The kernels are queued on the
 * context's stream and the call returns without waiting: the results are
 * ready after vps_synchronize(ctx).  Same kernels and results as
 * vps_mlp_loss_and_grad. */
int vps_mlp_loss_and_grad_device(vps_context* ctx, int64_t n, const float* d_features,
                                 const float* d_targets, double* d_loss, float* d_grad);

/* Inference over given features: out[n]. */
int vps_mlp_forward(vps_context* ctx, int64_t n, const float* features,
                    float* out);

/* Loss of the last in-situ Adam step (mean squared error of the scaled
 * pressure change on its batch, before the step) and the number of Adam
 * steps taken.  A step whose loss or reduced gradient is non-finite is
 * skipped (parameters and moments unchanged) and not counted. */
int vps_mlp_training_state(vps_context* ctx, double* last_loss, int64_t* steps);

/* Per-virtual-particle features of the current step in reference order
 * (n_virtual x 8 FP32). */
int vps_mlp_get_features(vps_context* ctx, float* out);

#ifdef __cplusplus
}
#endif

#endif /* VPS_H */
