/*
 * vps_oracle.h -- FP64 CPU restatement of the reference's dual-particle SPH
 * projection.  TEST INFRASTRUCTURE ONLY: used by tests/, by the smoke check
 * and by the CPU-baseline leg of bench.py, as the checker.  The product
 * (libvps) never links or calls it.
 *
 * Every function follows a reference file:line, cited at its definition in
 * vps_oracle.c.  Parameters come in the product's own vps_params block so one
 * description drives both sides.
 */
#ifndef VPS_ORACLE_H
#define VPS_ORACLE_H

#include <stdint.h>
#include "../include/vps.h"

#ifdef __cplusplus
extern "C" {
#endif

enum or_search_mode {
    OR_SEARCH_HASH = 0,   /* HashGridSearcher as the reference runs it */
    OR_SEARCH_BRUTE = 1   /* ParticlesNearbySearcher::forEach brute force */
};

typedef struct or_stats {
    int64_t n_virtual;
    int64_t cg_iterations;
    double cg_error;          /* Eigen's tol_error = sqrt(r.r / b.b) */
    double rhs_norm;
    int32_t cg_status;        /* 0 ok, 2 no convergence */
    int32_t pad;
    double s_refresh;         /* wall-clock seconds per phase */
    double s_generate;
    double s_build;
    double s_solve;
    double s_gradient;
    double s_total;
} or_stats;

typedef struct or_sim or_sim;

or_sim* or_create(const vps_params* p, int32_t search_mode, int32_t threads);
void or_destroy(or_sim* s);
int or_set_boundary(or_sim* s, int64_t n, const double* x, const double* nrm,
                    const double* v);
int or_set_particles(or_sim* s, int64_t n, const double* x, const double* v);
int or_get_particles(or_sim* s, double* x, double* v);
int or_initialize(or_sim* s);
int or_refresh_real(or_sim* s);
int or_generate(or_sim* s);
int or_project(or_sim* s, or_stats* st);
int or_projection_step(or_sim* s, or_stats* st);
int or_advect(or_sim* s, double dt, double gravity);
/* the rest of the SPH step: move + colliders + densities, gravity,
 * div-free viscosity, and the whole DualParticleSphLiquid::advance */
int or_move_particles(or_sim* s, double dt, const vps_collider* col, int32_t ncol);
int or_apply_gravity(or_sim* s, double dt, double gravity);
int or_apply_viscosity(or_sim* s, double dt, double coeff);
int or_advance(or_sim* s, double dt, double gravity, int32_t enable_gravity,
               double viscosity, const vps_collider* col, int32_t ncol, or_stats* st);
/* boundary densities and volumes of the current boundary set */
int or_refresh_boundary(or_sim* s);
/* one DEMParticleSand::advance of the grains held as real particles, the
 * walls as boundary particles (velocities of the start of the pass) */
int or_dem_step(or_sim* s, double dt, double gravity, double young,
                double poisson, double friction_angle);

int64_t or_virtual_count(or_sim* s);
int or_get_virtual(or_sim* s, int64_t* node, double* rho, double* vol,
                   double* vbar, double* div, double* diag, double* p);
int or_set_pressure(or_sim* s, const double* p);
/* CG initial guess of the last project, in virtual order */
int or_get_initial_guess(or_sim* s, double* x0);
/* applyPressureGradient with the current pressures (updates velocities) */
int or_pressure_gradient(or_sim* s);
int or_get_real_fields(or_sim* s, double* rho, double* vol);
int or_get_boundary_fields(or_sim* s, double* rho, double* vol);
int or_get_scalars(or_sim* s, double* rho0, double* kappa, double* alpha0);
/* weight of every lattice node (length = product of lattice dims) */
int or_get_node_weights(or_sim* s, double* w);

/* neighbour query: which 0 real, 1 boundary, 2 virtual; indices sorted */
int or_query(or_sim* s, int32_t which, int64_t nq, const double* q,
             int64_t* offsets, int64_t* idx, int64_t cap, int64_t* needed);
/* same query but indices in the searcher's visit order */
int or_query_ordered(or_sim* s, int32_t which, int64_t nq, const double* q,
                     int64_t* offsets, int64_t* idx, int64_t cap,
                     int64_t* needed);

/* CG operator M of the current system (M = A^T or A by op_mode), CSR with
 * ascending columns.  Call with rowptr only to get nnz in rowptr[n]. */
int64_t or_matrix_nnz(or_sim* s);
int or_get_matrix(or_sim* s, int64_t* rowptr, int32_t* col, double* val);
/* out = M p, FP64, in reference virtual order */
int or_apply_operator(or_sim* s, const double* p, double* out);
/* Plants mistake m in the oracle's formulas (tests only; 0 clears). */
void or_set_mutation(int32_t m);

/* Eigen ConjugateGradient restated on a CSR matrix (identity
 * preconditioner).  Returns 0 on success, 2 on no convergence. */
int or_cg_csr(int64_t n, const int64_t* rowptr, const int32_t* col,
              const double* val, const double* b, double* x, double tol,
              int64_t max_iter, int64_t* iters, double* err);

/* Kernel helpers (SmoothedParticles.h:60-72, .cpp:13-23, 74-80). */
double or_kernel(int32_t dim, double radius, double dist);
double or_kernel_d1(int32_t dim, double radius, double dist);
double or_packed_kernel_sum(int32_t dim, double radius);
/* HashGridSearcher::getHashKey for a cell coordinate and bucket count */
int64_t or_hash_key(int32_t dim, const int32_t* coord, int64_t nbuckets);

#ifdef __cplusplus
}
#endif

#endif
