/*
 * vps_oracle.c -- FP64 CPU restatement of the reference's virtual-particle
 * ("dual-particle") SPH projection.
 *
 * TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py as the checker and the CPU baseline; the
 * product library never calls it.
 *
 * Pinning (what checks each function; the reference ships no tests, golden
 * vectors or fixtures for this path):
 *   kernels / packed sum     closed forms, quadrature, finite differences
 *   hash searcher            brute force (same sets), hand-computed keys
 *   densities / sums         brute-force O(N^2) sums, lattice invariants
 *   virtual generation       brute force over all nodes, threshold cases
 *   Laplacian assembly       interior row sums, A vs A^T on a tiny system,
 *                            the quadratic pressure |x|^2/2 (interior
 *                            A p = -Dim V_I / s^Dim: sign and scale)
 *   CG                       dense solves, 1D Poisson, stopping rule
 *   pressure gradient        constant pressure, hand-built boundary cases,
 *                            a linear pressure (interior dv = -grad p),
 *                            one pressurised node (radial push out to H)
 * or_set_mutation() plants one of a list of plausible transcription
 * mistakes, so the tests can show that each of these checks fails on it.
 * The last-bit FP64 values of the reference depend on Eigen's SIMD reduction
 * order inside CG (dot/squaredNorm over long vectors); that part is "parity
 * unpinned" -- only compile of the reference against Eigen would fix it and
 * Eigen is not available here.
 *
 * Arithmetic follows the reference expression by expression, in FP64, with
 * FP contraction off (build flag -ffp-contract=off), so that a 3-vector
 * squared norm is ((dx*dx + dy*dy) + dz*dz) as Eigen's vectorised redux
 * evaluates it for a Vector3d with SSE2 packets.
 */
#include "vps_oracle.h"

#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* std::numbers::inv_pi, sqrt2, sqrt3 as doubles */
#define OR_INV_PI 0.31830988618379067154
#define OR_SQRT2 1.41421356237309504880
#define OR_SQRT3 1.73205080756887729353

/* Planted mistakes (tests only; 0 = the reference's formulas). */
enum {
    OR_MUT_NONE = 0,
    OR_MUT_GRAD_SIGN = 1,      /* virtual gradient term with the wrong sign */
    OR_MUT_GRAD_NO_VOL = 2,    /* ... without V_J */
    OR_MUT_GRAD_KERNEL_H = 3,  /* ... with the real kernel h instead of H */
    OR_MUT_COEF_HALF = 4,      /* alpha_IJ without the factor 2 */
    OR_MUT_COEF_SIGN = 5,      /* alpha_IJ with the wrong sign */
    OR_MUT_COEF_KERNEL_H = 6   /* alpha_IJ from W'_h instead of W'_H */
};
static int g_mutation = OR_MUT_NONE;
void or_set_mutation(int32_t m) { g_mutation = m; }

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* ------------------------------------------------------------------------
 * Cubic spline kernel.  SmoothedParticles constructor
 * (Cores/Structures/SmoothedParticles.cpp:13-23) and kernel /
 * firstDerivativeKernel (SmoothedParticles.h:60-72).
 * ---------------------------------------------------------------------- */
typedef struct {
    int dim;
    double radius, kr, kr2, inv_kr, inv_kr2, c0, c1;
} or_kern;

static void kern_init(or_kern* k, int dim, double radius) {
    const double kernel_rate = 4; /* SmoothedParticles.h:45 default */
    k->dim = dim;
    k->radius = radius;
    k->kr = radius * kernel_rate;
    k->kr2 = k->kr * k->kr;
    k->inv_kr = 1 / k->kr;
    k->inv_kr2 = 1 / k->kr2;
    k->c0 = OR_INV_PI * k->inv_kr2 * (dim == 2 ? (double)40 / 7 : 8 * k->inv_kr);
    k->c1 = 6 * k->inv_kr * k->c0;
}

static inline double kern_W(const or_kern* k, double distance) {
    const double x = distance * k->inv_kr;
    if (x < .5) return k->c0 * (6 * x * x * (x - 1) + 1);
    else if (x < 1) return k->c0 * 2 * (1 - x) * (1 - x) * (1 - x);
    else return 0;
}

static inline double kern_dW(const or_kern* k, double distance) {
    const double x = distance * k->inv_kr;
    if (x < .5) return k->c1 * x * (3 * x - 2);
    else if (x < 1) return k->c1 * (1 - x) * (x - 1);
    else return 0;
}

/* squaredNorm of a Dim-vector as Eigen evaluates it (see file header) */
static inline double sqnorm(int dim, const double* d) {
    if (dim == 2) return d[0] * d[0] + d[1] * d[1];
    return (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
}
static inline double dot3(int dim, const double* a, const double* b) {
    if (dim == 2) return a[0] * b[0] + a[1] * b[1];
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

/* kernel(deltaPos) = kernel(deltaPos.norm()) (SmoothedParticles.h:95) */
static inline double kern_Wv(const or_kern* k, const double* d) {
    return kern_W(k, sqrt(sqnorm(k->dim, d)));
}

/* gradientKernel(d) = -firstDerivativeKernel(|d|) * d.normalized()
 * (SmoothedParticles.h:96-98); Eigen normalized() divides by sqrt(z) when
 * z > 0 and returns d unchanged otherwise. */
static inline void kern_grad(const or_kern* k, const double* d, double* g) {
    const double z = sqnorm(k->dim, d);
    const double s = -kern_dW(k, sqrt(z));
    g[0] = g[1] = g[2] = 0;
    for (int a = 0; a < k->dim; a++) {
        const double nd = z > 0 ? d[a] / sqrt(z) : d[a];
        g[a] = s * nd;
    }
}

double or_kernel(int32_t dim, double radius, double dist) {
    or_kern k;
    kern_init(&k, dim, radius);
    return kern_W(&k, dist);
}
double or_kernel_d1(int32_t dim, double radius, double dist) {
    or_kern k;
    kern_init(&k, dim, radius);
    return kern_dW(&k, dist);
}

/* getPackedKernelSum (SmoothedParticles.cpp:74-80) */
static double packed_sum(const or_kern* k) {
    if (k->dim == 2)
        return kern_W(k, 0) + kern_W(k, k->radius * 2) * 6 +
               kern_W(k, k->radius * 2 * OR_SQRT3) * 6;
    return kern_W(k, 0) + kern_W(k, k->radius * 2) * 12 +
           kern_W(k, k->radius * 2 * OR_SQRT2) * 6 +
           kern_W(k, k->radius * 2 * OR_SQRT3) * 24;
}
double or_packed_kernel_sum(int32_t dim, double radius) {
    or_kern k;
    kern_init(&k, dim, radius);
    return packed_sum(&k);
}

/* ------------------------------------------------------------------------
 * Neighbour search.  HashGridSearcher (ParticlesNearbySearcher.h:37-74,
 * .cpp:10-35) on a Grid with spacing = kernel radius and origin 0
 * (ParticlesNearbySearcher.h:57), getQuadraticLower (Grid.h:70) and
 * quadraticNearbyDataPoints (Grid.cpp:26-45).  Buckets are stored as CSR
 * rather than vector<vector<int>>: same contents, same insertion order.
 * The brute-force mode is ParticlesNearbySearcher::forEach (.h:30-34).
 * ---------------------------------------------------------------------- */
typedef struct {
    int dim;
    double inv_spacing;
    int64_t nbuckets;
    int64_t* start;   /* nbuckets + 1 */
    int32_t* items;
} or_hash;

int64_t or_hash_key(int32_t dim, const int32_t* c, int64_t nbuckets) {
    /* (_kPrime0 * ullong(c0) | _kPrime1 * ullong(c1) [| _kPrime2 * ullong(c2)])
     * % buckets, ParticlesNearbySearcher.h:67-73 */
    uint64_t k = (uint64_t)73856093 * (uint64_t)(int64_t)c[0] |
                 (uint64_t)19349663 * (uint64_t)(int64_t)c[1];
    if (dim == 3) k |= (uint64_t)83492791 * (uint64_t)(int64_t)c[2];
    return (int64_t)(int)(k % (uint64_t)nbuckets);
}

static inline void quadratic_lower(const or_hash* h, const double* pos,
                                   int32_t* lo) {
    for (int a = 0; a < h->dim; a++)
        lo[a] = (int32_t)floor((pos[a] - 0.0) * h->inv_spacing - 1.0 / 2);
}

static void hash_free(or_hash* h) {
    free(h->start);
    free(h->items);
    h->start = NULL;
    h->items = NULL;
    h->nbuckets = 0;
}

/* HashGridSearcher::reset (ParticlesNearbySearcher.cpp:10-18) */
static void hash_reset(or_hash* h, int dim, double kr, int64_t n,
                       const double* x) {
    hash_free(h);
    h->dim = dim;
    h->inv_spacing = 1 / kr;
    h->nbuckets = n * 2;
    if (n == 0) return;
    h->start = (int64_t*)calloc((size_t)h->nbuckets + 1, sizeof(int64_t));
    h->items = (int32_t*)malloc((size_t)n * sizeof(int32_t));
    int64_t* key = (int64_t*)malloc((size_t)n * sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) {
        int32_t c[3];
        quadratic_lower(h, x + 3 * i, c);
        for (int a = 0; a < dim; a++) c[a] += 1;
        key[i] = or_hash_key(dim, c, h->nbuckets);
        h->start[key[i] + 1]++;
    }
    for (int64_t b = 0; b < h->nbuckets; b++) h->start[b + 1] += h->start[b];
    int64_t* fill = (int64_t*)malloc((size_t)h->nbuckets * sizeof(int64_t));
    memcpy(fill, h->start, (size_t)h->nbuckets * sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) h->items[fill[key[i]]++] = (int32_t)i;
    free(fill);
    free(key);
}

static int cmp_i64(const void* a, const void* b) {
    const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return x < y ? -1 : x > y;
}

typedef void (*or_cb)(void* ctx, int32_t j, const double* xj);

typedef struct {
    or_kern k;
    int64_t n;
    double* x;      /* n*3, z = 0 in 2D */
    double* rho;
    double* vol;
    double mass;
    int search_mode;
    or_hash hash;
} or_set;

/* SmoothedParticles::forEachNearby -> HashGridSearcher::forEach
 * (SmoothedParticles.h:111-113; ParticlesNearbySearcher.cpp:20-35) */
static void set_for_each(const or_set* s, const double* q, or_cb cb,
                         void* ctx) {
    const int dim = s->k.dim;
    if (s->n == 0) return;
    if (s->search_mode == OR_SEARCH_BRUTE) {
        for (int64_t j = 0; j < s->n; j++) {
            double d[3] = {s->x[3 * j] - q[0], s->x[3 * j + 1] - q[1],
                           s->x[3 * j + 2] - q[2]};
            if (sqnorm(dim, d) < s->k.kr2) cb(ctx, (int32_t)j, s->x + 3 * j);
        }
        return;
    }
    int64_t keys[27];
    int cnt = 0;
    int32_t lo[3];
    if (s->hash.nbuckets == 0) return; /* not built: the reference's % 0 */
    quadratic_lower(&s->hash, q, lo);
    if (dim == 2) {
        for (int j = 0; j < 3; j++)
            for (int i = 0; i < 3; i++) {
                int32_t c[3] = {lo[0] + i, lo[1] + j, 0};
                keys[cnt++] = or_hash_key(2, c, s->hash.nbuckets);
            }
    } else {
        for (int kk = 0; kk < 3; kk++)
            for (int j = 0; j < 3; j++)
                for (int i = 0; i < 3; i++) {
                    int32_t c[3] = {lo[0] + i, lo[1] + j, lo[2] + kk};
                    keys[cnt++] = or_hash_key(3, c, s->hash.nbuckets);
                }
    }
    qsort(keys, (size_t)cnt, sizeof(int64_t), cmp_i64);
    for (int i = 0; i < cnt; i++) {
        if (i && keys[i] == keys[i - 1]) continue;
        for (int64_t t = s->hash.start[keys[i]]; t < s->hash.start[keys[i] + 1];
             t++) {
            const int32_t j = s->hash.items[t];
            double d[3] = {s->x[3 * j] - q[0], s->x[3 * j + 1] - q[1],
                           s->x[3 * j + 2] - q[2]};
            if (sqnorm(dim, d) < s->k.kr2) cb(ctx, j, s->x + 3 * j);
        }
    }
}

static void set_reset(or_set* s) {
    if (s->search_mode == OR_SEARCH_HASH)
        hash_reset(&s->hash, s->k.dim, s->k.kr, s->n, s->x);
}

/* ---- sums over neighbours (SmoothedParticles.cpp:25-72,
 *      ParticlesBasedVectorField.cpp:5-21) ------------------------------- */
typedef struct {
    const or_set* s;
    const double* q;
    const or_kern* k; /* kernel used in the sum (may differ from s->k) */
    double acc;
    double vec[3];
    const double* field; /* vector field on s (3 per particle) */
    const double* other;
} sum_ctx;

static void cb_weight(void* c, int32_t j, const double* xj) {
    sum_ctx* t = (sum_ctx*)c;
    double d[3] = {t->q[0] - xj[0], t->q[1] - xj[1], t->q[2] - xj[2]};
    (void)j;
    t->acc += kern_Wv(t->k, d);
}
static void cb_bias_weight(void* c, int32_t j, const double* xj) {
    sum_ctx* t = (sum_ctx*)c;
    double d[3] = {t->q[0] - xj[0], t->q[1] - xj[1], t->q[2] - xj[2]};
    t->acc += t->s->vol[j] * kern_Wv(t->k, d);
}
static void cb_bias_grad(void* c, int32_t j, const double* xj) {
    sum_ctx* t = (sum_ctx*)c;
    double d[3] = {t->q[0] - xj[0], t->q[1] - xj[1], t->q[2] - xj[2]};
    double g[3];
    kern_grad(t->k, d, g);
    for (int a = 0; a < t->k->dim; a++) t->vec[a] += t->s->vol[j] * g[a];
}
/* ParticlesBasedVectorField::operator() (ParticlesBasedVectorField.cpp:5-12) */
static void cb_field_value(void* c, int32_t j, const double* xj) {
    sum_ctx* t = (sum_ctx*)c;
    double d[3] = {t->q[0] - xj[0], t->q[1] - xj[1], t->q[2] - xj[2]};
    const double w = kern_Wv(t->k, d);
    for (int a = 0; a < t->k->dim; a++)
        t->vec[a] += t->field[3 * j + a] * w / t->s->rho[j];
}
/* ParticlesBasedVectorField::divergence (ParticlesBasedVectorField.cpp:14-21) */
static void cb_field_div(void* c, int32_t j, const double* xj) {
    sum_ctx* t = (sum_ctx*)c;
    double d[3] = {t->q[0] - xj[0], t->q[1] - xj[1], t->q[2] - xj[2]};
    double g[3];
    kern_grad(t->k, d, g);
    t->acc += dot3(t->k->dim, t->field + 3 * j, g) * t->s->vol[j];
}

static double neighbor_weight(const or_set* s, const double* q) {
    sum_ctx c = {s, q, &s->k, 0, {0, 0, 0}, NULL, NULL};
    set_for_each(s, q, cb_weight, &c);
    return c.acc;
}
static double bias_weight(const or_set* s, const double* q) {
    sum_ctx c = {s, q, &s->k, 0, {0, 0, 0}, NULL, NULL};
    set_for_each(s, q, cb_bias_weight, &c);
    return c.acc;
}
static void bias_gradient(const or_set* s, const double* q, double* out) {
    sum_ctx c = {s, q, &s->k, 0, {0, 0, 0}, NULL, NULL};
    set_for_each(s, q, cb_bias_grad, &c);
    out[0] = c.vec[0];
    out[1] = c.vec[1];
    out[2] = c.vec[2];
}
static void field_value(const or_set* s, const double* field, const double* q,
                        double* out) {
    sum_ctx c = {s, q, &s->k, 0, {0, 0, 0}, field, NULL};
    set_for_each(s, q, cb_field_value, &c);
    for (int a = 0; a < 3; a++) out[a] = c.vec[a] * s->mass;
}
static double field_divergence(const or_set* s, const double* field,
                               const double* q) {
    sum_ctx c = {s, q, &s->k, 0, {0, 0, 0}, field, NULL};
    set_for_each(s, q, cb_field_div, &c);
    return c.acc;
}

/* SmoothedParticles::computeDensities / computeVolumes
 * (SmoothedParticles.cpp:25-36) */
static void set_compute_densities(or_set* s) {
    int64_t i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < s->n; i++)
        s->rho[i] = s->mass * neighbor_weight(s, s->x + 3 * i);
}
static void set_compute_volumes(or_set* s) {
    int64_t i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < s->n; i++) s->vol[i] = s->mass / s->rho[i];
}

static int set_resize(or_set* s, int64_t n) {
    s->n = n;
    s->x = (double*)realloc(s->x, (size_t)(n ? n : 1) * 3 * sizeof(double));
    s->rho = (double*)realloc(s->rho, (size_t)(n ? n : 1) * sizeof(double));
    s->vol = (double*)realloc(s->vol, (size_t)(n ? n : 1) * sizeof(double));
    return s->x && s->rho && s->vol ? 0 : VPS_ERR_OOM;
}
static void set_free(or_set* s) {
    free(s->x);
    free(s->rho);
    free(s->vol);
    hash_free(&s->hash);
}

/* ------------------------------------------------------------------------
 * Simulation state: DualParticleSphLiquid + VirtualParticle.
 * ---------------------------------------------------------------------- */
struct or_sim {
    vps_params p;
    int dim;
    or_set real, bnd, virt;
    double* v;        /* real velocities n*3 */
    double* bv;       /* boundary velocities */
    double* bn;       /* boundary normals */
    double rho0, kappa, alpha0;
    int initialized;
    /* virtual state */
    int64_t nnodes;
    double* w;        /* node weights */
    int64_t* node;    /* virtual -> linear node index */
    double* vbar;     /* n_virtual*3 */
    double* div;
    double* pres;
    int64_t npres;    /* current length of the pressure array */
    /* node-keyed history of the last solves (nnodes each, or NULL):
     * pressure p1, increments dp1 = p1 - p2 and dp2 = p2 - p3, and the depth
     * h (0: node absent from the last solve, else how many are known, <= 3) */
    double* x0;          /* CG initial guess of the last project (npres) */
    int start_fallback;  /* 1: the last guarded start gave way to the carried one */
    double* prev_node_p;
    double* prev_node_dp1;
    double* prev_node_dp2;
    unsigned char* prev_node_h;
    /* assembled operator M (CSR) */
    int64_t* rowptr;
    int32_t* col;
    double* val;
    double* diag;
};

static void lattice_pos(const or_sim* s, int64_t lin, double* out) {
    /* Grid::dataPosition: origin + coord * spacing (Grid.h:51); linear index
     * x fastest (Grid.h:54-58) */
    const int32_t nx = s->p.lattice_dims[0], ny = s->p.lattice_dims[1];
    const int64_t cx = lin % nx, cy = (lin / nx) % ny, cz = lin / nx / ny;
    const double sp = s->p.lattice_spacing;
    out[0] = s->p.lattice_origin[0] + (double)cx * sp;
    out[1] = s->p.lattice_origin[1] + (double)cy * sp;
    out[2] = s->dim == 3 ? s->p.lattice_origin[2] + (double)cz * sp : 0.0;
}

or_sim* or_create(const vps_params* p, int32_t search_mode, int32_t threads) {
    if (!p || (p->dim != 2 && p->dim != 3)) return NULL;
    or_sim* s = (or_sim*)calloc(1, sizeof(or_sim));
    s->p = *p;
    s->dim = p->dim;
    if (s->dim == 2) s->p.lattice_dims[2] = 1;
    kern_init(&s->real.k, s->dim, p->particle_radius);
    kern_init(&s->bnd.k, s->dim, p->particle_radius);
    kern_init(&s->virt.k, s->dim, p->lattice_spacing);
    s->real.mass = p->mass;
    s->bnd.mass = p->mass; /* DualParticleSphLiquid.cpp:84 */
    s->virt.mass = 1;      /* SmoothedParticles default; unused */
    s->real.search_mode = s->bnd.search_mode = s->virt.search_mode =
        search_mode;
    s->nnodes = (int64_t)s->p.lattice_dims[0] * s->p.lattice_dims[1] *
                s->p.lattice_dims[2];
    s->alpha0 = -1e6; /* VirtualParticle ctor (VirtualParticle.cpp:16) */
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    return s;
}

static void free_matrix(or_sim* s) {
    free(s->rowptr);
    free(s->col);
    free(s->val);
    free(s->diag);
    s->rowptr = NULL;
    s->col = NULL;
    s->val = NULL;
    s->diag = NULL;
}

void or_destroy(or_sim* s) {
    if (!s) return;
    set_free(&s->real);
    set_free(&s->bnd);
    set_free(&s->virt);
    free(s->v);
    free(s->bv);
    free(s->bn);
    free(s->w);
    free(s->node);
    free(s->vbar);
    free(s->div);
    free(s->pres);
    free(s->x0);
    free(s->prev_node_p);
    free(s->prev_node_dp1);
    free(s->prev_node_dp2);
    free(s->prev_node_h);
    free_matrix(s);
    free(s);
}

static void copy_in(int dim, int64_t n, const double* src, double* dst) {
    for (int64_t i = 0; i < n; i++) {
        dst[3 * i] = src[dim * i];
        dst[3 * i + 1] = src[dim * i + 1];
        dst[3 * i + 2] = dim == 3 ? src[dim * i + 2] : 0.0;
    }
}
static void copy_out(int dim, int64_t n, const double* src, double* dst) {
    for (int64_t i = 0; i < n; i++)
        for (int a = 0; a < dim; a++) dst[dim * i + a] = src[3 * i + a];
}

int or_set_boundary(or_sim* s, int64_t n, const double* x, const double* nrm,
                    const double* v) {
    if (n < 0 || (n > 0 && (!x || !nrm))) return VPS_ERR_INVALID_ARG;
    if (set_resize(&s->bnd, n)) return VPS_ERR_OOM;
    s->bv = (double*)realloc(s->bv, (size_t)(n ? n : 1) * 3 * sizeof(double));
    s->bn = (double*)realloc(s->bn, (size_t)(n ? n : 1) * 3 * sizeof(double));
    copy_in(s->dim, n, x, s->bnd.x);
    copy_in(s->dim, n, nrm, s->bn);
    if (v) copy_in(s->dim, n, v, s->bv);
    else memset(s->bv, 0, (size_t)(n ? n : 1) * 3 * sizeof(double));
    return 0;
}

/* BoundaryParticles::computeInfo after the boundary set changed
 * (DualParticleSphLiquid.cpp:88-89 runs it once for static walls) */
int or_refresh_boundary(or_sim* s) {
    set_reset(&s->bnd);
    set_compute_densities(&s->bnd);
    set_compute_volumes(&s->bnd);
    return 0;
}

int or_set_particles(or_sim* s, int64_t n, const double* x, const double* v) {
    if (n <= 0 || !x) return VPS_ERR_INVALID_ARG;
    if (set_resize(&s->real, n)) return VPS_ERR_OOM;
    s->v = (double*)realloc(s->v, (size_t)n * 3 * sizeof(double));
    copy_in(s->dim, n, x, s->real.x);
    if (v) copy_in(s->dim, n, v, s->v);
    else memset(s->v, 0, (size_t)n * 3 * sizeof(double));
    return 0;
}

int or_get_particles(or_sim* s, double* x, double* v) {
    if (x) copy_out(s->dim, s->real.n, s->real.x, x);
    if (v) copy_out(s->dim, s->real.n, s->v, v);
    return 0;
}


/* ---- VirtualParticle::generateParticles (VirtualParticle.cpp:45-62) ---- */
static void cb_vol_sq(void* c, int32_t j, const double* xj) {
    /* volumes[I] += V_j * V_j * real.kernel(p_I - p_j) (VirtualParticle.cpp:34) */
    sum_ctx* t = (sum_ctx*)c;
    double d[3] = {t->q[0] - xj[0], t->q[1] - xj[1], t->q[2] - xj[2]};
    t->acc += t->s->vol[j] * t->s->vol[j] * kern_Wv(t->k, d);
}

int or_generate(or_sim* s) {
    const int64_t nn = s->nnodes;
    if (s->real.n == 0) return VPS_ERR_STATE;
    if (!s->w) s->w = (double*)malloc((size_t)nn * sizeof(double));
    if (!s->w) return VPS_ERR_OOM;
    /* _weight.parallelForEach: w_I = sum_j real.kernel(p_I - p_j) over real
     * neighbours within h; OpenMP over z slices (y rows in 2D),
     * Grid.cpp:246-266 */
    const int64_t nx = s->p.lattice_dims[0], ny = s->p.lattice_dims[1],
                  nz = s->p.lattice_dims[2];
    const int64_t outer = s->dim == 3 ? nz : ny;
    const int64_t inner = s->dim == 3 ? nx * ny : nx;
    int64_t o;
#pragma omp parallel for schedule(static)
    for (o = 0; o < outer; o++) {
        for (int64_t t = 0; t < inner; t++) {
            const int64_t lin = o * inner + t;
            double q[3];
            lattice_pos(s, lin, q);
            sum_ctx c = {&s->real, q, &s->real.k, 0, {0, 0, 0}, NULL, NULL};
            set_for_each(&s->real, q, cb_weight, &c);
            s->w[lin] = c.acc;
        }
    }
    /* _weight.forEach: add(position) when weight > 1e-6, linear order */
    int64_t cnt = 0;
    for (int64_t lin = 0; lin < nn; lin++)
        if (s->w[lin] > 1e-6) cnt++;
    if (set_resize(&s->virt, cnt)) return VPS_ERR_OOM;
    s->node = (int64_t*)realloc(s->node, (size_t)(cnt ? cnt : 1) * sizeof(int64_t));
    cnt = 0;
    for (int64_t lin = 0; lin < nn; lin++)
        if (s->w[lin] > 1e-6) {
            lattice_pos(s, lin, s->virt.x + 3 * cnt);
            s->node[cnt++] = lin;
        }
    (void)nx; (void)nz;
    /* resetNearbySearcher(); computeInfo(realParticles) */
    set_reset(&s->virt);
    const or_set* r = &s->real;
    int64_t I;
    /* VirtualParticle::computeDensities (VirtualParticle.cpp:18-27): the
     * virtual kernel (H) over real neighbours found at h, times real mass */
#pragma omp parallel for schedule(static)
    for (I = 0; I < cnt; I++) {
        sum_ctx c = {r, s->virt.x + 3 * I, &s->virt.k, 0, {0, 0, 0}, NULL, NULL};
        set_for_each(r, s->virt.x + 3 * I, cb_weight, &c);
        s->virt.rho[I] = c.acc * r->mass;
    }
    /* VirtualParticle::computeVolumes (VirtualParticle.cpp:29-38) */
#pragma omp parallel for schedule(static)
    for (I = 0; I < cnt; I++) {
        sum_ctx c = {r, s->virt.x + 3 * I, &r->k, 0, {0, 0, 0}, NULL, NULL};
        set_for_each(r, s->virt.x + 3 * I, cb_vol_sq, &c);
        s->virt.vol[I] = c.acc;
    }
    return 0;
}

/* ---- VirtualParticle::calculateCoef (VirtualParticle.cpp:149-169) ------ *
 * Rows assembled serially in the searcher's visit order, diagonal triplet
 * appended; setFromTriplets sums the duplicate (I,I) pair (-0 + diag).  The
 * CG operator M is the transpose (Eigen's TransposeInput for a column-major
 * matrix with Lower|Upper) or A itself, stored as CSR with ascending
 * columns: for M = A^T that is column I of A with ascending rows, the order
 * in which Eigen's row-major sparse*dense product accumulates. */
typedef struct {
    int64_t cap, n;
    int32_t* col;
    double* val;
} row_buf;

typedef struct {
    const or_sim* s;
    const double* q;
    double sum;
    row_buf* rb;
} coef_ctx;

static void cb_coef(void* c, int32_t J, const double* pJ) {
    coef_ctx* t = (coef_ctx*)c;
    double d[3] = {t->q[0] - pJ[0], t->q[1] - pJ[1], t->q[2] - pJ[2]};
    const double r_ij = sqrt(sqnorm(t->s->dim, d));
    const or_kern* kk = g_mutation == OR_MUT_COEF_KERNEL_H ? &t->s->real.k : &t->s->virt.k;
    const double two = g_mutation == OR_MUT_COEF_HALF ? 1. : (g_mutation == OR_MUT_COEF_SIGN ? -2. : 2.);
    const double alpha_ij = two * t->s->virt.vol[J] * kern_dW(kk, r_ij) / (r_ij + 1e-6);
    t->sum -= alpha_ij;
    row_buf* rb = t->rb;
    if (rb->n == rb->cap) {
        rb->cap = rb->cap ? 2 * rb->cap : 1024;
        rb->col = (int32_t*)realloc(rb->col, (size_t)rb->cap * sizeof(int32_t));
        rb->val = (double*)realloc(rb->val, (size_t)rb->cap * sizeof(double));
    }
    rb->col[rb->n] = J;
    rb->val[rb->n] = alpha_ij;
    rb->n++;
}

static int calculate_coef(or_sim* s) {
    const int64_t n = s->virt.n;
    free_matrix(s);
    /* A as triplet rows (row-major CSR in visit order incl. the diagonal) */
    int64_t* arow = (int64_t*)malloc((size_t)(n + 1) * sizeof(int64_t));
    row_buf rb = {0, 0, NULL, NULL};
    arow[0] = 0;
    for (int64_t I = 0; I < n; I++) {
        coef_ctx c = {s, s->virt.x + 3 * I, 0, &rb};
        set_for_each(&s->virt, s->virt.x + 3 * I, cb_coef, &c);
        const double d = c.sum < s->alpha0 ? s->alpha0 : c.sum; /* std::max */
        if (rb.n == rb.cap) {
            rb.cap = rb.cap ? 2 * rb.cap : 1024;
            rb.col = (int32_t*)realloc(rb.col, (size_t)rb.cap * sizeof(int32_t));
            rb.val = (double*)realloc(rb.val, (size_t)rb.cap * sizeof(double));
        }
        rb.col[rb.n] = (int32_t)I;
        rb.val[rb.n] = d;
        rb.n++;
        arow[I + 1] = rb.n;
    }
    const int64_t nnz = rb.n;
    /* M rows with duplicates still present */
    int64_t* mrow = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    int32_t* mcol = (int32_t*)malloc((size_t)(nnz ? nnz : 1) * sizeof(int32_t));
    double* mval = (double*)malloc((size_t)(nnz ? nnz : 1) * sizeof(double));
    if (s->p.op_mode == VPS_OP_TRANSPOSE) {
        for (int64_t t = 0; t < nnz; t++) mrow[rb.col[t] + 1]++;
        for (int64_t I = 0; I < n; I++) mrow[I + 1] += mrow[I];
        int64_t* fill = (int64_t*)malloc((size_t)(n ? n : 1) * sizeof(int64_t));
        memcpy(fill, mrow, (size_t)n * sizeof(int64_t));
        for (int64_t I = 0; I < n; I++)
            for (int64_t t = arow[I]; t < arow[I + 1]; t++) {
                const int64_t dst = fill[rb.col[t]]++;
                mcol[dst] = (int32_t)I;
                mval[dst] = rb.val[t];
            }
        free(fill);
    } else {
        memcpy(mrow, arow, (size_t)(n + 1) * sizeof(int64_t));
        for (int64_t I = 0; I < n; I++) {
            /* stable insertion sort of row I by column */
            for (int64_t t = arow[I] + 1; t < arow[I + 1]; t++) {
                const int32_t c = rb.col[t];
                const double v = rb.val[t];
                int64_t u = t - 1;
                while (u >= arow[I] && rb.col[u] > c) {
                    rb.col[u + 1] = rb.col[u];
                    rb.val[u + 1] = rb.val[u];
                    u--;
                }
                rb.col[u + 1] = c;
                rb.val[u + 1] = v;
            }
        }
        memcpy(mcol, rb.col, (size_t)nnz * sizeof(int32_t));
        memcpy(mval, rb.val, (size_t)nnz * sizeof(double));
    }
    free(rb.col);
    free(rb.val);
    free(arow);
    /* collapse duplicates (first += later, SparseMatrix::collapseDuplicates) */
    s->rowptr = (int64_t*)malloc((size_t)(n + 1) * sizeof(int64_t));
    s->col = (int32_t*)malloc((size_t)(nnz ? nnz : 1) * sizeof(int32_t));
    s->val = (double*)malloc((size_t)(nnz ? nnz : 1) * sizeof(double));
    s->diag = (double*)malloc((size_t)(n ? n : 1) * sizeof(double));
    int64_t w = 0;
    s->rowptr[0] = 0;
    for (int64_t I = 0; I < n; I++) {
        const int64_t row_begin = w;
        for (int64_t t = mrow[I]; t < mrow[I + 1]; t++) {
            if (w > row_begin && s->col[w - 1] == mcol[t]) {
                s->val[w - 1] += mval[t];
            } else {
                s->col[w] = mcol[t];
                s->val[w] = mval[t];
                w++;
            }
        }
        s->rowptr[I + 1] = w;
        s->diag[I] = 0;
        for (int64_t t = row_begin; t < w; t++)
            if (s->col[t] == I) s->diag[I] = s->val[t];
    }
    free(mrow);
    free(mcol);
    free(mval);
    return 0;
}

/* ---- Eigen ConjugateGradient (IterativeSolver.h:19-40 and Eigen's
 * internal::conjugate_gradient, identity preconditioner) ----------------- */
static void csr_mv(int64_t n, const int64_t* rp, const int32_t* col,
                   const double* val, const double* x, double* y) {
    int64_t i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; i++) {
        double tmp = 0;
        for (int64_t t = rp[i]; t < rp[i + 1]; t++) tmp += val[t] * x[col[t]];
        y[i] = tmp;
    }
}
static double vdot(int64_t n, const double* a, const double* b) {
    double s = 0;
    for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

int or_cg_csr(int64_t n, const int64_t* rp, const int32_t* col,
              const double* val, const double* b, double* x, double tol_in,
              int64_t max_iter, int64_t* iters, double* err) {
    /* solver.setTolerance(max(tolerance, eps / b.norm())) */
    const double bnorm = sqrt(vdot(n, b, b));
    const double tol = tol_in > DBL_EPSILON / bnorm ? tol_in : DBL_EPSILON / bnorm;
    const int64_t max_iters = max_iter < 0 ? 2 * n : max_iter;
    double* r = (double*)malloc((size_t)(n ? n : 1) * sizeof(double));
    double* p = (double*)malloc((size_t)(n ? n : 1) * sizeof(double));
    double* t = (double*)malloc((size_t)(n ? n : 1) * sizeof(double));
    int status = 0;
    csr_mv(n, rp, col, val, x, t);
    for (int64_t i = 0; i < n; i++) r[i] = b[i] - t[i];
    const double rhs2 = vdot(n, b, b);
    if (rhs2 == 0) {
        for (int64_t i = 0; i < n; i++) x[i] = 0;
        *iters = 0;
        *err = 0;
        goto done;
    }
    {
        double threshold = tol * tol * rhs2;
        if (threshold < DBL_MIN) threshold = DBL_MIN;
        double r2 = vdot(n, r, r);
        if (r2 < threshold) {
            *iters = 0;
            *err = sqrt(r2 / rhs2);
            goto done;
        }
        memcpy(p, r, (size_t)n * sizeof(double));
        double abs_new = vdot(n, r, p);
        int64_t it = 0;
        while (it < max_iters) {
            csr_mv(n, rp, col, val, p, t);
            const double alpha = abs_new / vdot(n, p, t);
            for (int64_t i = 0; i < n; i++) x[i] += alpha * p[i];
            for (int64_t i = 0; i < n; i++) r[i] -= alpha * t[i];
            r2 = vdot(n, r, r);
            if (r2 < threshold) break;
            const double abs_old = abs_new;
            abs_new = vdot(n, r, r);
            const double beta = abs_new / abs_old;
            for (int64_t i = 0; i < n; i++) p[i] = r[i] + beta * p[i];
            it++;
        }
        *err = sqrt(r2 / rhs2);
        *iters = it;
        status = *err <= tol ? 0 : VPS_ERR_NO_CONVERGENCE;
    }
done:
    free(r);
    free(p);
    free(t);
    return status;
}

/* ---- DualParticleSphLiquid::initialize (DualParticleSphLiquid.cpp:78-110),
 * boundary sampling left to the caller --------------------------------- */
int or_initialize(or_sim* s) {
    if (s->real.n == 0) return VPS_ERR_STATE;
    if (s->bnd.n == 0) return VPS_ERR_INVALID_ARG; /* empty set: % 0 in ref */
    set_reset(&s->real);
    set_compute_densities(&s->real);
    set_compute_volumes(&s->real);
    if (s->p.rest_density_mode == VPS_FROM_INIT) {
        double m = s->real.rho[0]; /* ParticlesAttribute::max */
        for (int64_t i = 1; i < s->real.n; i++)
            if (m < s->real.rho[i]) m = s->real.rho[i];
        s->rho0 = m;
    } else {
        s->rho0 = s->p.rest_density;
    }
    set_reset(&s->bnd);
    set_compute_densities(&s->bnd);
    set_compute_volumes(&s->bnd);
    /* VirtualParticle::setAlpha (VirtualParticle.cpp:64-73) */
    s->alpha0 = -1e6;
    int rc = or_generate(s);
    if (rc) return rc;
    rc = calculate_coef(s);
    if (rc) return rc;
    double a = s->virt.n ? s->diag[0] : 0;
    for (int64_t I = 1; I < s->virt.n; I++)
        if (a < s->diag[I]) a = s->diag[I];
    s->alpha0 = a;
    /* setKappa(1 / packed kernel sum of the real particles) (:98) */
    s->kappa = s->p.kappa_mode == VPS_FROM_INIT ? 1. / packed_sum(&s->real.k)
                                                : s->p.kappa;
    /* pressures start empty (zero-filled on the first resize) */
    s->npres = 0;
    free(s->prev_node_p);
    free(s->prev_node_dp1);
    free(s->prev_node_dp2);
    free(s->prev_node_h);
    s->prev_node_p = s->prev_node_dp1 = s->prev_node_dp2 = NULL;
    s->prev_node_h = NULL;
    s->initialized = 1;
    return 0;
}

/* SmthParticleHydrodLiquid::moveParticles lines 80-81 */
int or_refresh_real(or_sim* s) {
    if (!s->initialized) return VPS_ERR_STATE;
    set_reset(&s->real);
    set_compute_densities(&s->real);
    return 0;
}

/* force += volumes[J] * _pressures[J] * gradientKernel(p_i - p_J)
 * (VirtualParticle.cpp:118-121), virtual kernel H */
typedef struct {
    const or_sim* s;
    const double* q;
    double f[3];
} grad_virtual_ctx;
static void cb_grad_virtual(void* c, int32_t J, const double* pJ) {
    grad_virtual_ctx* t = (grad_virtual_ctx*)c;
    const or_sim* s = t->s;
    double d[3] = {t->q[0] - pJ[0], t->q[1] - pJ[1], t->q[2] - pJ[2]};
    double g[3];
    kern_grad(g_mutation == OR_MUT_GRAD_KERNEL_H ? &s->real.k : &s->virt.k, d, g);
    double sc = s->virt.vol[J] * s->pres[J];
    if (g_mutation == OR_MUT_GRAD_SIGN) sc = -sc;
    else if (g_mutation == OR_MUT_GRAD_NO_VOL) sc = s->pres[J];
    for (int a = 0; a < s->dim; a++) t->f[a] += sc * g[a];
}
/* force -= V_b * min(n_b.(v_i - v_b), 0) * n_b * kernel_b(p_i - p_b)
 * (VirtualParticle.cpp:124-128); v_i is the pre-update velocity */
typedef struct {
    const or_sim* s;
    const double* q;
    const double* vi;
    double f[3];
} grad_boundary_ctx;
static void cb_grad_boundary(void* c, int32_t b, const double* pb) {
    grad_boundary_ctx* t = (grad_boundary_ctx*)c;
    const or_sim* s = t->s;
    double d[3] = {t->q[0] - pb[0], t->q[1] - pb[1], t->q[2] - pb[2]};
    double dv[3] = {t->vi[0] - s->bv[3 * b], t->vi[1] - s->bv[3 * b + 1],
                    t->vi[2] - s->bv[3 * b + 2]};
    const double* nb = s->bn + 3 * b;
    double ng[3];
    if (nb[0] == 0. && nb[1] == 0. && nb[2] == 0.) {
        /* a grain (zero normal): the unit vector from the grain to the
         * particle; the reference has no grain normal */
        const double r = sqrt(dot3(s->dim, d, d));
        for (int a = 0; a < 3; a++) ng[a] = r > 0. ? d[a] / r : 0.;
        nb = ng;
    }
    const double dn = dot3(s->dim, nb, dv);
    const double m = 0. < dn ? 0. : dn; /* std::min(dn, 0.) */
    const double sc = s->bnd.vol[b] * m;
    const double w = kern_Wv(&s->bnd.k, d);
    for (int a = 0; a < s->dim; a++) t->f[a] -= sc * nb[a] * w;
}

/* applyPressureGradient (VirtualParticle.cpp:110-134) */
static void pressure_gradient(or_sim* s) {
    const int dim = s->dim;
    const or_set* vs = &s->virt;
    const or_set* bs = &s->bnd;
    int64_t i;
    double* vnew = (double*)malloc((size_t)s->real.n * 3 * sizeof(double));
#pragma omp parallel for schedule(static)
    for (i = 0; i < s->real.n; i++) {
        const double* pi = s->real.x + 3 * i;
        double force[3] = {0, 0, 0};
        grad_virtual_ctx cv = {s, pi, {0, 0, 0}};
        set_for_each(vs, pi, cb_grad_virtual, &cv);
        grad_boundary_ctx cb = {s, pi, s->v + 3 * i, {cv.f[0], cv.f[1], cv.f[2]}};
        set_for_each(bs, pi, cb_grad_boundary, &cb);
        force[0] = cb.f[0];
        force[1] = cb.f[1];
        force[2] = cb.f[2];
        for (int a = 0; a < 3; a++) vnew[3 * i + a] = s->v[3 * i + a] + force[a];
        if (dim == 2) vnew[3 * i + 2] = 0;
    }
    memcpy(s->v, vnew, (size_t)s->real.n * 3 * sizeof(double));
    free(vnew);
}

int or_pressure_gradient(or_sim* s) {
    if (!s->initialized || !s->pres || s->npres != s->virt.n) return VPS_ERR_STATE;
    pressure_gradient(s);
    return 0;
}

/* The learned start of the HIP library with an untrained network (no
 * reference counterpart): the Galerkin projection onto the span of
 * P = p1 and E = e (the extrapolation's increment), x0 = y0 P + y1 E with
 * (V^T M V) y = V^T b, V = [P E], over the fields with v . M v > 0 and by
 * elimination with partial pivoting (a pivot not above 1e-10 of its
 * diagonal keeps the extrapolation P + E).  Nodes without a history stay
 * zero; s->pres holds the extrapolation on entry. */
static void learned_start(or_sim* s, int64_t n) {
    double* V[2];
    double* MV[2];
    for (int k = 0; k < 2; k++) {
        V[k] = (double*)calloc((size_t)(n ? n : 1), sizeof(double));
        MV[k] = (double*)malloc((size_t)(n ? n : 1) * sizeof(double));
    }
    for (int64_t I = 0; I < n; I++) {
        const int64_t nd = s->node[I];
        const int h = s->prev_node_h[nd];
        if (h >= 1) {
            V[0][I] = s->prev_node_p[nd];
            V[1][I] = h >= 3 ? 2 * s->prev_node_dp1[nd] - s->prev_node_dp2[nd]
                             : (h == 2 ? s->prev_node_dp1[nd] : 0.);
        }
    }
    for (int k = 0; k < 2; k++) csr_mv(n, s->rowptr, s->col, s->val, V[k], MV[k]);
    double t[2][2] = {{0, 0}, {0, 0}}, rb[2] = {0, 0};
    for (int64_t I = 0; I < n; I++)
        for (int p = 0; p < 2; p++) {
            for (int q = 0; q < 2; q++) t[p][q] += V[p][I] * MV[q][I];
            rb[p] += V[p][I] * s->div[I];
        }
    int idx[2], m = 0;
    for (int i = 0; i < 2; i++)
        if (t[i][i] > 0 && isfinite(t[i][i])) idx[m++] = i;
    double G[2][2], r[2], y[2] = {1, 1};
    for (int p = 0; p < m; p++) {
        r[p] = rb[idx[p]];
        for (int q = 0; q < m; q++) G[p][q] = t[idx[p]][idx[q]];
    }
    int ok = m > 0;
    double dg[2] = {m > 0 ? fabs(G[0][0]) : 0, m > 1 ? fabs(G[1][1]) : 0};
    if (m == 2 && fabs(G[1][0]) > fabs(G[0][0])) {   /* partial pivoting */
        for (int q = 0; q < 2; q++) {
            const double tmp = G[0][q];
            G[0][q] = G[1][q];
            G[1][q] = tmp;
        }
        const double tr = r[0];
        r[0] = r[1];
        r[1] = tr;
    }
    for (int k = 0; k < m && ok; k++) {
        if (!(fabs(G[k][k]) > 1e-10 * dg[k])) {
            ok = 0;
            break;
        }
        for (int p = k + 1; p < m; p++) {
            const double f = G[p][k] / G[k][k];
            for (int q = k; q < m; q++) G[p][q] -= f * G[k][q];
            r[p] -= f * r[k];
        }
    }
    if (ok) {
        double x[2];
        for (int p = m - 1; p >= 0; p--) {
            double sum = r[p];
            for (int q = p + 1; q < m; q++) sum -= G[p][q] * x[q];
            x[p] = sum / G[p][p];
        }
        int fin = 1;
        for (int p = 0; p < m; p++) fin = fin && isfinite(x[p]);
        if (fin) {
            y[0] = y[1] = 0;
            for (int p = 0; p < m; p++) y[idx[p]] = x[p];
        }
    }
    if (ok)
        for (int64_t I = 0; I < n; I++) s->pres[I] = y[0] * V[0][I] + y[1] * V[1][I];
    memcpy(s->x0, s->pres, (size_t)n * sizeof(double));
    for (int k = 0; k < 2; k++) {
        free(V[k]);
        free(MV[k]);
    }
}

/* ---- VirtualParticle::project (VirtualParticle.cpp:75-134) ------------- */
int or_project(or_sim* s, or_stats* st) {
    if (!s->initialized) return VPS_ERR_STATE;
    const int dim = s->dim;
    const int64_t n = s->virt.n;
    double t0 = now_s();
    /* _pressures.resize(this): keep the first min(old,new), zero the rest
     * (ParticlesBasedData.h:25-29); other warm starts are options */
    double* p_new = (double*)calloc((size_t)(n ? n : 1), sizeof(double));
    double* p_carry = (double*)calloc((size_t)(n ? n : 1), sizeof(double));
    {
        const int64_t keep = s->npres < n ? s->npres : n;
        if (s->pres) memcpy(p_carry, s->pres, (size_t)keep * sizeof(double));
    }
    if (s->p.warm_start == VPS_WARM_INDEX_CARRY) {
        memcpy(p_new, p_carry, (size_t)n * sizeof(double));
    } else if (s->p.warm_start == VPS_WARM_NODE_KEYED) {
        if (s->prev_node_p)
            for (int64_t I = 0; I < n; I++)
                if (s->prev_node_h[s->node[I]])
                    p_new[I] = s->prev_node_p[s->node[I]];
    } else if (s->p.warm_start == VPS_WARM_EXTRAPOLATE ||
               s->p.warm_start == VPS_WARM_LEARNED) {
        /* polynomial extrapolation through the known history; the learned
         * start (of the HIP library with an untrained network) is formed
         * from it once the matrix exists, below */
        if (s->prev_node_p)
            for (int64_t I = 0; I < n; I++) {
                const int64_t nd = s->node[I];
                const int h = s->prev_node_h[nd];
                if (h >= 1)
                    p_new[I] = s->prev_node_p[nd] +
                               (h >= 3 ? 2 * s->prev_node_dp1[nd] - s->prev_node_dp2[nd]
                                       : (h == 2 ? s->prev_node_dp1[nd] : 0.));
            }
    }
    free(s->pres);
    s->pres = p_new;
    s->npres = n;
    s->x0 = (double*)realloc(s->x0, (size_t)(n ? n : 1) * sizeof(double));
    memcpy(s->x0, p_new, (size_t)n * sizeof(double));
    s->div = (double*)realloc(s->div, (size_t)(n ? n : 1) * sizeof(double));
    s->vbar = (double*)realloc(s->vbar, (size_t)(n ? n : 1) * 3 * sizeof(double));
    int64_t I;
    /* calculateVelocity (VirtualParticle.cpp:136-147) */
#pragma omp parallel for schedule(static)
    for (I = 0; I < n; I++) {
        const double* q = s->virt.x + 3 * I;
        double vr[3], vb[3];
        field_value(&s->real, s->v, q, vr);
        field_value(&s->bnd, s->bv, q, vb);
        const double den = bias_weight(&s->real, q) + bias_weight(&s->bnd, q) + 1e-6;
        for (int a = 0; a < 3; a++) s->vbar[3 * I + a] = (vr[a] + vb[a]) / den;
    }
    /* calculateDiv (VirtualParticle.cpp:171-201) */
#pragma omp parallel for schedule(static)
    for (I = 0; I < n; I++) {
        const double* q = s->virt.x + 3 * I;
        const double* vb = s->vbar + 3 * I;
        double gr[3], gb[3];
        bias_gradient(&s->real, q, gr);
        bias_gradient(&s->bnd, q, gb);
        const double drho = s->virt.rho[I] - s->rho0;
        s->div[I] = field_divergence(&s->real, s->v, q) - dot3(dim, vb, gr) +
                    field_divergence(&s->bnd, s->bv, q) - dot3(dim, vb, gb) -
                    s->kappa * (drho < 0. ? 0. : drho) / s->rho0;
    }
    int rc = calculate_coef(s);
    if (rc) {
        free(p_carry);
        return rc;
    }
    if (s->p.warm_start == VPS_WARM_LEARNED && s->prev_node_p) learned_start(s, n);
    /* guarded starts (extrapolated or learned, no reference counterpart):
     * one whose residual is not finite or above ||b|| gives way to the
     * index-carried start, as in the HIP library */
    s->start_fallback = 0;
    if (s->p.warm_start == VPS_WARM_EXTRAPOLATE || s->p.warm_start == VPS_WARM_LEARNED) {
        double* t = (double*)malloc((size_t)(n ? n : 1) * sizeof(double));
        csr_mv(n, s->rowptr, s->col, s->val, s->pres, t);
        double rr = 0, bb = 0;
        for (I = 0; I < n; I++) {
            const double r = s->div[I] - t[I];
            rr += r * r;
            bb += s->div[I] * s->div[I];
        }
        free(t);
        if (!(rr <= bb)) {
            memcpy(s->pres, p_carry, (size_t)n * sizeof(double));
            memcpy(s->x0, p_carry, (size_t)n * sizeof(double));
            s->start_fallback = 1;
        }
    }
    free(p_carry);
    double t1 = now_s();
    /* solveLinearSystem -> IterativeSolver::solve */
    int64_t iters = 0;
    double err = 0;
    rc = or_cg_csr(n, s->rowptr, s->col, s->val, s->div, s->pres,
                   s->p.tolerance, s->p.max_iterations, &iters, &err);
    double t2 = now_s();
    /* node-keyed history: this solve's pressures, their increments over
     * the previous solve where the node was in it, depth h + 1 (at most 3);
     * nodes absent from this solve lose their history */
    if (!s->prev_node_p) {
        s->prev_node_p = (double*)calloc((size_t)s->nnodes, sizeof(double));
        s->prev_node_dp1 = (double*)calloc((size_t)s->nnodes, sizeof(double));
        s->prev_node_dp2 = (double*)calloc((size_t)s->nnodes, sizeof(double));
        s->prev_node_h = (unsigned char*)calloc((size_t)s->nnodes, 1);
    }
    {
        unsigned char* hnew = (unsigned char*)calloc((size_t)s->nnodes, 1);
        for (I = 0; I < n; I++) {
            const int64_t nd = s->node[I];
            const int h = s->prev_node_h[nd];
            const double dp1 = h >= 1 ? s->pres[I] - s->prev_node_p[nd] : 0.;
            s->prev_node_dp2[nd] = h >= 2 ? s->prev_node_dp1[nd] : 0.;
            s->prev_node_dp1[nd] = dp1;
            s->prev_node_p[nd] = s->pres[I];
            hnew[nd] = (unsigned char)(h + 1 > 3 ? 3 : h + 1);
        }
        free(s->prev_node_h);
        s->prev_node_h = hnew;
    }
    pressure_gradient(s);
    double t3 = now_s();
    if (st) {
        st->n_virtual = n;
        st->cg_iterations = iters;
        st->cg_error = err;
        st->rhs_norm = sqrt(vdot(n, s->div, s->div));
        st->cg_status = rc;
        st->s_build = t1 - t0;
        st->s_solve = t2 - t1;
        st->s_gradient = t3 - t2;
    }
    return rc;
}

int or_projection_step(or_sim* s, or_stats* st) {
    double t0 = now_s();
    int rc = or_refresh_real(s);
    if (rc) return rc;
    double t1 = now_s();
    rc = or_generate(s);
    if (rc) return rc;
    double t2 = now_s();
    rc = or_project(s, st);
    double t3 = now_s();
    if (st) {
        st->s_refresh = t1 - t0;
        st->s_generate = t2 - t1;
        st->s_total = t3 - t0;
    }
    return rc;
}

/* x += v dt; v.y -= g dt (SmthParticleHydrodLiquid.cpp:73, 84-88) */
int or_advect(or_sim* s, double dt, double gravity) {
    int64_t i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < s->real.n; i++) {
        for (int a = 0; a < s->dim; a++) s->real.x[3 * i + a] += s->v[3 * i + a] * dt;
        s->v[3 * i + 1] -= gravity * dt;
    }
    return 0;
}

/* ------------------------------------------------------------------------
 * DEM sand step.  DEMParticleSand::advance (DEMParticleSand.cpp:90-94):
 *   moveParticles (:98-109)       x += v dt, rebuild the grains' searcher
 *   applyPressureForce (:118-152) v_i += dt * getForceSum(i) + wall terms
 *   applyExternalForces (:111-115) v_i.y -= g dt
 * The grains' searcher has kernel radius 2r (kernel_rate 2,
 * DEMParticle.cpp:17); the walls keep 4r.  Contact force ComputeDemForces
 * (DEMParticle.cpp:69-96) with K_norm = Young r, K_tang = K_norm Poisson
 * (:45-55) and tan of the friction angle (:58-60).  The capillary force
 * (:98-124) needs 1e-6 < |d| - 2r, which no neighbour within 2r has, so it
 * is zero on every pair the searcher yields and is left out.  The reference
 * adds each force to velocities[i] while other threads read velocities[j]
 * (a data race); here every force reads the velocities of the start of the
 * pass.
 * ---------------------------------------------------------------------- */
typedef struct {
    int dim;
    const double* q;
    const double* vi;
    const double* v;      /* start-of-pass velocities, 3 per grain */
    double two_r, min_dist, k_norm, k_tang, tan_phi;
    double f[3];
} dem_ctx;

static void cb_dem_contact(void* c, int32_t j, const double* xj) {
    dem_ctx* t = (dem_ctx*)c;
    double d[3], w[3];
    for (int a = 0; a < 3; a++) {
        d[a] = xj[a] - t->q[a];            /* positions[j] - positions[i] */
        w[a] = t->v[3 * j + a] - t->vi[a]; /* velocities[j] - velocities[i] */
    }
    double dist = sqrt(sqnorm(t->dim, d));
    const double pen = t->two_r - dist;
    if (!(pen > 0.)) return;
    if (dist <= t->min_dist) dist = t->min_dist;
    double n[3], vt[3], fn[3], fs[3];
    for (int a = 0; a < 3; a++) n[a] = d[a] * (1 / dist);
    const double vn = dot3(t->dim, w, n);
    for (int a = 0; a < 3; a++) {
        vt[a] = w[a] - vn * n[a];
        fn[a] = t->k_norm * pen * n[a];
        fs[a] = -t->k_tang * vt[a];
    }
    const double max_fs = sqrt(sqnorm(t->dim, fn)) * t->tan_phi;
    const double fs_norm = sqrt(sqnorm(t->dim, fs));
    if (fs_norm > max_fs)
        for (int a = 0; a < 3; a++) fs[a] = fs[a] * max_fs / fs_norm;
    for (int a = 0; a < t->dim; a++) t->f[a] += -fn[a] - fs[a];
}

typedef struct {
    const or_sim* s;
    const double* q;
    const double* vi;
    double k_tang, tan_phi;
    double f[3];
} dem_wall_ctx;

/* the boundary lambda of DEMParticleSand::applyPressureForce (:138-149) */
static void cb_dem_wall(void* c, int32_t js, const double* xj) {
    dem_wall_ctx* t = (dem_wall_ctx*)c;
    const int dim = t->s->dim;
    const double* nb = t->s->bn + 3 * js;
    double dv[3], d[3], Fn[3], Ft[3];
    for (int a = 0; a < 3; a++) {
        dv[a] = t->vi[a] - t->s->bv[3 * js + a];
        d[a] = t->q[a] - xj[a];
    }
    const double un = dot3(dim, nb, dv);
    const double w = kern_Wv(&t->s->bnd.k, d);
    const double m = un < 0. ? un : 0.;
    for (int a = 0; a < 3; a++) {
        Fn[a] = -t->s->bnd.vol[js] * m * nb[a] * w;
        Ft[a] = -t->k_tang * (dv[a] - un * nb[a]);
    }
    const double tn = sqrt(sqnorm(dim, Ft));
    if (tn <= 0.000001) {
        for (int a = 0; a < dim; a++) t->f[a] += Fn[a] + Ft[a];
    } else {
        const double lim = sqrt(sqnorm(dim, Fn)) * t->tan_phi;
        for (int a = 0; a < dim; a++)
            t->f[a] += Fn[a] + (lim > tn ? Ft[a] : (lim / tn) * Ft[a]);
    }
}

int or_dem_step(or_sim* s, double dt, double gravity, double young,
                double poisson, double friction_angle) {
    const int dim = s->dim;
    const int64_t n = s->real.n;
    if (n == 0) return VPS_ERR_STATE;
    int64_t i;
    /* moveParticles */
    for (i = 0; i < n; i++)
        for (int a = 0; a < dim; a++) s->real.x[3 * i + a] += s->v[3 * i + a] * dt;
    or_set grains = s->real; /* the same positions behind a 2r searcher */
    memset(&grains.hash, 0, sizeof(grains.hash));
    grains.k.kr = s->p.particle_radius * 2;
    grains.k.kr2 = grains.k.kr * grains.k.kr;
    set_reset(&grains);
    /* wall volumes: computeInfo of the boundary set (DEMParticleSand.cpp:84-85) */
    if (s->bnd.n > 0) {
        set_reset(&s->bnd);
        set_compute_densities(&s->bnd);
        set_compute_volumes(&s->bnd);
    }
    double* v0 = (double*)malloc((size_t)n * 3 * sizeof(double));
    if (!v0) {
        hash_free(&grains.hash);
        return VPS_ERR_OOM;
    }
    memcpy(v0, s->v, (size_t)n * 3 * sizeof(double));
    const double r = s->p.particle_radius;
    const double k_norm = young * r, k_tang = k_norm * poisson;
    const double tan_phi = tan(friction_angle);
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; i++) {
        const double* q = grains.x + 3 * i;
        dem_ctx c = {dim, q, v0 + 3 * i, v0, 2 * r, 0.0001 * r, k_norm, k_tang,
                     tan_phi, {0, 0, 0}};
        set_for_each(&grains, q, cb_dem_contact, &c);
        /* force = dt * getForceSum(i), then each wall term added to it */
        dem_wall_ctx cw = {s, q, v0 + 3 * i, k_tang, tan_phi,
                           {c.f[0] * dt, c.f[1] * dt, c.f[2] * dt}};
        set_for_each(&s->bnd, q, cb_dem_wall, &cw);
        for (int a = 0; a < dim; a++) s->v[3 * i + a] = v0[3 * i + a] + cw.f[a];
        /* applyExternalForces */
        s->v[3 * i + 1] -= gravity * dt;
    }
    free(v0);
    hash_free(&grains.hash);
    set_reset(&s->real);
    return 0;
}

/* ------------------------------------------------------------------------
 * The rest of the SPH step (SmthParticleHydrodLiquid.cpp:59-98).
 * Implicit surfaces of Cores/Geometries/ImplicitSurface.h:27-113 and
 * ComplementarySurface (Surface.h:50-69): signed distance, closest normal
 * and closest position, each in the reference's order of operations.
 * ---------------------------------------------------------------------- */
static void or_normalize(int dim, double* v) {
    const double z = sqnorm(dim, v);
    if (z > 0) {
        const double n = sqrt(z);
        for (int a = 0; a < dim; a++) v[a] = v[a] / n;
    }
}

/* signed distance, normal and closest position of the surface itself */
static double surf_inner(const vps_surface* s, int dim, const double* p, double* n,
                         double* cp) {
    n[0] = n[1] = n[2] = 0;
    if (s->kind == VPS_SURFACE_BOX) {
        double d[3] = {0, 0, 0}, phi[3] = {0, 0, 0};
        int all_in = 1, ax = 0;
        for (int a = 0; a < dim; a++) {
            const double c = s->a[a] + s->b[a] / 2, h = s->b[a] / 2;
            d[a] = p[a] - c;
            phi[a] = fabs(d[a]) - h;
            all_in = all_in && phi[a] <= 0;
            if (a > 0 && phi[a] > phi[ax]) ax = a;
        }
        double sd;
        if (all_in) {
            sd = phi[ax];
            n[ax] = 1;
        } else {
            double m[3] = {0, 0, 0};
            for (int a = 0; a < dim; a++) m[a] = phi[a] > 0 ? phi[a] : 0;
            sd = sqrt(sqnorm(dim, m));
            for (int a = 0; a < dim; a++) n[a] = m[a];
        }
        for (int a = 0; a < dim; a++) n[a] = n[a] * (d[a] > 0 ? 1. : (d[a] < 0 ? -1. : 0.));
        or_normalize(dim, n);
        /* ImplicitSurface::closestPosition: pos - signedDistance * normal */
        for (int a = 0; a < dim; a++) cp[a] = p[a] - sd * n[a];
        return sd;
    }
    if (s->kind == VPS_SURFACE_SPHERE) {
        for (int a = 0; a < dim; a++) n[a] = p[a] - s->a[a];
        const double sd = sqrt(sqnorm(dim, n)) - s->b[0];
        or_normalize(dim, n);
        /* ImplicitSphere::closestPosition: center + normal * radius */
        for (int a = 0; a < dim; a++) cp[a] = s->a[a] + n[a] * s->b[0];
        return sd;
    }
    if (s->kind == VPS_SURFACE_PLANE) {
        double u[3] = {s->b[0], s->b[1], s->b[2]}, d[3] = {0, 0, 0};
        or_normalize(dim, u);
        for (int a = 0; a < dim; a++) {
            d[a] = p[a] - s->a[a];
            n[a] = u[a];
        }
        const double sd = dot3(dim, d, u);
        for (int a = 0; a < dim; a++) cp[a] = p[a] - sd * n[a];
        return sd;
    }
    /* ellipsoid (measured from the origin, as the reference does) */
    double q[3] = {0, 0, 0};
    for (int a = 0; a < dim; a++) q[a] = p[a] / s->b[a];
    const double s2 = sqnorm(dim, q);
    const int in = (s2 - 1) <= 0;
    const double sq = sqrt(s2);
    for (int a = 0; a < dim; a++) {
        cp[a] = p[a] / sq;
        n[a] = p[a] - cp[a];
    }
    const double dist = sqrt(sqnorm(dim, n));
    or_normalize(dim, n);
    const double sg = in ? -1. : 1.;
    for (int a = 0; a < dim; a++) n[a] = n[a] * sg;
    return dist * sg;
}

static double surf_eval(const vps_surface* s, int dim, const double* p, double* n,
                        double* cp) {
    double sd = surf_inner(s, dim, p, n, cp);
    if (s->complement) {
        sd = -sd;
        for (int a = 0; a < dim; a++) n[a] = -n[a];
    }
    return sd;
}

/* Collider::collide(pos, vel, radius) with a static collider
 * (Collider.cpp:21-59): inside when signedDistance - radius <= 0 */
static void collide_one(const vps_collider* c, int dim, double r, double* p, double* v) {
    double n[3], cp[3];
    const double sd = surf_eval(&c->surface, dim, p, n, cp);
    if (!(sd - r <= 0)) return;
    const double nd = dot3(dim, n, v);   /* colliderVel = 0 */
    if (nd < 0) {
        double rn[3] = {0, 0, 0}, rt[3] = {0, 0, 0}, drn[3] = {0, 0, 0};
        for (int a = 0; a < dim; a++) {
            rn[a] = nd * n[a];
            rt[a] = v[a] - rn[a];
            drn[a] = -(c->restitution + 1) * rn[a];
            rn[a] = rn[a] * -c->restitution;
        }
        int any = 0;
        for (int a = 0; a < dim; a++) any |= rt[a] != 0;
        if (any) {
            double f = 1 - c->friction * sqrt(sqnorm(dim, drn)) / sqrt(sqnorm(dim, rt));
            f = f > 0 ? f : 0;
            for (int a = 0; a < dim; a++) rt[a] = rt[a] * f;
        }
        for (int a = 0; a < dim; a++) v[a] = (rn[a] + rt[a]) + 0.0;
    }
    for (int a = 0; a < dim; a++) p[a] = cp[a] + r * n[a];
}

/* SmthParticleHydrodLiquid::moveParticles (:72-82): x += v dt, every
 * collider over every particle, neighbour rebuild and densities */
int or_move_particles(or_sim* s, double dt, const vps_collider* col, int32_t ncol) {
    int64_t i;
    const int dim = s->dim;
    for (i = 0; i < s->real.n; i++)
        for (int a = 0; a < dim; a++) s->real.x[3 * i + a] += s->v[3 * i + a] * dt;
    for (int c = 0; c < ncol; c++)
        for (i = 0; i < s->real.n; i++)
            collide_one(&col[c], dim, s->p.particle_radius, s->real.x + 3 * i, s->v + 3 * i);
    set_reset(&s->real);
    set_compute_densities(&s->real);
    return 0;
}

/* applyExternalForces (:84-88) */
int or_apply_gravity(or_sim* s, double dt, double gravity) {
    for (int64_t i = 0; i < s->real.n; i++) s->v[3 * i + 1] -= gravity * dt;
    return 0;
}

typedef struct {
    const or_sim* s;
    const double* q;
    const double* vi;
    double lap[3];
} visc_ctx;

/* ParticlesBasedVectorField::divFreeLaplacianAtDataPoint
 * (ParticlesBasedVectorField.cpp:56-69): deltaPos = x_j - x_i, skipped
 * when zero */
static void cb_visc(void* c, int32_t j, const double* xj) {
    visc_ctx* t = (visc_ctx*)c;
    const or_sim* s = t->s;
    const int dim = s->dim;
    double d[3] = {xj[0] - t->q[0], xj[1] - t->q[1], xj[2] - t->q[2]};
    if (!(d[0] != 0 || d[1] != 0 || (dim == 3 && d[2] != 0))) return;
    double dv[3] = {0, 0, 0}, g[3];
    for (int a = 0; a < dim; a++) dv[a] = s->v[3 * j + a] - t->vi[a];
    const double coef = dot3(dim, dv, d) / sqnorm(dim, d);
    kern_grad(&s->real.k, d, g);
    for (int a = 0; a < dim; a++) t->lap[a] += coef * g[a] / s->real.rho[j];
}

/* applyViscosityForce (:90-98): v_i += nu * lapl_i * dt on a copy */
int or_apply_viscosity(or_sim* s, double dt, double coeff) {
    if (coeff == 0) return 0;
    const int dim = s->dim;
    const int64_t n = s->real.n;
    double* vnew = (double*)malloc((size_t)n * 3 * sizeof(double));
    if (!vnew) return VPS_ERR_OOM;
    memcpy(vnew, s->v, (size_t)n * 3 * sizeof(double));
    int64_t i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; i++) {
        visc_ctx c = {s, s->real.x + 3 * i, s->v + 3 * i, {0, 0, 0}};
        set_for_each(&s->real, c.q, cb_visc, &c);
        for (int a = 0; a < dim; a++) {
            const double lap = c.lap[a] * 2 * (dim + 2.0) * s->real.mass;
            vnew[3 * i + a] += coeff * lap * dt;
        }
    }
    memcpy(s->v, vnew, (size_t)n * 3 * sizeof(double));
    free(vnew);
    return 0;
}

/* DualParticleSphLiquid::advance (DualParticleSphLiquid.cpp:107-116) */
int or_advance(or_sim* s, double dt, double gravity, int32_t enable_gravity,
               double viscosity, const vps_collider* col, int32_t ncol, or_stats* st) {
    if (!s->initialized) return VPS_ERR_STATE;
    int rc = or_move_particles(s, dt, col, ncol);
    if (rc) return rc;
    if (enable_gravity) or_apply_gravity(s, dt, gravity);
    rc = or_apply_viscosity(s, dt, viscosity);
    if (rc) return rc;
    rc = or_generate(s);
    if (rc) return rc;
    return or_project(s, st);
}

/* ---- read-back ---------------------------------------------------------- */
int64_t or_virtual_count(or_sim* s) { return s->virt.n; }

int or_get_virtual(or_sim* s, int64_t* node, double* rho, double* vol,
                   double* vbar, double* div, double* diag, double* p) {
    const int64_t n = s->virt.n;
    for (int64_t I = 0; I < n; I++) {
        if (node) node[I] = s->node[I];
        if (rho) rho[I] = s->virt.rho[I];
        if (vol) vol[I] = s->virt.vol[I];
        if (vbar && s->vbar)
            for (int a = 0; a < s->dim; a++) vbar[s->dim * I + a] = s->vbar[3 * I + a];
        if (div && s->div) div[I] = s->div[I];
        if (diag && s->diag) diag[I] = s->diag[I];
        if (p && s->pres && I < s->npres) p[I] = s->pres[I];
    }
    return 0;
}

int or_get_initial_guess(or_sim* s, double* x0) {
    if (!s->x0) return VPS_ERR_STATE;
    memcpy(x0, s->x0, (size_t)s->npres * sizeof(double));
    return 0;
}

int or_set_pressure(or_sim* s, const double* p) {
    const int64_t n = s->virt.n;
    double* np = (double*)malloc((size_t)(n ? n : 1) * sizeof(double));
    memcpy(np, p, (size_t)n * sizeof(double));
    free(s->pres);
    s->pres = np;
    s->npres = n;
    return 0;
}

int or_get_real_fields(or_sim* s, double* rho, double* vol) {
    for (int64_t i = 0; i < s->real.n; i++) {
        if (rho) rho[i] = s->real.rho[i];
        if (vol) vol[i] = s->real.vol[i];
    }
    return 0;
}
int or_get_boundary_fields(or_sim* s, double* rho, double* vol) {
    for (int64_t i = 0; i < s->bnd.n; i++) {
        if (rho) rho[i] = s->bnd.rho[i];
        if (vol) vol[i] = s->bnd.vol[i];
    }
    return 0;
}
int or_get_scalars(or_sim* s, double* rho0, double* kappa, double* alpha0) {
    if (rho0) *rho0 = s->rho0;
    if (kappa) *kappa = s->kappa;
    if (alpha0) *alpha0 = s->alpha0;
    return 0;
}
int or_get_node_weights(or_sim* s, double* w) {
    if (!s->w) return VPS_ERR_STATE;
    memcpy(w, s->w, (size_t)s->nnodes * sizeof(double));
    return 0;
}

typedef struct {
    int64_t* buf;
    int64_t n, cap;
} list_ctx;
static void cb_list(void* c, int32_t j, const double* xj) {
    list_ctx* t = (list_ctx*)c;
    (void)xj;
    if (t->n < t->cap) t->buf[t->n] = j;
    t->n++;
}

static int query_impl(or_sim* s, int32_t which, int64_t nq, const double* q,
                      int64_t* offsets, int64_t* idx, int64_t cap,
                      int64_t* needed, int sorted) {
    const or_set* set = which == 0 ? &s->real : which == 1 ? &s->bnd : &s->virt;
    int64_t total = 0;
    if (set->search_mode == OR_SEARCH_HASH && set->n > 0 && set->hash.nbuckets == 0)
        return VPS_ERR_STATE;
    offsets[0] = 0;
    for (int64_t k = 0; k < nq; k++) {
        double qq[3] = {q[s->dim * k], q[s->dim * k + 1],
                        s->dim == 3 ? q[s->dim * k + 2] : 0.0};
        list_ctx c = {idx ? idx + total : NULL, 0, idx ? cap - total : 0};
        if (c.cap < 0) c.cap = 0;
        set_for_each(set, qq, cb_list, &c);
        if (sorted && c.n <= c.cap && idx)
            qsort(idx + total, (size_t)c.n, sizeof(int64_t), cmp_i64);
        total += c.n;
        offsets[k + 1] = total;
    }
    if (needed) *needed = total;
    return total > cap ? VPS_ERR_INVALID_ARG : 0;
}
int or_query(or_sim* s, int32_t which, int64_t nq, const double* q,
             int64_t* offsets, int64_t* idx, int64_t cap, int64_t* needed) {
    return query_impl(s, which, nq, q, offsets, idx, cap, needed, 1);
}
int or_query_ordered(or_sim* s, int32_t which, int64_t nq, const double* q,
                     int64_t* offsets, int64_t* idx, int64_t cap,
                     int64_t* needed) {
    return query_impl(s, which, nq, q, offsets, idx, cap, needed, 0);
}

int64_t or_matrix_nnz(or_sim* s) {
    return s->rowptr ? s->rowptr[s->virt.n] : 0;
}
int or_get_matrix(or_sim* s, int64_t* rowptr, int32_t* col, double* val) {
    if (!s->rowptr) return VPS_ERR_STATE;
    const int64_t n = s->virt.n;
    memcpy(rowptr, s->rowptr, (size_t)(n + 1) * sizeof(int64_t));
    if (col) memcpy(col, s->col, (size_t)s->rowptr[n] * sizeof(int32_t));
    if (val) memcpy(val, s->val, (size_t)s->rowptr[n] * sizeof(double));
    return 0;
}
int or_apply_operator(or_sim* s, const double* p, double* out) {
    if (!s->rowptr) return VPS_ERR_STATE;
    csr_mv(s->virt.n, s->rowptr, s->col, s->val, p, out);
    return 0;
}
