// vps_api.hip -- the extern "C" entry points declared in include/vps.h.
//
// Orchestration follows the reference's call order:
//   initialize  DualParticleSphLiquid::initialize
//               (Cores/Physics/DualParticleSphLiquid.cpp:78-110)
//   step        moveParticles' hash rebuild + densities
//               (SmthParticleHydrodLiquid.cpp:80-81), generateParticles,
//               project (DualParticleSphLiquid.cpp:112-126)
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "vps_internal.h"

using namespace vps;

namespace {

thread_local std::string g_create_error;

inline unsigned blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

__global__ void k_upload(int64_t n, int dim, const double* x, const double* v,
                         double4* pos, float4* vel, int32_t* id) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    pos[i] = make_double4(x[dim * i], x[dim * i + 1], dim == 3 ? x[dim * i + 2] : 0.0, 0.0);
    vel[i] = v ? make_float4((float)v[dim * i], (float)v[dim * i + 1],
                             dim == 3 ? (float)v[dim * i + 2] : 0.f, 0.f)
               : make_float4(0.f, 0.f, 0.f, 0.f);
    id[i] = (int32_t)i;
}

// Overwrites x and v of the resident (cell-sorted) state from caller-order
// arrays, keeping the frozen volumes.
__global__ void k_upload_resident(int64_t n, int dim, const double* x,
                                  const double* v, const int32_t* id,
                                  double4* pos, float4* vel) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const int64_t i = id[k];
    pos[k] = make_double4(x[dim * i], x[dim * i + 1], dim == 3 ? x[dim * i + 2] : 0.0, 0.0);
    if (v) {
        float4 u = vel[k];
        u.x = (float)v[dim * i];
        u.y = (float)v[dim * i + 1];
        u.z = dim == 3 ? (float)v[dim * i + 2] : 0.f;
        vel[k] = u;
    }
}

__global__ void k_upload_boundary(int64_t n, int dim, const double* x,
                                  const double* nr, const double* v,
                                  double4* pos, float4* vel, float4* nrm,
                                  int32_t* id) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    pos[i] = make_double4(x[dim * i], x[dim * i + 1], dim == 3 ? x[dim * i + 2] : 0.0, 0.0);
    vel[i] = v ? make_float4((float)v[dim * i], (float)v[dim * i + 1],
                             dim == 3 ? (float)v[dim * i + 2] : 0.f, 0.f)
               : make_float4(0.f, 0.f, 0.f, 0.f);
    nrm[i] = make_float4((float)nr[dim * i], (float)nr[dim * i + 1],
                         dim == 3 ? (float)nr[dim * i + 2] : 0.f, 0.f);
    id[i] = (int32_t)i;
}

int fail(vps_context* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

double packed_kernel_sum(int dim, double radius) {
    // SmoothedParticles::getPackedKernelSum (SmoothedParticles.cpp:74-80)
    const double kr = radius * 4;
    const double inv = 1 / kr;
    const double c0 = 0.31830988618379067154 * (1 / (kr * kr)) *
                      (dim == 2 ? (double)40 / 7 : 8 * inv);
    auto W = [&](double d) {
        const double x = d * inv;
        if (x < .5) return c0 * (6 * x * x * (x - 1) + 1);
        else if (x < 1) return c0 * 2 * (1 - x) * (1 - x) * (1 - x);
        return 0.0;
    };
    if (dim == 2)
        return W(0) + W(radius * 2) * 6 + W(radius * 2 * 1.73205080756887729353) * 6;
    return W(0) + W(radius * 2) * 12 + W(radius * 2 * 1.41421356237309504880) * 6 +
           W(radius * 2 * 1.73205080756887729353) * 24;
}

void fill_kconst(vps_context* c) {
    const vps_params& p = c->p;
    KConst& k = c->k;
    k.dim = p.dim;
    // SmoothedParticles ctor (SmoothedParticles.cpp:13-23): kr = radius * 4
    k.h = p.particle_radius * 4;
    k.h2 = k.h * k.h;
    k.H = p.lattice_spacing * 4;
    k.H2 = k.H * k.H;
    const double inv_h = 1 / k.h, inv_H = 1 / k.H;
    const double c0h = 0.31830988618379067154 * (1 / k.h2) *
                       (p.dim == 2 ? (double)40 / 7 : 8 * inv_h);
    const double c0H = 0.31830988618379067154 * (1 / k.H2) *
                       (p.dim == 2 ? (double)40 / 7 : 8 * inv_H);
    k.inv_h64 = inv_h;
    k.c0h64 = c0h;
    k.hf = (float)k.h;
    k.inv_h = (float)inv_h;
    k.c0h = (float)c0h;
    k.c1h = (float)(6 * inv_h * c0h);
    k.Hf = (float)k.H;
    k.inv_H = (float)inv_H;
    k.c0H = (float)c0H;
    k.c1H = (float)(6 * inv_H * c0H);
    for (int a = 0; a < 3; a++) {
        k.origin[a] = a < p.dim ? p.lattice_origin[a] : 0.0;
        k.ndims[a] = a < p.dim ? p.lattice_dims[a] : 1;
        k.bdims[a] = a < p.dim ? (k.ndims[a] + kBrick - 1) / kBrick : 1;
    }
    k.spacing = p.lattice_spacing;
    // cells a hair above h: every neighbour within h stays within one cell
    // of the FP64-rounded cell of a point, while a lattice node on a cell
    // face (nodes every h / 2 on a grid aligned with the cells) stays within
    // 1e-9 h of it over any domain, so the sums can skip the cell behind it
    const double cs = k.h * (1 + 1e-12);
    k.cell_inv = 1 / cs;
    k.cell_size = cs;
    for (int a = 0; a < 3; a++) {
        if (a < p.dim) {
            k.cell_lo[a] = p.domain_lo[a];
            const double ext = p.domain_hi[a] - p.domain_lo[a];
            k.cdims[a] = std::max(1, (int)std::ceil(ext / cs) + 1);
        } else {
            k.cell_lo[a] = 0;
            k.cdims[a] = 1;
        }
    }
    k.mass = (float)p.mass;
    k.rho0 = (float)c->rho0;
    k.kappa = (float)c->kappa;
    {
        // per-axis error of the FP32 difference: the roundings of the cell
        // offsets (< 2 cells), of the corner term and of the two FP32 sums
        // are each 2^-24 of at most 4 cells, plus the node coordinate of a
        // brick (< 14 s); 16 cells + 14 s + 2 h bounds their sum
        const double u = std::ldexp(1.0, -24);
        const double delta = u * (14 * k.spacing + 16 * cs + 2 * k.h);
        const double err = 2 * std::sqrt(3.0) * k.h * delta + 4 * u * k.h2;
        const double band = 4 * err;
        k.h2f_lo = std::nextafter((float)(k.h2 - band), 0.f);
        k.h2f_hi = std::nextafter((float)(k.h2 + band), 1e30f);
    }
}

bool finite_all(const double* a, int64_t n) {
    for (int64_t i = 0; i < n; i++)
        if (!std::isfinite(a[i])) return false;
    return true;
}

void free_all(vps_context* c) {
    for (int b = 0; b < 2; b++) {
        dfree(c->rpos[b]);
        dfree(c->rvel[b]);
        dfree(c->rid[b]);
        dfree(c->brick_slot[b]);
        dfree(c->slot_brick[b]);
        dfree(c->nbr_slot);
        dfree(c->col_run);
        dfree(c->col_order);
        dfree(c->vx[b]);
        dfree(c->vp[b]);
        dfree(c->p_by_rank[b]);
    }
    dfree(c->rrho);
    dfree(c->rrinv);
    dfree(c->rofs);
    dfree(c->bofs);
    dfree(c->vel_next);
    dfree(c->grad_chunks);
    dfree(c->grad_over);
    dfree(c->couple_stage);
    dfree(c->cell_of);
    dfree(c->slot_of);
    dfree(c->perm);
    dfree(c->cell_count);
    dfree(c->cell_start);
    dfree(c->bpos);
    dfree(c->bvel);
    dfree(c->bnrm);
    dfree(c->bid);
    dfree(c->bpos_spare);
    dfree(c->bvel_spare);
    dfree(c->bnrm_spare);
    dfree(c->bid_spare);
    dfree(c->bcell_start);
    dfree(c->brick_flag);
    dfree(c->vrho);
    dfree(c->vvol);
    dfree(c->vb);
    dfree(c->vdiag);
    dfree(c->vr);
    dfree(c->vt);
    dfree(c->vr2);
    dfree(c->vt2);
    dfree(c->vrt);
    dfree(c->vq);
    dfree(c->vdiag64);
    dfree(c->vxb);
    dfree(c->vrowsum);
    dfree(c->vrank);
    dfree(c->bits);
    dfree(c->word_prefix);
    dfree(c->partials);
    dfree(c->cg);
    dfree(c->scan_tmp);
    dfree(c->counters);
}

int ensure_particle_capacity(vps_context* c, int64_t n) {
    for (int b = 0; b < 2; b++) {
        VPS_TRY(dalloc(c, c->rpos[b], (size_t)n));
        VPS_TRY(dalloc(c, c->rvel[b], (size_t)n));
        VPS_TRY(dalloc(c, c->rid[b], (size_t)n));
    }
    VPS_TRY(dalloc(c, c->rrho, (size_t)n));
    VPS_TRY(dalloc(c, c->rrinv, (size_t)n));
    VPS_TRY(dalloc(c, c->rofs, (size_t)n));
    const int64_t m = std::max(n, c->nb);
    VPS_TRY(dalloc(c, c->cell_of, (size_t)m));
    VPS_TRY(dalloc(c, c->slot_of, (size_t)m));
    VPS_TRY(dalloc(c, c->perm, (size_t)m));
    return VPS_OK;
}

double elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
    return ms;
}

void fill_common_stats(vps_context* c, vps_stats* st) {
    st->n_real = c->n;
    st->n_boundary = c->nb;
    st->n_virtual = c->nvirt;
    st->n_active_bricks = c->nslots;
    st->alpha0 = c->alpha0;
    st->kappa = c->kappa;
    st->rest_density = c->rho0;
}

}  // namespace

namespace vps {
// Boundary set from device FP64 AoS arrays (n x dim each): upload into the
// cell-sorted layout, densities rho_b and volumes V_b = m / rho_b.
int upload_boundary_device(vps_context* c, int64_t n, const double* dx,
                           const double* dn, const double* dv) {
    const int D = c->dim;
    if (c->bpos.n < (size_t)n) {
        dfree(c->bpos);
        dfree(c->bvel);
        dfree(c->bnrm);
        dfree(c->bid);
        VPS_TRY(dalloc(c, c->bpos, (size_t)n));
        VPS_TRY(dalloc(c, c->bvel, (size_t)n));
        VPS_TRY(dalloc(c, c->bnrm, (size_t)n));
        VPS_TRY(dalloc(c, c->bid, (size_t)n));
    }
    if (!c->bcell_start.p) VPS_TRY(dalloc(c, c->bcell_start, (size_t)c->ncells + 1));
    c->nb = n;
    hipLaunchKernelGGL(k_upload_boundary, dim3(blocks(n, 256)), dim3(256), 0,
                       c->stream, n, D, dx, dn, dv, c->bpos.p, c->bvel.p,
                       c->bnrm.p, c->bid.p);
    VPS_HIP(c, hipGetLastError());
    VPS_TRY(cells_sort_boundary(c));
    VPS_TRY(boundary_density(c));
    return VPS_OK;
}
}  // namespace vps

extern "C" {

int32_t vps_abi_version(void) { return VPS_ABI_VERSION; }
int64_t vps_sizeof_params(void) { return (int64_t)sizeof(vps_params); }
int64_t vps_sizeof_stats(void) { return (int64_t)sizeof(vps_stats); }

const char* vps_last_error(const vps_context* ctx) {
    if (!ctx) return g_create_error.c_str();
    return ctx->err.c_str();
}

int vps_create(const vps_params* params, vps_context** out) {
    if (!out) return fail(nullptr, VPS_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (!params) return fail(nullptr, VPS_ERR_INVALID_ARG, "params is NULL");
    const vps_params& p = *params;
    if (p.dim != 2 && p.dim != 3)
        return fail(nullptr, VPS_ERR_INVALID_ARG, "dim must be 2 or 3");
    if (!(p.particle_radius > 0) || !(p.lattice_spacing > 0) || !(p.mass > 0))
        return fail(nullptr, VPS_ERR_INVALID_ARG,
                    "particle_radius, lattice_spacing and mass must be > 0");
    for (int a = 0; a < p.dim; a++) {
        if (p.lattice_dims[a] <= 0)
            return fail(nullptr, VPS_ERR_INVALID_ARG, "lattice_dims must be > 0");
        if (!(p.domain_hi[a] > p.domain_lo[a]))
            return fail(nullptr, VPS_ERR_INVALID_ARG, "domain_hi must exceed domain_lo");
    }
    if (p.op_mode != VPS_OP_TRANSPOSE && p.op_mode != VPS_OP_ROW)
        return fail(nullptr, VPS_ERR_INVALID_ARG, "unknown op_mode");
    if (p.warm_start < VPS_WARM_INDEX_CARRY || p.warm_start > VPS_WARM_EXTRAPOLATE)
        return fail(nullptr, VPS_ERR_INVALID_ARG, "unknown warm_start");
    if (p.rest_density_mode == VPS_FROM_PARAMS && !(p.rest_density > 0))
        return fail(nullptr, VPS_ERR_INVALID_ARG, "rest_density must be > 0");
    vps_context* c = new vps_context();
    c->p = p;
    c->dim = p.dim;
    if (p.dim == 2) c->p.lattice_dims[2] = 1;
    c->brick_nodes = p.dim == 3 ? kBrickNodes3 : kBrickNodes2;
    if (hipGetDevice(&c->device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(nullptr, VPS_ERR_HIP, "no HIP device or stream");
    }
    fill_kconst(c);
    const KConst& k = c->k;
    c->ncells = (int64_t)k.cdims[0] * k.cdims[1] * k.cdims[2];
    c->nnodes = (int64_t)k.ndims[0] * k.ndims[1] * k.ndims[2];
    c->nbricks = (int64_t)k.bdims[0] * k.bdims[1] * k.bdims[2];
    c->nwords = (c->nnodes + 31) / 32;
    if (c->ncells > (int64_t)1 << 30 || c->nnodes > ((int64_t)1 << 31) - 1024) {
        delete c;
        return fail(nullptr, VPS_ERR_INVALID_ARG, "domain or lattice too large");
    }
    // the 3D cell offsets carry the cell's x index + 1 exactly in FP32
    if (p.dim == 3 && k.cdims[0] >= (1 << 24)) {
        delete c;
        return fail(nullptr, VPS_ERR_INVALID_ARG, "domain too large (2^24 or more cells along x)");
    }
    // the 3D CG iteration addresses the per-node arrays (bricks and the zero
    // brick, 2 KB each) with 32-bit byte offsets
    if (p.dim == 3 && (c->nbricks + 1) * 2048 > (int64_t)UINT32_MAX) {
        delete c;
        return fail(nullptr, VPS_ERR_INVALID_ARG, "lattice too large (more than 2^21 - 1 bricks)");
    }
    // every per-node array has one extra all-zero brick at index nbricks,
    // read by the stencil in place of absent neighbour bricks
    const size_t cap = (size_t)(c->nbricks + 1) * c->brick_nodes;
    int rc = VPS_OK;
#define ALLOC(buf, n)                                   \
    if (rc == VPS_OK) rc = dalloc(c, buf, (size_t)(n));
    ALLOC(c->cell_count, c->ncells + 1);
    ALLOC(c->cell_start, c->ncells + 1);
    ALLOC(c->brick_flag, c->nbricks);
    for (int b = 0; b < 2; b++) {
        ALLOC(c->brick_slot[b], c->nbricks);
        ALLOC(c->slot_brick[b], c->nbricks);
        ALLOC(c->vx[b], cap);
        ALLOC(c->vp[b], cap);
        ALLOC(c->p_by_rank[b], cap);
    }
    ALLOC(c->vrho, cap);
    ALLOC(c->vvol, cap);
    ALLOC(c->vb, cap);
    ALLOC(c->vdiag, cap);
    ALLOC(c->vdiag64, cap);
    ALLOC(c->vxb, cap);
    ALLOC(c->vrowsum, cap);
    ALLOC(c->vr, cap);
    ALLOC(c->vt, cap);
    ALLOC(c->vr2, cap);
    ALLOC(c->vt2, cap);
    ALLOC(c->vrt, cap);
    ALLOC(c->vq, cap);
    ALLOC(c->vrank, cap);
    ALLOC(c->nbr_slot, (size_t)(c->nbricks + 1) * 27);
    ALLOC(c->col_run, c->nbricks + 1);
    ALLOC(c->col_order, c->nbricks + 1);
    ALLOC(c->bits, c->nwords + 1);
    ALLOC(c->word_prefix, c->nwords + 1);
    ALLOC(c->partials, partials_size(c->nbricks));
    ALLOC(c->cg, 1);
    ALLOC(c->counters, 64);
#undef ALLOC
    if (rc == VPS_OK) {
        if (hipHostMalloc((void**)&c->host_cg, sizeof(CgState)) != hipSuccess ||
            hipHostMalloc((void**)&c->host_counters, 64 * sizeof(int64_t)) !=
                hipSuccess)
            rc = VPS_ERR_OOM;
    }
    if (rc == VPS_OK) {
        for (auto& e : c->ev)
            if (hipEventCreate(&e) != hipSuccess) rc = VPS_ERR_HIP;
        c->kev.resize(4 * 256);
        for (auto& e : c->kev)
            if (hipEventCreate(&e) != hipSuccess) rc = VPS_ERR_HIP;
    }
    if (rc == VPS_OK) {
        CgState z{};
        if (hipMemset(c->cg.p, 0, sizeof(CgState)) != hipSuccess ||
            hipMemset(c->counters.p, 0, 64 * sizeof(int64_t)) != hipSuccess ||
            // counting arrays start at zero; their consumers clear them again
            hipMemset(c->cell_count.p, 0, (size_t)(c->ncells + 1) * 4) != hipSuccess ||
            hipMemset(c->brick_flag.p, 0, (size_t)c->nbricks * 4) != hipSuccess ||
            hipMemset(c->vq.p, 0, cap * 4) != hipSuccess ||
            hipMemset(c->vrho.p, 0, cap * 4) != hipSuccess ||
            hipMemset(c->vvol.p, 0, cap * 4) != hipSuccess ||
            hipMemset(c->vb.p, 0, cap * 4) != hipSuccess ||
            hipMemset(c->vdiag.p, 0, cap * 4) != hipSuccess ||
            hipMemset(c->vdiag64.p, 0, cap * 8) != hipSuccess ||
            hipMemset(c->vxb.p, 0, cap * 8) != hipSuccess ||
            hipMemset(c->vrowsum.p, 0, cap * 4) != hipSuccess ||
            hipMemset(c->vr.p, 0, cap * 4) != hipSuccess ||
            hipMemset(c->vt.p, 0, cap * 4) != hipSuccess ||
            hipMemset(c->vr2.p, 0, cap * 4) != hipSuccess ||
            hipMemset(c->vt2.p, 0, cap * 4) != hipSuccess ||
            hipMemset(c->vrt.p, 0, cap * 4) != hipSuccess ||
            hipMemset(c->vx[0].p, 0, cap * 4) != hipSuccess ||
            hipMemset(c->vx[1].p, 0, cap * 4) != hipSuccess ||
            hipMemset(c->vp[0].p, 0, cap * 4) != hipSuccess ||
            hipMemset(c->vp[1].p, 0, cap * 4) != hipSuccess ||
            hipMemset(c->brick_slot[0].p, 0xff, (size_t)c->nbricks * 4) != hipSuccess ||
            hipMemset(c->brick_slot[1].p, 0xff, (size_t)c->nbricks * 4) != hipSuccess)
            rc = VPS_ERR_HIP;
        (void)z;
    }
    if (rc == VPS_OK) rc = mlp_create(c);
    if (rc == VPS_OK) rc = gradient_setup(c);
    if (rc != VPS_OK) {
        const std::string msg = c->err.empty() ? "allocation failed" : c->err;
        vps_destroy(c);
        return fail(nullptr, rc, msg);
    }
    *out = c;
    return VPS_OK;
}

void vps_destroy(vps_context* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    mlp_destroy(c);
    free_all(c);
    if (c->host_cg) (void)hipHostFree(c->host_cg);
    if (c->host_counters) (void)hipHostFree(c->host_counters);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : c->kev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int vps_set_boundary(vps_context* c, int64_t n, const double* x,
                     const double* normals, const double* v) {
    if (!c) return VPS_ERR_INVALID_ARG;
    if (n < 0 || (n > 0 && (!x || !normals)))
        return fail(c, VPS_ERR_INVALID_ARG, "boundary: bad size or NULL array");
    const int D = c->dim;
    if (!finite_all(x, n * D) || !finite_all(normals, n * D) ||
        (v && !finite_all(v, n * D)))
        return fail(c, VPS_ERR_NONFINITE, "boundary: non-finite input");
    c->nb = n;
    c->host_bx.assign(x, x + n * D);
    c->couple_walls = -1;   // staged walls of vps_couple_grains are stale
    c->host_bn.assign(normals, normals + n * D);
    if (v) c->host_bv.assign(v, v + n * D);
    else c->host_bv.assign((size_t)(n * D), 0.0);
    if (n == 0) return VPS_OK;
    VPS_TRY(dalloc(c, c->bpos, (size_t)n));
    VPS_TRY(dalloc(c, c->bvel, (size_t)n));
    VPS_TRY(dalloc(c, c->bnrm, (size_t)n));
    VPS_TRY(dalloc(c, c->bid, (size_t)n));
    VPS_TRY(dalloc(c, c->bcell_start, (size_t)c->ncells + 1));
    return vps_refresh_boundary(c);
}

int vps_refresh_boundary(vps_context* c) {
    if (!c) return VPS_ERR_INVALID_ARG;
    const int64_t n = (int64_t)(c->host_bx.size() / c->dim);
    if (n == 0) return VPS_OK;
    const int D = c->dim;
    DBuf<double> dx, dn, dv;
    VPS_TRY(dalloc(c, dx, (size_t)(n * D)));
    VPS_TRY(dalloc(c, dn, (size_t)(n * D)));
    VPS_TRY(dalloc(c, dv, (size_t)(n * D)));
    VPS_HIP(c, hipMemcpyAsync(dx.p, c->host_bx.data(), (size_t)n * D * 8,
                              hipMemcpyHostToDevice, c->stream));
    VPS_HIP(c, hipMemcpyAsync(dn.p, c->host_bn.data(), (size_t)n * D * 8,
                              hipMemcpyHostToDevice, c->stream));
    VPS_HIP(c, hipMemcpyAsync(dv.p, c->host_bv.data(), (size_t)n * D * 8,
                              hipMemcpyHostToDevice, c->stream));
    VPS_TRY(upload_boundary_device(c, n, dx.p, dn.p, dv.p));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    dfree(dx);
    dfree(dn);
    dfree(dv);
    return VPS_OK;
}


int vps_set_particles(vps_context* c, int64_t n, const double* x,
                      const double* v) {
    if (!c) return VPS_ERR_INVALID_ARG;
    if (n <= 0 || !x)
        return fail(c, VPS_ERR_INVALID_ARG, "particles: empty set or NULL x");
    if (n > ((int64_t)1 << 31) - 1)
        return fail(c, VPS_ERR_INVALID_ARG, "particles: too many");
    const int D = c->dim;
    if (!finite_all(x, n * D) || (v && !finite_all(v, n * D)))
        return fail(c, VPS_ERR_NONFINITE, "particles: non-finite input");
    DBuf<double> dx, dv;
    VPS_TRY(dalloc(c, dx, (size_t)(n * D)));
    if (v) VPS_TRY(dalloc(c, dv, (size_t)(n * D)));
    VPS_HIP(c, hipMemcpyAsync(dx.p, x, (size_t)n * D * 8, hipMemcpyHostToDevice,
                              c->stream));
    if (v)
        VPS_HIP(c, hipMemcpyAsync(dv.p, v, (size_t)n * D * 8,
                                  hipMemcpyHostToDevice, c->stream));
    if (c->initialized && n == c->n) {
        // keep the frozen volumes: overwrite x, v in the resident order
        const int s = c->cur;
        hipLaunchKernelGGL(k_upload_resident, dim3(blocks(n, 256)), dim3(256), 0,
                           c->stream, n, D, dx.p, v ? dv.p : nullptr,
                           c->rid[s].p, c->rpos[s].p, c->rvel[s].p);
    } else {
        c->n = n;
        c->initialized = false;
        c->cg_drift = 1.0;   // a new particle set: no drift measured on it yet
        VPS_TRY(ensure_particle_capacity(c, n));
        c->cur = 0;
        hipLaunchKernelGGL(k_upload, dim3(blocks(n, 256)), dim3(256), 0,
                           c->stream, n, D, dx.p, v ? dv.p : nullptr,
                           c->rpos[0].p, c->rvel[0].p, c->rid[0].p);
    }
    c->generated = false;
    VPS_HIP(c, hipGetLastError());
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    dfree(dx);
    dfree(dv);
    return VPS_OK;
}

int vps_get_particles(vps_context* c, double* x, double* v) {
    if (!c) return VPS_ERR_INVALID_ARG;
    if (c->n == 0) return fail(c, VPS_ERR_STATE, "no particles");
    const int64_t n = c->n;
    const int D = c->dim;
    std::vector<double4> pos((size_t)n);
    std::vector<float4> vel((size_t)n);
    std::vector<int32_t> id((size_t)n);
    const int s = c->cur;
    VPS_HIP(c, hipMemcpyAsync(pos.data(), c->rpos[s].p, (size_t)n * 32,
                              hipMemcpyDeviceToHost, c->stream));
    VPS_HIP(c, hipMemcpyAsync(vel.data(), c->rvel[s].p, (size_t)n * 16,
                              hipMemcpyDeviceToHost, c->stream));
    VPS_HIP(c, hipMemcpyAsync(id.data(), c->rid[s].p, (size_t)n * 4,
                              hipMemcpyDeviceToHost, c->stream));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    for (int64_t k = 0; k < n; k++) {
        const int64_t i = id[(size_t)k];
        const double4& p = pos[(size_t)k];
        const float4& u = vel[(size_t)k];
        if (x) {
            x[D * i] = p.x;
            x[D * i + 1] = p.y;
            if (D == 3) x[D * i + 2] = p.z;
        }
        if (v) {
            v[D * i] = u.x;
            v[D * i + 1] = u.y;
            if (D == 3) v[D * i + 2] = u.z;
        }
    }
    return VPS_OK;
}

int vps_initialize(vps_context* c, vps_stats* st) {
    if (!c) return VPS_ERR_INVALID_ARG;
    if (c->n == 0) return fail(c, VPS_ERR_STATE, "initialize: no particles");
    if (c->nb == 0)
        return fail(c, VPS_ERR_INVALID_ARG,
                    "initialize: empty boundary set (the reference hashes into "
                    "zero buckets there)");
    VPS_TRY(ensure_particle_capacity(c, c->n));
    VPS_HIP(c, hipEventRecord(c->ev[0], c->stream));
    VPS_TRY(cells_sort_real(c));
    VPS_TRY(real_density(c, true));
    if (c->p.rest_density_mode == VPS_FROM_INIT) {
        double* d = (double*)(c->counters.p + 8);
        VPS_TRY(reduce_max_f32(c, c->rrho.p, c->n, d));
        VPS_HIP(c, hipMemcpyAsync(c->host_counters + 8, d, 8,
                                  hipMemcpyDeviceToHost, c->stream));
        VPS_HIP(c, hipStreamSynchronize(c->stream));
        c->rho0 = *(double*)(c->host_counters + 8);
    } else {
        c->rho0 = c->p.rest_density;
    }
    c->kappa = c->p.kappa_mode == VPS_FROM_INIT
                   ? 1. / packed_kernel_sum(c->dim, c->p.particle_radius)
                   : c->p.kappa;
    c->k.rho0 = (float)c->rho0;
    c->k.kappa = (float)c->kappa;
    c->alpha0 = -1e6;   // VirtualParticle ctor (VirtualParticle.cpp:16)
    c->last_solve_buf = -1;
    c->last_solve_nvirt = 0;
    c->prev_iters = 0;
    c->cg_drift = 1.0;
    c->nvirt = 0;
    VPS_TRY(virtual_generate(c));
    VPS_TRY(virtual_diag(c, true));
    VPS_HIP(c, hipEventRecord(c->ev[1], c->stream));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    c->initialized = true;
    c->generated = true;
    if (st) {
        std::memset(st, 0, sizeof(*st));
        fill_common_stats(c, st);
        st->ms_total = elapsed(c->ev[0], c->ev[1]);
        st->cg_true_rel_residual = -1;
    }
    return VPS_OK;
}

int vps_refresh_real(vps_context* c) {
    if (!c) return VPS_ERR_INVALID_ARG;
    if (!c->initialized) return fail(c, VPS_ERR_STATE, "not initialized");
    VPS_TRY(cells_sort_real(c));
    VPS_TRY(real_density(c, false));
    c->generated = false;
    return VPS_OK;
}

int vps_generate_particles(vps_context* c) {
    if (!c) return VPS_ERR_INVALID_ARG;
    if (!c->initialized) return fail(c, VPS_ERR_STATE, "not initialized");
    VPS_TRY(virtual_generate(c));
    c->generated = true;
    return VPS_OK;
}

static int project_impl(vps_context* c, vps_stats* st, int e0) {
    if (!c->initialized) return fail(c, VPS_ERR_STATE, "not initialized");
    if (!c->generated)
        return fail(c, VPS_ERR_STATE,
                    "project: call vps_generate_particles after the last state "
                    "change");
    VPS_HIP(c, hipEventRecord(c->ev[e0], c->stream));
    VPS_TRY(virtual_diag(c, false));
    VPS_HIP(c, hipEventRecord(c->ev[e0 + 1], c->stream));
    const int w = c->which;
    float* x0 = c->vx[w].p;
    const int mode = c->p.warm_start;
    // the history of pressures (and the learned start) is kept only when
    // the network is used or trained
    if (mlp_active(c)) {
        VPS_TRY(mlp_warm_start(c, mode, x0));
    } else {
        VPS_TRY(warm_start(c, mode, x0));
    }
    vps_stats local{};
    vps_stats* s = st ? st : &local;
    // a learned or extrapolated start is guarded: rejected (and the solve run
    // from the plain start) when its residual is not finite or larger than
    // ||b||, or when the iteration from it overflows
    c->start_guard = mode == VPS_WARM_LEARNED || mode == VPS_WARM_EXTRAPOLATE;
    // cg_solve queues the fold, after_solve and the gradient behind its
    // batches (guarded on the device) and records ev[e0 + 2] before the
    // gradient; spec_done says the accepted batch's copy stands
    c->spec_post = e0 == 4;
    int rc = cg_solve(c, s);
    if (rc == kStartRejected) {
        // the plain start: the reference's index-carried pressures
        VPS_TRY(warm_start(c, VPS_WARM_INDEX_CARRY, x0));
        c->start_fallbacks++;
        c->start_guard = false;
        rc = cg_solve(c, s);
        s->reserved[2] = 1;
    }
    c->start_guard = false;
    c->spec_post = false;
    if (rc != VPS_OK && rc != VPS_ERR_NO_CONVERGENCE) {
        c->spec_done = false;
        return rc;
    }
    // The network's history and training read this solve's pressures in the
    // slot map of this solve (index `which`), which after_solve keeps; they
    // touch neither the particles nor V p, so they are queued behind the
    // gradient, which then starts without waiting for the host to launch
    // the short history kernels.
    if (c->spec_done) {
        after_solve_commit(c);   // its kernel and the gradient are queued
        c->spec_done = false;
    } else {
        VPS_TRY(after_solve(c));
        VPS_HIP(c, hipEventRecord(c->ev[e0 + 2], c->stream));
        if (rc == VPS_OK) VPS_TRY(pressure_gradient(c));
    }
    if (mlp_active(c) && rc == VPS_OK) VPS_TRY(mlp_after_solve(c, s));
    VPS_HIP(c, hipEventRecord(c->ev[e0 + 3], c->stream));
    return rc;
}

int vps_project(vps_context* c, vps_stats* st) {
    if (!c) return VPS_ERR_INVALID_ARG;
    if (st) std::memset(st, 0, sizeof(*st));
    const int rc = project_impl(c, st, 4);
    if (rc != VPS_OK && rc != VPS_ERR_NO_CONVERGENCE) return rc;
    if (c->async_steps) {
        // the step's last kernels stay queued: no phase times
        if (st) fill_common_stats(c, st);
        c->generated = false;
        return rc;
    }
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    if (st) {
        fill_common_stats(c, st);
        st->ms_diag = elapsed(c->ev[4], c->ev[5]);
        st->ms_cg = elapsed(c->ev[5], c->ev[6]);
        st->ms_gradient = elapsed(c->ev[6], c->ev[7]);
        st->ms_total = elapsed(c->ev[4], c->ev[7]);
    }
    c->generated = false;
    return rc;
}

int vps_move_particles(vps_context* c, double dt, const vps_collider* colliders,
                       int32_t n_colliders) {
    if (!c) return VPS_ERR_INVALID_ARG;
    if (!c->initialized) return fail(c, VPS_ERR_STATE, "not initialized");
    if (!std::isfinite(dt)) return fail(c, VPS_ERR_INVALID_ARG, "move_particles: dt");
    VPS_TRY(move_collide(c, dt, colliders, n_colliders));
    VPS_TRY(cells_sort_real(c));
    VPS_TRY(real_density(c, false));
    c->generated = false;
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    return VPS_OK;
}

int vps_apply_gravity(vps_context* c, double dt, double gravity) {
    if (!c) return VPS_ERR_INVALID_ARG;
    if (c->n == 0) return fail(c, VPS_ERR_STATE, "apply_gravity: no particles");
    if (!std::isfinite(dt) || !std::isfinite(gravity))
        return fail(c, VPS_ERR_INVALID_ARG, "apply_gravity: non-finite dt or gravity");
    VPS_TRY(apply_gravity(c, dt, gravity));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    return VPS_OK;
}

int vps_apply_viscosity(vps_context* c, double dt, double coefficient) {
    if (!c) return VPS_ERR_INVALID_ARG;
    if (!c->initialized) return fail(c, VPS_ERR_STATE, "not initialized");
    if (!std::isfinite(dt) || !std::isfinite(coefficient))
        return fail(c, VPS_ERR_INVALID_ARG, "apply_viscosity: non-finite dt or coefficient");
    VPS_TRY(apply_viscosity(c, dt, coefficient));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    return VPS_OK;
}

int vps_advance(vps_context* c, const vps_step_params* sp, vps_stats* st) {
    if (!c || !sp) return VPS_ERR_INVALID_ARG;
    if (!c->initialized) return fail(c, VPS_ERR_STATE, "not initialized");
    if (!std::isfinite(sp->dt) || !std::isfinite(sp->gravity) ||
        !std::isfinite(sp->viscosity))
        return fail(c, VPS_ERR_INVALID_ARG, "advance: non-finite dt, gravity or viscosity");
    if (st) std::memset(st, 0, sizeof(*st));
    VPS_HIP(c, hipEventRecord(c->ev[0], c->stream));
    VPS_TRY(move_collide(c, sp->dt, sp->colliders, sp->n_colliders));
    VPS_TRY(cells_sort_real(c));
    VPS_TRY(real_density(c, false));
    if (sp->enable_gravity) VPS_TRY(apply_gravity(c, sp->dt, sp->gravity));
    VPS_TRY(apply_viscosity(c, sp->dt, sp->viscosity));
    VPS_HIP(c, hipEventRecord(c->ev[1], c->stream));
    VPS_TRY(virtual_generate(c));
    c->generated = true;
    VPS_HIP(c, hipEventRecord(c->ev[2], c->stream));
    const int rc = project_impl(c, st, 4);
    if (rc != VPS_OK && rc != VPS_ERR_NO_CONVERGENCE) return rc;
    if (c->async_steps) {
        // the step's last kernels stay queued: no phase times
        if (st) fill_common_stats(c, st);
        c->generated = false;
        return rc;
    }
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    if (st) {
        fill_common_stats(c, st);
        st->ms_sort = elapsed(c->ev[0], c->ev[1]);
        st->ms_generate = elapsed(c->ev[1], c->ev[2]);
        st->ms_diag = elapsed(c->ev[4], c->ev[5]);
        st->ms_cg = elapsed(c->ev[5], c->ev[6]);
        st->ms_gradient = elapsed(c->ev[6], c->ev[7]);
        st->ms_total = elapsed(c->ev[0], c->ev[7]);
    }
    c->generated = false;
    return rc;
}

int vps_projection_step(vps_context* c, vps_stats* st) {
    if (!c) return VPS_ERR_INVALID_ARG;
    if (!c->initialized) return fail(c, VPS_ERR_STATE, "not initialized");
    if (st) std::memset(st, 0, sizeof(*st));
    VPS_HIP(c, hipEventRecord(c->ev[0], c->stream));
    VPS_TRY(cells_sort_real(c));
    VPS_TRY(real_density(c, false));
    VPS_HIP(c, hipEventRecord(c->ev[1], c->stream));
    VPS_TRY(virtual_generate(c));
    c->generated = true;
    VPS_HIP(c, hipEventRecord(c->ev[2], c->stream));
    const int rc = project_impl(c, st, 4);
    if (rc != VPS_OK && rc != VPS_ERR_NO_CONVERGENCE) return rc;
    if (c->async_steps) {
        // the step's last kernels stay queued: no phase times
        if (st) fill_common_stats(c, st);
        c->generated = false;
        return rc;
    }
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    if (st) {
        fill_common_stats(c, st);
        st->ms_sort = elapsed(c->ev[0], c->ev[1]);
        st->ms_generate = elapsed(c->ev[1], c->ev[2]);
        st->ms_diag = elapsed(c->ev[4], c->ev[5]);
        st->ms_cg = elapsed(c->ev[5], c->ev[6]);
        st->ms_gradient = elapsed(c->ev[6], c->ev[7]);
        st->ms_total = elapsed(c->ev[0], c->ev[7]);
    }
    c->generated = false;
    return rc;
}

int vps_advect(vps_context* c, double dt, double gravity) {
    if (!c) return VPS_ERR_INVALID_ARG;
    if (c->n == 0) return fail(c, VPS_ERR_STATE, "no particles");
    if (!std::isfinite(dt) || !std::isfinite(gravity))
        return fail(c, VPS_ERR_NONFINITE, "advect: non-finite dt or gravity");
    VPS_TRY(advect(c, dt, gravity));
    c->generated = false;
    return VPS_OK;
}

int vps_set_check_true_residual(vps_context* c, int32_t enable) {
    if (!c) return VPS_ERR_INVALID_ARG;
    c->check_true = enable != 0;
    return VPS_OK;
}

int vps_set_kernel_timing(vps_context* c, int32_t enable) {
    if (!c) return VPS_ERR_INVALID_ARG;
    c->kernel_timing = enable != 0;
    return VPS_OK;
}

int vps_set_async_steps(vps_context* c, int32_t enable) {
    if (!c) return VPS_ERR_INVALID_ARG;
    c->async_steps = enable != 0;
    return VPS_OK;
}

int vps_synchronize(vps_context* c) {
    if (!c) return VPS_ERR_INVALID_ARG;
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    return VPS_OK;
}

int vps_get_virtual_count(vps_context* c, int64_t* n) {
    if (!c || !n) return VPS_ERR_INVALID_ARG;
    *n = c->nvirt;
    return VPS_OK;
}

int vps_get_virtual(vps_context* c, int64_t* node, double* rho, double* vol,
                    double* div, double* diag, double* pressure) {
    if (!c) return VPS_ERR_INVALID_ARG;
    return gather_virtual(c, node, rho, vol, div, diag, pressure);
}

int vps_get_real_fields(vps_context* c, double* rho, double* vol) {
    if (!c) return VPS_ERR_INVALID_ARG;
    const int64_t n = c->n;
    std::vector<float> r((size_t)n);
    std::vector<float4> v((size_t)n);
    std::vector<int32_t> id((size_t)n);
    const int s = c->cur;
    VPS_HIP(c, hipMemcpyAsync(r.data(), c->rrho.p, (size_t)n * 4,
                              hipMemcpyDeviceToHost, c->stream));
    VPS_HIP(c, hipMemcpyAsync(v.data(), c->rvel[s].p, (size_t)n * 16,
                              hipMemcpyDeviceToHost, c->stream));
    VPS_HIP(c, hipMemcpyAsync(id.data(), c->rid[s].p, (size_t)n * 4,
                              hipMemcpyDeviceToHost, c->stream));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    for (int64_t k = 0; k < n; k++) {
        if (rho) rho[id[(size_t)k]] = r[(size_t)k];
        if (vol) vol[id[(size_t)k]] = v[(size_t)k].w;
    }
    return VPS_OK;
}

int vps_get_boundary_fields(vps_context* c, double* rho, double* vol) {
    if (!c) return VPS_ERR_INVALID_ARG;
    const int64_t n = c->nb;
    if (n == 0) return VPS_OK;
    std::vector<float4> nr((size_t)n), v((size_t)n);
    std::vector<int32_t> id((size_t)n);
    VPS_HIP(c, hipMemcpyAsync(nr.data(), c->bnrm.p, (size_t)n * 16,
                              hipMemcpyDeviceToHost, c->stream));
    VPS_HIP(c, hipMemcpyAsync(v.data(), c->bvel.p, (size_t)n * 16,
                              hipMemcpyDeviceToHost, c->stream));
    VPS_HIP(c, hipMemcpyAsync(id.data(), c->bid.p, (size_t)n * 4,
                              hipMemcpyDeviceToHost, c->stream));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    for (int64_t k = 0; k < n; k++) {
        if (rho) rho[id[(size_t)k]] = nr[(size_t)k].w;
        if (vol) vol[id[(size_t)k]] = v[(size_t)k].w;
    }
    return VPS_OK;
}

int vps_get_scalars(vps_context* c, double* rest_density, double* kappa,
                    double* alpha0) {
    if (!c) return VPS_ERR_INVALID_ARG;
    if (rest_density) *rest_density = c->rho0;
    if (kappa) *kappa = c->kappa;
    if (alpha0) *alpha0 = c->alpha0;
    return VPS_OK;
}

int vps_query_neighbors(vps_context* c, int32_t which, int64_t nq,
                        const double* q, int64_t* offsets, int64_t* indices,
                        int64_t capacity, int64_t* needed) {
    if (!c) return VPS_ERR_INVALID_ARG;
    if (which < 0 || which > 2 || nq < 0 || (nq > 0 && (!q || !offsets)))
        return fail(c, VPS_ERR_INVALID_ARG, "query: bad arguments");
    if (which == 0 && !c->initialized)
        return fail(c, VPS_ERR_STATE, "query: real set not built");
    if (which == 1 && c->nb == 0)
        return fail(c, VPS_ERR_STATE, "query: no boundary");
    if (which == 2 && c->nslots == 0)
        return fail(c, VPS_ERR_STATE, "query: no virtual set");
    if (nq == 0) {
        if (offsets) offsets[0] = 0;
        if (needed) *needed = 0;
        return VPS_OK;
    }
    return query_neighbors(c, which, nq, q, offsets, indices, capacity, needed);
}

int vps_debug_apply_operator(vps_context* c, const double* p, double* out) {
    if (!c || !p || !out) return VPS_ERR_INVALID_ARG;
    if (c->nslots == 0) return fail(c, VPS_ERR_STATE, "no virtual set");
    const int64_t nn = c->nslots * c->brick_nodes;
    std::vector<int32_t> rank((size_t)nn);
    VPS_HIP(c, hipMemcpyAsync(rank.data(), c->vrank.p, (size_t)nn * 4,
                              hipMemcpyDeviceToHost, c->stream));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<float> in((size_t)nn, 0.f), res((size_t)nn);
    for (int64_t g = 0; g < nn; g++)
        if (rank[(size_t)g] >= 0) in[(size_t)g] = (float)p[rank[(size_t)g]];
    // input needs the zero brick at index nbricks, like every node array
    const size_t cap = (size_t)(c->nbricks + 1) * c->brick_nodes;
    DBuf<float> din, dout;
    VPS_TRY(dalloc(c, din, cap));
    VPS_TRY(dalloc(c, dout, (size_t)nn));
    VPS_HIP(c, hipMemsetAsync(din.p, 0, cap * 4, c->stream));
    VPS_HIP(c, hipMemcpyAsync(din.p, in.data(), (size_t)nn * 4,
                              hipMemcpyHostToDevice, c->stream));
    VPS_TRY(apply_operator_f32(c, din.p, dout.p, c->p.op_mode));
    VPS_HIP(c, hipMemcpyAsync(res.data(), dout.p, (size_t)nn * 4,
                              hipMemcpyDeviceToHost, c->stream));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    for (int64_t g = 0; g < nn; g++)
        if (rank[(size_t)g] >= 0) out[rank[(size_t)g]] = res[(size_t)g];
    dfree(din);
    dfree(dout);
    return VPS_OK;
}

int vps_debug_member_counts(vps_context* c, int32_t which, int32_t* counts) {
    if (!c || !counts || which < 0 || which > 3) return VPS_ERR_INVALID_ARG;
    if (c->dim != 3) return fail(c, VPS_ERR_INVALID_ARG, "member counts are 3D only");
    if (which == 2 && (c->nslots == 0 || c->grad_span_cells <= 0))
        return fail(c, VPS_ERR_STATE, "no virtual set or no staged gradient");
    const int64_t n = which == 1 ? c->nb : c->n;
    if (n == 0) return VPS_OK;
    DBuf<int32_t> dm;
    VPS_TRY(dalloc(c, dm, (size_t)n));
    VPS_HIP(c, hipMemsetAsync(dm.p, 0xff, (size_t)n * 4, c->stream));   // -1
    const int rc = which == 2   ? gradient_member_counts(c, dm.p)
                   : which == 3 ? viscosity_member_counts(c, dm.p)
                                : density_member_counts(c, which, dm.p);
    if (rc != VPS_OK) {
        dfree(dm);
        return rc;
    }
    std::vector<int32_t> m((size_t)n), id((size_t)n);
    const int32_t* did = which == 1 ? c->bid.p : c->rid[c->cur].p;
    VPS_HIP(c, hipMemcpyAsync(m.data(), dm.p, (size_t)n * 4, hipMemcpyDeviceToHost,
                              c->stream));
    VPS_HIP(c, hipMemcpyAsync(id.data(), did, (size_t)n * 4, hipMemcpyDeviceToHost,
                              c->stream));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    dfree(dm);
    for (int64_t k = 0; k < n; k++) counts[id[(size_t)k]] = m[(size_t)k];
    return VPS_OK;
}

}  // extern "C"

namespace vps {

int gather_virtual(vps_context* c, int64_t* node, double* rho, double* vol,
                   double* div, double* diag, double* pressure) {
    const int64_t nn = c->nslots * c->brick_nodes;
    if (nn == 0) return VPS_OK;
    const int w = c->which;
    std::vector<int32_t> rank((size_t)nn), sb((size_t)c->nslots);
    VPS_HIP(c, hipMemcpyAsync(rank.data(), c->vrank.p, (size_t)nn * 4,
                              hipMemcpyDeviceToHost, c->stream));
    VPS_HIP(c, hipMemcpyAsync(sb.data(), c->slot_brick[w].p,
                              (size_t)c->nslots * 4, hipMemcpyDeviceToHost,
                              c->stream));
    std::vector<float> buf((size_t)nn);
    std::vector<double> buf64((size_t)nn);
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    auto pull = [&](const float* src, double* dst) -> int {
        if (!dst) return VPS_OK;
        VPS_HIP(c, hipMemcpy(buf.data(), src, (size_t)nn * 4, hipMemcpyDeviceToHost));
        for (int64_t g = 0; g < nn; g++)
            if (rank[(size_t)g] >= 0) dst[rank[(size_t)g]] = buf[(size_t)g];
        return VPS_OK;
    };
    VPS_TRY(pull(c->vrho.p, rho));
    VPS_TRY(pull(c->vvol.p, vol));
    VPS_TRY(pull(c->vb.p, div));
    VPS_TRY(pull(c->vx[w].p, pressure));
    if (diag) {
        VPS_HIP(c, hipMemcpy(buf64.data(), c->vdiag64.p, (size_t)nn * 8,
                             hipMemcpyDeviceToHost));
        for (int64_t g = 0; g < nn; g++)
            if (rank[(size_t)g] >= 0) diag[rank[(size_t)g]] = buf64[(size_t)g];
    }
    if (node) {
        const int NB = c->brick_nodes;
        const KConst& k = c->k;
        for (int64_t g = 0; g < nn; g++) {
            if (rank[(size_t)g] < 0) continue;
            const int brick = sb[(size_t)(g / NB)];
            const int t = (int)(g % NB);
            const int bx = brick % k.bdims[0];
            const int by = (brick / k.bdims[0]) % k.bdims[1];
            const int bz = brick / k.bdims[0] / k.bdims[1];
            const int64_t cx = bx * kBrick + (t & 7);
            const int64_t cy = by * kBrick + ((t >> 3) & 7);
            const int64_t cz = c->dim == 3 ? bz * kBrick + (t >> 6) : 0;
            node[rank[(size_t)g]] = (cz * k.ndims[1] + cy) * k.ndims[0] + cx;
        }
    }
    return VPS_OK;
}

}  // namespace vps
