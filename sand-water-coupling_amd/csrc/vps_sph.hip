// vps_sph.hip -- the rest of the SPH step on the GPU: advection with
// colliders, gravity and the divergence-free viscosity.
//
// Replaces SmthParticleHydrodLiquid::moveParticles / applyExternalForces /
// applyViscosityForce (Cores/Physics/SmthParticleHydrodLiquid.cpp:72-98),
// StaticCollider::collide (Cores/Geometries/Collider.cpp:13-63) and
// ParticlesBasedVectorField::divFreeLaplacianAtDataPoint
// (Cores/Structures/ParticlesBasedVectorField.cpp:56-69).  Positions stay
// FP64 and velocities FP32 as in the projection; the collider arithmetic is
// FP64 in the reference's order.
#include <cmath>

#include "vps_internal.h"
#include "vps_surface.h"

namespace vps {
namespace {

inline unsigned blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

constexpr int kMaxColliders = 8;
struct CollK {
    SurfK s;
    double e, mu;   // restitution, friction
};
struct CollSet {
    CollK c[kMaxColliders];
    int n;
    double r;   // particle radius
};

// Collider::collide(pos, vel, radius) of a static collider: inside when the
// signed distance - radius <= 0; the normal velocity is reversed with the
// restitution, the tangential one reduced by friction (never reversed), and
// the position moved to the closest surface point + radius along the normal.
__device__ void collide_one(const CollK& k, int dim, double r, double* p, double* v) {
#pragma clang fp contract(off)
    double n[3], cp[3];
    const double sd = surf_eval(k.s, dim, p, n, cp);
    if (!(sd - r <= 0)) return;
    double nd = n[0] * v[0] + n[1] * v[1];
    if (dim == 3) nd = nd + n[2] * v[2];
    if (nd < 0) {
        double rn[3] = {0, 0, 0}, rt[3] = {0, 0, 0}, drn[3] = {0, 0, 0};
        bool any = false;
        for (int a = 0; a < dim; a++) {
            rn[a] = nd * n[a];
            rt[a] = v[a] - rn[a];
            drn[a] = -(k.e + 1) * rn[a];
            rn[a] = rn[a] * -k.e;
            any = any || rt[a] != 0;
        }
        if (any) {
            double f = 1 - k.mu * sqrt(surf_sqnorm(dim, drn)) / sqrt(surf_sqnorm(dim, rt));
            f = f > 0 ? f : 0;
            for (int a = 0; a < dim; a++) rt[a] = rt[a] * f;
        }
        for (int a = 0; a < dim; a++) v[a] = (rn[a] + rt[a]) + 0.0;
    }
    for (int a = 0; a < dim; a++) p[a] = cp[a] + r * n[a];
}

template <int D>
__global__ void __launch_bounds__(256) k_move_collide(int64_t n, double4* pos, float4* vel,
                                                      double dt, CollSet cs) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double4 p0 = pos[i];
    const float4 v0 = vel[i];
    double p[3] = {p0.x + (double)v0.x * dt, p0.y + (double)v0.y * dt,
                   D == 3 ? p0.z + (double)v0.z * dt : 0.0};
    double v[3] = {(double)v0.x, (double)v0.y, D == 3 ? (double)v0.z : 0.0};
    for (int k = 0; k < cs.n; k++) collide_one(cs.c[k], D, cs.r, p, v);
    pos[i] = make_double4(p[0], p[1], p[2], p0.w);
    vel[i] = make_float4((float)v[0], (float)v[1], D == 3 ? (float)v[2] : 0.f, v0.w);
}

__global__ void __launch_bounds__(256) k_gravity(int64_t n, float4* vel, float gdt) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) vel[i].y -= gdt;
}

// One neighbour's term of the divergence-free Laplacian: d = x_j - x_i,
// ((v_j - v_i).d / |d|^2) gradK(d) / rho_j with gradK(d) = -W'(r) d / r.
struct ViscAcc {
    float lx = 0.f, ly = 0.f, lz = 0.f;
    __device__ __forceinline__ void add(const KConst& k, float d2, float dx, float dy,
                                        float dz, const float4& vi, const float4& vj,
                                        float rho_j) {
        const float rinv = __builtin_amdgcn_rsqf(d2);
        const float x = d2 * rinv * k.inv_h;
        const float dw = x < 0.5f ? k.c1h * x * (3.f * x - 2.f) : k.c1h * (1.f - x) * (x - 1.f);
        const float coef = ((vj.x - vi.x) * dx + (vj.y - vi.y) * dy + (vj.z - vi.z) * dz) *
                           (rinv * rinv);
        const float s = coef * (-dw * rinv) / rho_j;
        lx += s * dx;
        ly += s * dy;
        lz += s * dz;
    }
};

// 3D from the FP32 cell-corner offsets (as the density): W' is zero at h,
// so the FP32 distance decides which terms enter.  CNT (test support): only
// the number of terms into members[i].
template <bool CNT = false>
__global__ void __launch_bounds__(256) k_viscosity3f(const double4* pos, const float4* ofs,
                                                     const int32_t* cstart, const float4* vel,
                                                     const float* rho, int64_t n, KConst k,
                                                     float scale, float4* vout,
                                                     int32_t* members = nullptr) {
    const int nb = (int)((n + 255) / 256);
    const int64_t i = (int64_t)xcd_remap(blockIdx.x, nb) * 256 + threadIdx.x;
    if (i >= n) return;
    const double4 q = pos[i];
    const float4 vi = vel[i];
    const int cc0 = cell_coord(q.x, k.cell_lo[0], k.cell_inv, k.cdims[0]);
    const int cc1 = cell_coord(q.y, k.cell_lo[1], k.cell_inv, k.cdims[1]);
    const int cc2 = cell_coord(q.z, k.cell_lo[2], k.cell_inv, k.cdims[2]);
    const int x0 = cc0 > 0 ? cc0 - 1 : 0;
    const int x1 = cc0 + 1 < k.cdims[0] ? cc0 + 1 : k.cdims[0] - 1;
    const int zlo = cc2 > 0 ? cc2 - 1 : 0, zhi = cc2 + 1 < k.cdims[2] ? cc2 + 1 : k.cdims[2] - 1;
    const int ylo = cc1 > 0 ? cc1 - 1 : 0, yhi = cc1 + 1 < k.cdims[1] ? cc1 + 1 : k.cdims[1] - 1;
    // q - C_c per candidate cell, rounded once from FP64
    const float tx0 = (float)(q.x - cell_corner(k, 0, x0));
    const float tx1 = (float)(q.x - cell_corner(k, 0, x0 + 1));
    const float tx2 = (float)(q.x - cell_corner(k, 0, x0 + 2));
    const float far = 2.f * (float)k.cell_size;
    const float qo0 = (float)(q.x - cell_corner(k, 0, cc0));
    const float qo1 = (float)(q.y - cell_corner(k, 1, cc1));
    const float qo2 = (float)(q.z - cell_corner(k, 2, cc2));
    const bool qfar = !(fabsf(qo0) <= far && fabsf(qo1) <= far && fabsf(qo2) <= far);
    const float h2 = (float)k.h2;
    ViscAcc acc;
    int32_t nterms = 0;
    for (int z = zlo; z <= zhi; z++) {
        const float tz = (float)(q.z - cell_corner(k, 2, z));
        for (int y = ylo; y <= yhi; y++) {
            const float ty = (float)(q.y - cell_corner(k, 1, y));
            const int64_t row = ((int64_t)z * k.cdims[1] + y) * k.cdims[0];
            const int b = cstart[row + x0], e = cstart[row + x1 + 1];
            const int b1 = x0 + 1 <= x1 ? cstart[row + x0 + 1] : e;
            const int b2 = x0 + 2 <= x1 ? cstart[row + x0 + 2] : e;
            for (int j = b; j < e; j++) {
                const float4 p = ofs[j];
                const float tx = j < b1 ? tx0 : (j < b2 ? tx1 : tx2);
                float dx, dy, dz;   // x_j - x_i
                if (p.w >= 0.f && !qfar) {
                    dx = p.x - tx;
                    dy = p.y - ty;
                    dz = p.z - tz;
                } else {
                    const double4 pj = pos[j];
                    dx = (float)(pj.x - q.x);
                    dy = (float)(pj.y - q.y);
                    dz = (float)(pj.z - q.z);
                }
                const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                if (!(d2 > 0.f) || !(d2 < h2)) continue;   // d = 0 is skipped
                if constexpr (CNT) {
                    nterms++;
                    continue;
                }
                acc.add(k, d2, dx, dy, dz, vi, vel[j], rho[j]);
            }
        }
    }
    if constexpr (CNT) {
        members[i] = nterms;
        return;
    }
    float4 v = vi;
    v.x += scale * acc.lx;
    v.y += scale * acc.ly;
    v.z += scale * acc.lz;
    vout[i] = v;
}

__global__ void __launch_bounds__(256) k_viscosity2(const double4* pos, const int32_t* cstart,
                                                    const float4* vel, const float* rho,
                                                    int64_t n, KConst k, float scale,
                                                    float4* vout) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double4 q = pos[i];
    const float4 vi = vel[i];
    ViscAcc acc;
    for_each_candidate<2>(k, cstart, pos, q.x, q.y, q.z, [&](int j, const double4& pj) {
        const double d2 = sq_dist2(pj.x, pj.y, q.x, q.y);
        if (!(d2 > 0) || !(d2 < k.h2)) return;
        acc.add(k, (float)d2, (float)(pj.x - q.x), (float)(pj.y - q.y), 0.f, vi, vel[j],
                rho[j]);
    });
    float4 v = vi;
    v.x += scale * acc.lx;
    v.y += scale * acc.ly;
    vout[i] = v;
}

}  // namespace

int move_collide(vps_context* c, double dt, const vps_collider* col, int32_t ncol) {
    if (ncol < 0 || ncol > kMaxColliders || (ncol > 0 && !col)) {
        c->err = "move_particles: 0 to 8 colliders";
        return VPS_ERR_INVALID_ARG;
    }
    CollSet cs{};
    cs.n = ncol;
    cs.r = c->p.particle_radius;
    for (int i = 0; i < ncol; i++) {
        std::string why;
        if (!make_surfk(col[i].surface, c->dim, &cs.c[i].s, &why)) {
            c->err = "move_particles: collider " + std::to_string(i) + ": " + why;
            return VPS_ERR_INVALID_ARG;
        }
        if (!std::isfinite(col[i].restitution) || !std::isfinite(col[i].friction)) {
            c->err = "move_particles: non-finite collider coefficients";
            return VPS_ERR_INVALID_ARG;
        }
        cs.c[i].e = col[i].restitution;
        cs.c[i].mu = col[i].friction;
    }
    const int s = c->cur;
    const int64_t n = c->n;
    if (c->dim == 3)
        hipLaunchKernelGGL(k_move_collide<3>, dim3(blocks(n, 256)), dim3(256), 0, c->stream,
                           n, c->rpos[s].p, c->rvel[s].p, dt, cs);
    else
        hipLaunchKernelGGL(k_move_collide<2>, dim3(blocks(n, 256)), dim3(256), 0, c->stream,
                           n, c->rpos[s].p, c->rvel[s].p, dt, cs);
    VPS_HIP(c, hipGetLastError());
    return VPS_OK;
}

int apply_gravity(vps_context* c, double dt, double g) {
    hipLaunchKernelGGL(k_gravity, dim3(blocks(c->n, 256)), dim3(256), 0, c->stream, c->n,
                       c->rvel[c->cur].p, (float)(g * dt));
    VPS_HIP(c, hipGetLastError());
    return VPS_OK;
}

// Terms of the 3D viscosity walk per real particle (test support, sorted
// order), on offsets recomputed from the current positions.
int viscosity_member_counts(vps_context* c, int32_t* members) {
    const int64_t n = c->n;
    if (c->dim != 3) return VPS_ERR_INVALID_ARG;
    if (n == 0) return VPS_OK;
    VPS_TRY(real_offsets(c));
    const int s = c->cur;
    hipLaunchKernelGGL(k_viscosity3f<true>, dim3(blocks(n, 256)), dim3(256), 0, c->stream,
                       c->rpos[s].p, c->rofs.p, c->cell_start.p, c->rvel[s].p, c->rrho.p, n,
                       c->k, 0.f, (float4*)nullptr, members);
    VPS_HIP(c, hipGetLastError());
    return VPS_OK;
}

int apply_viscosity(vps_context* c, double dt, double nu) {
    const int64_t n = c->n;
    if (nu == 0 || n == 0) return VPS_OK;
    if (c->vel_next.n < (size_t)n) {
        dfree(c->vel_next);
        VPS_TRY(dalloc(c, c->vel_next, (size_t)n));
    }
    const int s = c->cur;
    // v += nu lapl dt with lapl = 2 (Dim + 2) m sum(...)
    const float scale = (float)(nu * (2.0 * (c->dim + 2.0) * c->p.mass) * dt);
    if (c->dim == 3)
        hipLaunchKernelGGL(k_viscosity3f<false>, dim3(blocks(n, 256)), dim3(256), 0, c->stream,
                           c->rpos[s].p, c->rofs.p, c->cell_start.p, c->rvel[s].p, c->rrho.p,
                           n, c->k, scale, c->vel_next.p);
    else
        hipLaunchKernelGGL(k_viscosity2, dim3(blocks(n, 256)), dim3(256), 0, c->stream,
                           c->rpos[s].p, c->cell_start.p, c->rvel[s].p, c->rrho.p, n, c->k,
                           scale, c->vel_next.p);
    VPS_HIP(c, hipGetLastError());
    VPS_HIP(c, hipMemcpyAsync(c->rvel[s].p, c->vel_next.p, (size_t)n * sizeof(float4),
                              hipMemcpyDeviceToDevice, c->stream));
    return VPS_OK;
}

}  // namespace vps
