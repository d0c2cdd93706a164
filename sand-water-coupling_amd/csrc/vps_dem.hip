// vps_dem.hip -- GPU step of the DEM sand grains and their coupling to the
// water projection as moving boundary particles.
//
// Replaces DEMParticleSand<Dim>::advance (Cores/Physics/DEMParticleSand.cpp:
// 89-152) with the contact model of DEMParticle<Dim>
// (Cores/Structures/DEMParticle.cpp:13-134):
//   moveParticles        x += v dt
//   applyPressureForce   v += dt * sum_{j: |x_j - x_i| < 2r} F(x_j - x_i, v_j - v_i)
//                          + sum_{b: |x_b - x_i| < 4r} (F_norm + F_tang)
//   applyExternalForces  v.y -= g dt
// with the soft-sphere contact F (penetration 2r - |d| > 0, n = d / max(|d|,
// 1e-4 r), normal K_n pen n, tangential -K_t v_t limited to tan(phi) |F_n|,
// F = -F_n - F_t) and the wall force F_norm = -V_b min(n_b.(v_i - v_b), 0)
// n_b W_4r, F_tang = -K_t (dv - (n_b.dv) n_b) limited to tan(phi) |F_norm|
// (used as is when |F_tang| <= 1e-6).  The grains are the context's real
// particles (kernel radius h = 4r: grain-grain pairs are the h-walk's
// candidates within 2r), the walls its boundary particles.
//
// The capillary bridge force (DEMParticle.cpp:97-125) acts only for gaps
// 1e-6 < |d| - 2r < d_rupture, i.e. beyond the 2r neighbourhood the reference
// searches, so it never contributes there and is not evaluated.  The
// reference updates velocities in place while other threads read them (an
// OpenMP race); here every force uses the velocities at the start of the pass.
#include <cmath>

#include "vps_internal.h"

namespace vps {
namespace {

inline unsigned blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <int D>
__global__ void __launch_bounds__(256) k_dem_move(int64_t n, double4* pos,
                                                  const float4* vel, double dt) {
#pragma clang fp contract(off)   // x + v dt rounded twice, as the reference
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double4 p = pos[i];
    const float4 v = vel[i];
    p.x += (double)v.x * dt;
    p.y += (double)v.y * dt;
    if (D == 3) p.z += (double)v.z * dt;
    pos[i] = p;
}

struct DemK {
    float k_norm, k_tang, tan_phi;
    float two_r;         // contact distance 2r
    float min_dist;      // 1e-4 r
    double two_r2;       // (2r)^2, the searcher's strict radius
    float dt, g_dt;
};

template <int D>
__global__ void __launch_bounds__(256) k_dem_forces(
    KConst k, DemK dk, int64_t n, const double4* pos, const float4* ofs, const float4* vel,
    const int32_t* cstart, const double4* bpos, const float4* bvel,
    const float4* bnrm, const int32_t* bcstart, float4* vout) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double4 q = pos[i];
    const float4 vi = vel[i];
    float cx = 0.f, cy = 0.f, cz = 0.f;   // sum of contact forces
    auto contact = [&](int j, const double4& pj) {
        const double d2 = D == 3 ? sq_dist3(pj.x, pj.y, pj.z, q.x, q.y, q.z)
                                 : sq_dist2(pj.x, pj.y, q.x, q.y);
        if (!(d2 < dk.two_r2) || j == i) return;   // self: zero force
        // d = x_j - x_i, v = v_j - v_i
        const float dx = (float)(pj.x - q.x), dy = (float)(pj.y - q.y),
                    dz = D == 3 ? (float)(pj.z - q.z) : 0.f;
        float dist = sqrtf((float)d2);
        const float pen = dk.two_r - dist;
        if (!(pen > 0.f)) return;
        if (dist <= dk.min_dist) dist = dk.min_dist;
        const float inv = 1.f / dist;
        const float nx = dx * inv, ny = dy * inv, nz = dz * inv;
        const float4 vj = vel[j];
        const float wx = vj.x - vi.x, wy = vj.y - vi.y, wz = D == 3 ? vj.z - vi.z : 0.f;
        const float vn = wx * nx + wy * ny + wz * nz;
        const float tx = wx - vn * nx, ty = wy - vn * ny, tz = wz - vn * nz;
        const float fn = dk.k_norm * pen;   // |normal force|
        float sx = -dk.k_tang * tx, sy = -dk.k_tang * ty, sz = -dk.k_tang * tz;
        const float max_fs = fabsf(fn) * dk.tan_phi;
        const float fs = sqrtf(sx * sx + sy * sy + sz * sz);
        if (fs > max_fs) {
            const float s = max_fs / fs;
            sx *= s;
            sy *= s;
            sz *= s;
        }
        cx += -fn * nx - sx;
        cy += -fn * ny - sy;
        cz += -fn * nz - sz;
    };
    if constexpr (D == 3) {
        // Two passes per row: the FP32 distance from the cell offsets keeps
        // the candidates within 2r (margin 1e-4 of (2r)^2, far above its
        // error; particles far from their cell always), then the FP64 test
        // and the force for those only, in candidate order (about 2% of the
        // 27-cell candidates are contacts).
        const int cc0 = cell_coord(q.x, k.cell_lo[0], k.cell_inv, k.cdims[0]);
        const int cc1 = cell_coord(q.y, k.cell_lo[1], k.cell_inv, k.cdims[1]);
        const int cc2 = cell_coord(q.z, k.cell_lo[2], k.cell_inv, k.cdims[2]);
        const int x0 = cc0 > 0 ? cc0 - 1 : 0;
        const int x1 = cc0 + 1 < k.cdims[0] ? cc0 + 1 : k.cdims[0] - 1;
        const int zlo = cc2 > 0 ? cc2 - 1 : 0, zhi = cc2 + 1 < k.cdims[2] ? cc2 + 1 : k.cdims[2] - 1;
        const int ylo = cc1 > 0 ? cc1 - 1 : 0, yhi = cc1 + 1 < k.cdims[1] ? cc1 + 1 : k.cdims[1] - 1;
        float tx0 = (float)(q.x - cell_corner(k, 0, x0));
        float tx1 = (float)(q.x - cell_corner(k, 0, x0 + 1));
        float tx2 = (float)(q.x - cell_corner(k, 0, x0 + 2));
        asm volatile("" : "+v"(tx0), "+v"(tx1), "+v"(tx2));
        const float far = 2.f * (float)k.cell_size;
        const bool qfar = !(fabsf((float)(q.x - cell_corner(k, 0, cc0))) <= far &&
                            fabsf((float)(q.y - cell_corner(k, 1, cc1))) <= far &&
                            fabsf((float)(q.z - cell_corner(k, 2, cc2))) <= far);
        const float thr = (float)(dk.two_r2 * 1.0001);
        for (int z = zlo; z <= zhi; z++) {
            const float tz = (float)(q.z - cell_corner(k, 2, z));
            for (int y = ylo; y <= yhi; y++) {
                const float ty = (float)(q.y - cell_corner(k, 1, y));
                const int64_t row = ((int64_t)z * k.cdims[1] + y) * k.cdims[0];
                const int b = cstart[row + x0], e = cstart[row + x1 + 1];
                const int b1 = x0 + 1 <= x1 ? cstart[row + x0 + 1] : e;
                const int b2 = x0 + 2 <= x1 ? cstart[row + x0 + 2] : e;
                for (int c0 = b; c0 < e; c0 += 32) {
                    const int c1 = c0 + 32 < e ? c0 + 32 : e;
                    uint32_t m = 0u;
                    int nbits = 0;
                    for (int j0 = c0; j0 < c1; j0 += 4, nbits += 4) {
                        float4 pv[4];
#pragma unroll
                        for (int u = 0; u < 4; u++)
                            pv[u] = j0 + u < c1 ? ofs[j0 + u] : make_float4(1e30f, 0.f, 0.f, 1.f);
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            const int j = j0 + u;
                            const float tx = j < b1 ? tx0 : (j < b2 ? tx1 : tx2);
                            const float dx = tx - pv[u].x, dy = ty - pv[u].y, dz = tz - pv[u].z;
                            const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                            const bool cand = j < c1 && (d2 < thr || pv[u].w < 0.f || qfar);
                            m = m + m + (uint32_t)cand;
                        }
                    }
                    while (m) {
                        const int bit = 31 - __builtin_clz(m);
                        m ^= 1u << bit;
                        const int j = c0 + nbits - 1 - bit;
                        contact(j, pos[j]);
                    }
                }
            }
        }
    } else {
        for_each_candidate<D>(k, cstart, pos, q.x, q.y, q.z, contact);
    }
    float fx = cx * dk.dt, fy = cy * dk.dt, fz = cz * dk.dt;
    // walls within 4r
    for_each_candidate<D>(k, bcstart, bpos, q.x, q.y, q.z, [&](int j, const double4& pj) {
        const double d2 = D == 3 ? sq_dist3(pj.x, pj.y, pj.z, q.x, q.y, q.z)
                                 : sq_dist2(pj.x, pj.y, q.x, q.y);
        if (!(d2 < k.h2)) return;
        const KW kh = kernel_f32(d2, k.h2, k.hf, k.inv_h, k.c0h, k.c1h);
        const float4 vb = bvel[j];   // (v_b, V_b)
        const float4 nb = bnrm[j];   // (n_b, rho_b)
        const float ux = vi.x - vb.x, uy = vi.y - vb.y, uz = D == 3 ? vi.z - vb.z : 0.f;
        const float un = nb.x * ux + nb.y * uy + (D == 3 ? nb.z * uz : 0.f);
        const float m = un < 0.f ? un : 0.f;
        const float an = -vb.w * m * kh.w;
        const float Fx = an * nb.x, Fy = an * nb.y, Fz = D == 3 ? an * nb.z : 0.f;
        float Tx = -dk.k_tang * (ux - un * nb.x), Ty = -dk.k_tang * (uy - un * nb.y),
              Tz = D == 3 ? -dk.k_tang * (uz - un * nb.z) : 0.f;
        const float tn = sqrtf(Tx * Tx + Ty * Ty + Tz * Tz);
        if (tn > 1e-6f) {
            const float lim = sqrtf(Fx * Fx + Fy * Fy + Fz * Fz) * dk.tan_phi;
            if (!(lim > tn)) {
                const float s = lim / tn;
                Tx *= s;
                Ty *= s;
                Tz *= s;
            }
        }
        fx += Fx + Tx;
        fy += Fy + Ty;
        fz += Fz + Tz;
    });
    float4 v = vi;
    v.x += fx;
    v.y += fy;
    if (D == 3) v.z += fz;
    v.y -= dk.g_dt;
    vout[i] = v;
}

// Grain positions / velocities as moving boundary particles of a water
// context: the walls first (FP64 AoS from the host copy), then the grains in
// their caller order (boundary index = wall count + grain index) with a zero
// normal (the water side takes the unit vector from the grain to the fluid
// particle).
template <int D>
__global__ void __launch_bounds__(256) k_pack_grains(int64_t n, const double4* gpos,
                                                     const float4* gvel,
                                                     const int32_t* gid, int64_t off,
                                                     double* x, double* nrm, double* v) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double4 p = gpos[i];
    const float4 u = gvel[i];
    const int64_t b = off + gid[i];
    const double pc[3] = {p.x, p.y, p.z};
    const double uc[3] = {(double)u.x, (double)u.y, (double)u.z};
    for (int a = 0; a < D; a++) {
        x[b * D + a] = pc[a];
        nrm[b * D + a] = 0.0;
        v[b * D + a] = uc[a];
    }
}

}  // namespace

int dem_step(vps_context* c, double dt, double g, const vps_dem_params* dp) {
    const int D = c->dim;
    const int64_t n = c->n;
    int s = c->cur;
    if (D == 3)
        hipLaunchKernelGGL(k_dem_move<3>, dim3(blocks(n, 256)), dim3(256), 0, c->stream,
                           n, c->rpos[s].p, c->rvel[s].p, dt);
    else
        hipLaunchKernelGGL(k_dem_move<2>, dim3(blocks(n, 256)), dim3(256), 0, c->stream,
                           n, c->rpos[s].p, c->rvel[s].p, dt);
    VPS_TRY(cells_sort_real(c));
    s = c->cur;
    if (D == 3) VPS_TRY(real_offsets(c));
    if (c->vel_next.n < (size_t)n) {
        dfree(c->vel_next);
        VPS_TRY(dalloc(c, c->vel_next, (size_t)n));
    }
    const double r = c->p.particle_radius;
    DemK dk;
    dk.k_norm = (float)(dp->young * r);                  // K_norm = Young r
    dk.k_tang = (float)(dp->young * r * dp->poisson);    // K_tang = K_norm Poisson
    dk.tan_phi = (float)std::tan(dp->friction_angle);
    dk.two_r = (float)(2 * r);
    dk.min_dist = (float)(1e-4 * r);
    dk.two_r2 = (2 * r) * (2 * r);
    dk.dt = (float)dt;
    dk.g_dt = (float)(g * dt);
    const double4* bp = c->nb ? c->bpos.p : nullptr;
    const int32_t* bcs = c->nb ? c->bcell_start.p : c->cell_start.p;
    // with no walls the boundary walk must see empty cells
    if (c->nb == 0) {
        if (!c->bcell_start.p) VPS_TRY(dalloc(c, c->bcell_start, (size_t)c->ncells + 1));
        VPS_HIP(c, hipMemsetAsync(c->bcell_start.p, 0, ((size_t)c->ncells + 1) * 4,
                                  c->stream));
        bcs = c->bcell_start.p;
    }
    if (D == 3)
        hipLaunchKernelGGL(k_dem_forces<3>, dim3(blocks(n, 256)), dim3(256), 0, c->stream,
                           c->k, dk, n, c->rpos[s].p, c->rofs.p, c->rvel[s].p, c->cell_start.p,
                           bp, c->bvel.p, c->bnrm.p, bcs, c->vel_next.p);
    else
        hipLaunchKernelGGL(k_dem_forces<2>, dim3(blocks(n, 256)), dim3(256), 0, c->stream,
                           c->k, dk, n, c->rpos[s].p, (const float4*)nullptr, c->rvel[s].p,
                           c->cell_start.p,
                           bp, c->bvel.p, c->bnrm.p, bcs, c->vel_next.p);
    VPS_HIP(c, hipGetLastError());
    VPS_HIP(c, hipMemcpyAsync(c->rvel[s].p, c->vel_next.p, (size_t)n * sizeof(float4),
                              hipMemcpyDeviceToDevice, c->stream));
    c->generated = false;
    return VPS_OK;
}

// Boundary of the water context = its walls + the grains of the DEM context.
int couple_grains(vps_context* w, vps_context* g) {
    const int D = w->dim;
    const int64_t nw = (int64_t)(w->host_bx.size() / D);
    const int64_t ng = g->n;
    const int64_t nt = nw + ng;
    // staging (x, n, v) kept across calls; the walls are copied once
    const size_t need = (size_t)(3 * nt * D);
    if (w->couple_stage.n != need || w->couple_walls != nw) {
        dfree(w->couple_stage);
        VPS_TRY(dalloc(w, w->couple_stage, need));
        w->couple_walls = -1;
    }
    double* dx = w->couple_stage.p;
    double* dn = dx + nt * D;
    double* dv = dn + nt * D;
    if (w->couple_walls != nw && nw > 0) {
        VPS_HIP(w, hipMemcpyAsync(dx, w->host_bx.data(), (size_t)nw * D * 8,
                                  hipMemcpyHostToDevice, w->stream));
        VPS_HIP(w, hipMemcpyAsync(dn, w->host_bn.data(), (size_t)nw * D * 8,
                                  hipMemcpyHostToDevice, w->stream));
        VPS_HIP(w, hipMemcpyAsync(dv, w->host_bv.data(), (size_t)nw * D * 8,
                                  hipMemcpyHostToDevice, w->stream));
    }
    w->couple_walls = nw;
    // the grains are read on the water stream once the DEM stream is done
    VPS_HIP(g, hipStreamSynchronize(g->stream));
    if (ng > 0) {
        const int s = g->cur;
        if (D == 3)
            hipLaunchKernelGGL(k_pack_grains<3>, dim3(blocks(ng, 256)), dim3(256), 0,
                               w->stream, ng, g->rpos[s].p, g->rvel[s].p, g->rid[s].p,
                               nw, dx, dn, dv);
        else
            hipLaunchKernelGGL(k_pack_grains<2>, dim3(blocks(ng, 256)), dim3(256), 0,
                               w->stream, ng, g->rpos[s].p, g->rvel[s].p, g->rid[s].p,
                               nw, dx, dn, dv);
        VPS_HIP(w, hipGetLastError());
    }
    VPS_TRY(upload_boundary_device(w, nt, dx, dn, dv));
    VPS_HIP(w, hipStreamSynchronize(w->stream));
    return VPS_OK;
}

}  // namespace vps

using namespace vps;

extern "C" {

int vps_dem_step(vps_context* c, double dt, double gravity,
                 const vps_dem_params* dp) {
    if (!c || !dp) return VPS_ERR_INVALID_ARG;
    if (c->n == 0) {
        c->err = "dem step: no grains (vps_set_particles)";
        return VPS_ERR_STATE;
    }
    if (!std::isfinite(dt) || !std::isfinite(gravity) || !(dt >= 0) ||
        !std::isfinite(dp->young) || !std::isfinite(dp->poisson) ||
        !std::isfinite(dp->friction_angle) || !(dp->young > 0)) {
        c->err = "dem step: dt >= 0, finite gravity and parameters, young > 0";
        return VPS_ERR_INVALID_ARG;
    }
    VPS_TRY(dem_step(c, dt, gravity, dp));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    return VPS_OK;
}

int vps_couple_grains(vps_context* water, vps_context* grains) {
    if (!water || !grains) return VPS_ERR_INVALID_ARG;
    if (water->dim != grains->dim) {
        water->err = "couple grains: dimensions differ";
        return VPS_ERR_INVALID_ARG;
    }
    if (water->device != grains->device) {
        water->err = "couple grains: contexts on different devices";
        return VPS_ERR_INVALID_ARG;
    }
    return couple_grains(water, grains);
}

}  // extern "C"
