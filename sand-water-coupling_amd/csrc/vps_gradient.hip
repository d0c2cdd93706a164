// vps_gradient.hip -- pressure gradient + velocity update, advection, and the
// neighbour query used by tests.
//
// Replaces VirtualParticle::applyPressureGradient
// (Cores/Structures/VirtualParticle.cpp:110-134):
//   v_i += sum_{J in N_H(i)} V_J p_J gradK_H(x_i - x_J)
//        - sum_{b in N_h(i)} V_b min(n_b.(v_i - v_b), 0) n_b W_h(x_i - x_b)
// (no dt, no 1/rho; v_i on the right is the pre-update velocity).
// Virtual particles are lattice nodes, so the virtual neighbours of x_i are
// found by scanning the node box around x_i: the node coordinates of each
// axis are formed once per particle (origin + c * spacing, as
// Grid::dataPosition rounds them) and the exact FP64 test
// ((dx^2 + dy^2) + dz^2) < H^2 decides membership.  The brick map gives
// q_J = V_J p_J (zero on non-virtual nodes).
#include "vps_internal.h"

namespace vps {
namespace {

inline unsigned blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

constexpr int kBox = 10;   // max node candidates per axis (2H/s + 2)

template <int D>
__global__ void __launch_bounds__(256) k_gradient(
    KConst k, int64_t n, const double4* rpos, float4* rvel,
    const int32_t* slot_of_brick, const float* vq, const double4* bpos,
    const float4* bvel, const float4* bnrm, const int32_t* bcstart) {
    const int nblk = (int)((n + 255) / 256);
    const int64_t i = (int64_t)xcd_remap(blockIdx.x, nblk) * 256 + threadIdx.x;
    if (i >= n) return;
    const double4 q = rpos[i];
    const double qc[3] = {q.x, q.y, q.z};
    float4 vi = rvel[i];
    float fx = 0.f, fy = 0.f, fz = 0.f;
    const double inv = 1.0 / k.spacing;
    int lo[3] = {0, 0, 0}, cnt[3] = {1, 1, 1};
    bool empty = false;
#pragma unroll
    for (int a = 0; a < D; a++) {
        int l = (int)floor((qc[a] - k.H - k.origin[a]) * inv);
        int u = (int)floor((qc[a] + k.H - k.origin[a]) * inv) + 1;
        l = l < 0 ? 0 : l;
        u = u >= k.ndims[a] ? k.ndims[a] - 1 : u;
        if (u - l + 1 > kBox) u = l + kBox - 1;  // cannot happen for H = 4s
        lo[a] = l;
        cnt[a] = u - l + 1;
        if (cnt[a] <= 0) empty = true;
    }
    if (!empty) {
        // x-axis differences and squares, kept in registers
        double dx2[kBox];
        float dxf[kBox];
#pragma unroll
        for (int c = 0; c < kBox; c++) {
            const double nx = node_coord(k.origin[0], lo[0] + c, k.spacing);
            const double dxs = dsub(nx, q.x);
            dx2[c] = dsq(dxs);
            dxf[c] = (float)dxs;
        }
        const int bx0 = lo[0] >> 3;
        int64_t brow_cached = -1;
        int s3[3] = {-1, -1, -1};
        for (int iz = 0; iz < cnt[2]; iz++) {
            const int cz = lo[2] + iz;
            double dz = 0, dz2 = 0;
            if (D == 3) {
                dz = dsub(node_coord(k.origin[2], cz, k.spacing), q.z);
                dz2 = dsq(dz);
                if (!(dz2 < k.H2)) continue;   // no node of this plane within H
            }
            for (int iy = 0; iy < cnt[1]; iy++) {
                const int cy = lo[1] + iy;
                const double dy = dsub(node_coord(k.origin[1], cy, k.spacing), q.y);
                const double dy2 = dsq(dy);
                // dx^2 >= 0, so (dx^2 + dy^2) + dz^2 >= dy^2 + dz^2 in FP64
                // as well: a row with dy^2 + dz^2 >= H^2 has no member
                if (!(dadd(dy2, dz2) < k.H2)) continue;
                const int64_t brow = ((int64_t)(D == 3 ? cz >> 3 : 0) * k.bdims[1] +
                                      (cy >> 3)) * k.bdims[0];
                const int lrow = 8 * ((cy & 7) + (D == 3 ? 8 * (cz & 7) : 0));
                constexpr int NB = D == 3 ? 512 : 64;
                // the row's <= 10 nodes lie in at most 3 bricks along x; the
                // slots change only when the row crosses into another brick row
                if (brow != brow_cached) {
                    brow_cached = brow;
#pragma unroll
                    for (int u = 0; u < 3; u++)
                        s3[u] = bx0 + u < k.bdims[0] ? slot_of_brick[brow + bx0 + u] : -1;
                }
                // issue the q loads of every node inside the sphere at once
                float qv[kBox], d2f[kBox];
#pragma unroll
                for (int c = 0; c < kBox; c++) {
                    const double d2 = D == 3 ? dadd(dadd(dx2[c], dy2), dz2)
                                             : dadd(dx2[c], dy2);
                    d2f[c] = (float)d2;
                    const int cx = lo[0] + c;
                    const int u = (cx >> 3) - bx0;
                    const int s = u == 0 ? s3[0] : (u == 1 ? s3[1] : s3[2]);
                    qv[c] = (c < cnt[0] && d2 < k.H2 && s >= 0)
                                ? vq[(int64_t)s * NB + lrow + (cx & 7)] : 0.f;
                }
                const float dyf = (float)dy, dzf = (float)dz;
#pragma unroll
                for (int c = 0; c < kBox; c++) {
                    if (qv[c] == 0.f) continue;
                    // gradK_H(x_i - x_J) = S (x_i - x_J), S = -W'(r) / r; at
                    // r = 0 the difference is zero
                    const float rinv = __builtin_amdgcn_rsqf(d2f[c]);
                    const float r = d2f[c] > 0.f ? d2f[c] * rinv : 0.f;
                    const float S = spline_s(r * k.inv_H, rinv, k.c1H, k.inv_H);
                    const float sc = qv[c] * S;
                    fx -= sc * dxf[c];
                    fy -= sc * dyf;
                    if (D == 3) fz -= sc * dzf;
                }
            }
        }
    }
    // boundary term
    for_each_candidate<D>(k, bcstart, bpos, q.x, q.y, q.z, [&](int j, const double4& pj) {
        const double d2 = D == 3 ? sq_dist3(pj.x, pj.y, pj.z, q.x, q.y, q.z)
                                 : sq_dist2(pj.x, pj.y, q.x, q.y);
        if (d2 < k.h2) {
            const KW kh = kernel_f32(d2, k.h2, k.hf, k.inv_h, k.c0h, k.c1h);
            const float4 vb = bvel[j];
            float4 nb = bnrm[j];
            if (nb.x == 0.f && nb.y == 0.f && (D == 2 || nb.z == 0.f)) {
                // a grain: the unit vector from the grain to the particle
                // (none when they coincide)
                const float inv = kh.r > 0.f ? 1.f / kh.r : 0.f;
                nb.x = (float)(q.x - pj.x) * inv;
                nb.y = (float)(q.y - pj.y) * inv;
                nb.z = D == 3 ? (float)(q.z - pj.z) * inv : 0.f;
            }
            const float dn = nb.x * (vi.x - vb.x) + nb.y * (vi.y - vb.y) +
                             (D == 3 ? nb.z * (vi.z - vb.z) : 0.f);
            const float m = dn < 0.f ? dn : 0.f;
            const float sc = vb.w * m * kh.w;
            fx -= sc * nb.x;
            fy -= sc * nb.y;
            if (D == 3) fz -= sc * nb.z;
        }
    });
    vi.x += fx;
    vi.y += fy;
    if (D == 3) vi.z += fz;
    rvel[i] = vi;
}

template <int D>
__global__ void __launch_bounds__(256) k_advect(int64_t n, double4* pos,
                                                float4* vel, double dt,
                                                double g) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double4 p = pos[i];
    float4 v = vel[i];
    p.x += (double)v.x * dt;
    p.y += (double)v.y * dt;
    if (D == 3) p.z += (double)v.z * dt;
    v.y -= (float)(g * dt);
    pos[i] = p;
    vel[i] = v;
}

// ---- neighbour query (tests) ---------------------------------------------
template <int D>
__global__ void __launch_bounds__(128) k_query(
    KConst k, int which, int64_t nq, const double* qs, const double4* pos,
    const int32_t* cstart, const int32_t* ids, const int32_t* slot_of_brick,
    const float* vvol, const int32_t* rank, const int64_t* offsets,
    int64_t* out_count, int64_t* out_idx) {
    const int64_t t = (int64_t)blockIdx.x * 128 + threadIdx.x;
    if (t >= nq) return;
    const double qx = qs[D * t], qy = qs[D * t + 1], qz = D == 3 ? qs[D * t + 2] : 0.0;
    int64_t cnt = 0;
    const int64_t base = offsets ? offsets[t] : 0;
    if (which < 2) {
        for_each_candidate<D>(k, cstart, pos, qx, qy, qz, [&](int j, const double4& pj) {
            const double d2 = D == 3 ? sq_dist3(pj.x, pj.y, pj.z, qx, qy, qz)
                                     : sq_dist2(pj.x, pj.y, qx, qy);
            if (d2 < k.h2) {
                if (out_idx) out_idx[base + cnt] = ids[j];
                cnt++;
            }
        });
    } else {
        const double qc[3] = {qx, qy, qz};
        const double inv = 1.0 / k.spacing;
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (int a = 0; a < D; a++) {
            int l = (int)floor((qc[a] - k.H - k.origin[a]) * inv);
            int u = (int)floor((qc[a] + k.H - k.origin[a]) * inv) + 1;
            lo[a] = l < 0 ? 0 : l;
            hi[a] = u >= k.ndims[a] ? k.ndims[a] - 1 : u;
        }
        const int NB = D == 3 ? 512 : 64;
        for (int cz = lo[2]; cz <= hi[2]; cz++)
            for (int cy = lo[1]; cy <= hi[1]; cy++)
                for (int cx = lo[0]; cx <= hi[0]; cx++) {
                    const double nx = node_coord(k.origin[0], cx, k.spacing);
                    const double ny = node_coord(k.origin[1], cy, k.spacing);
                    const double nz = D == 3 ? node_coord(k.origin[2], cz, k.spacing) : 0.0;
                    const double d2 = D == 3 ? sq_dist3(nx, ny, nz, qx, qy, qz)
                                             : sq_dist2(nx, ny, qx, qy);
                    if (!(d2 < k.H2)) continue;
                    const int64_t b = ((int64_t)(D == 3 ? cz >> 3 : 0) * k.bdims[1] +
                                       (cy >> 3)) * k.bdims[0] + (cx >> 3);
                    const int s = slot_of_brick[b];
                    if (s < 0) continue;
                    const int64_t g = (int64_t)s * NB + (cx & 7) + 8 * (cy & 7) +
                                      (D == 3 ? 64 * (cz & 7) : 0);
                    if (!(vvol[g] > 0.f)) continue;
                    if (out_idx) out_idx[base + cnt] = rank[g];
                    cnt++;
                }
    }
    if (out_count) out_count[t] = cnt;
}

}  // namespace

int pressure_gradient(vps_context* c) {
    const int64_t n = c->n;
    const int s = c->cur;
    const int w = c->which;
    if (c->dim == 3)
        hipLaunchKernelGGL(k_gradient<3>, dim3(blocks(n, 256)), dim3(256), 0,
                           c->stream, c->k, n, c->rpos[s].p, c->rvel[s].p,
                           c->brick_slot[w].p, c->vq.p, c->bpos.p, c->bvel.p,
                           c->bnrm.p, c->bcell_start.p);
    else
        hipLaunchKernelGGL(k_gradient<2>, dim3(blocks(n, 256)), dim3(256), 0,
                           c->stream, c->k, n, c->rpos[s].p, c->rvel[s].p,
                           c->brick_slot[w].p, c->vq.p, c->bpos.p, c->bvel.p,
                           c->bnrm.p, c->bcell_start.p);
    VPS_HIP(c, hipGetLastError());
    return VPS_OK;
}

int advect(vps_context* c, double dt, double g) {
    const int s = c->cur;
    if (c->dim == 3)
        hipLaunchKernelGGL(k_advect<3>, dim3(blocks(c->n, 256)), dim3(256), 0,
                           c->stream, c->n, c->rpos[s].p, c->rvel[s].p, dt, g);
    else
        hipLaunchKernelGGL(k_advect<2>, dim3(blocks(c->n, 256)), dim3(256), 0,
                           c->stream, c->n, c->rpos[s].p, c->rvel[s].p, dt, g);
    VPS_HIP(c, hipGetLastError());
    return VPS_OK;
}

int query_neighbors(vps_context* c, int which, int64_t nq, const double* q,
                    int64_t* offsets, int64_t* idx, int64_t cap,
                    int64_t* needed) {
    const int D = c->dim;
    DBuf<double> dq;
    DBuf<int64_t> dcnt, doff, didx;
    VPS_TRY(dalloc(c, dq, (size_t)(nq * D)));
    VPS_TRY(dalloc(c, dcnt, (size_t)nq));
    VPS_TRY(dalloc(c, doff, (size_t)nq + 1));
    VPS_HIP(c, hipMemcpyAsync(dq.p, q, (size_t)nq * D * 8, hipMemcpyHostToDevice,
                              c->stream));
    const double4* pos = which == 0 ? c->rpos[c->cur].p : c->bpos.p;
    const int32_t* cs = which == 0 ? c->cell_start.p : c->bcell_start.p;
    const int32_t* ids = which == 0 ? c->rid[c->cur].p : c->bid.p;
    const int32_t* sob = c->brick_slot[c->which].p;
    auto run = [&](const int64_t* off, int64_t* cntp, int64_t* ix) {
        if (D == 3)
            hipLaunchKernelGGL(k_query<3>, dim3(blocks(nq, 128)), dim3(128), 0,
                               c->stream, c->k, which, nq, dq.p, pos, cs, ids,
                               sob, c->vvol.p, c->vrank.p, off, cntp, ix);
        else
            hipLaunchKernelGGL(k_query<2>, dim3(blocks(nq, 128)), dim3(128), 0,
                               c->stream, c->k, which, nq, dq.p, pos, cs, ids,
                               sob, c->vvol.p, c->vrank.p, off, cntp, ix);
    };
    run(nullptr, dcnt.p, nullptr);
    std::vector<int64_t> cnt((size_t)nq);
    VPS_HIP(c, hipMemcpyAsync(cnt.data(), dcnt.p, (size_t)nq * 8,
                              hipMemcpyDeviceToHost, c->stream));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    offsets[0] = 0;
    for (int64_t i = 0; i < nq; i++) offsets[i + 1] = offsets[i] + cnt[(size_t)i];
    const int64_t total = offsets[nq];
    if (needed) *needed = total;
    int rc = VPS_OK;
    if (!idx) {
        // size query: offsets and *needed only
    } else if (total > cap) {
        rc = VPS_ERR_INVALID_ARG;
        c->err = "vps_query_neighbors: capacity too small";
    } else if (total > 0) {
        VPS_TRY(dalloc(c, didx, (size_t)total));
        VPS_HIP(c, hipMemcpyAsync(doff.p, offsets, (size_t)(nq + 1) * 8,
                                  hipMemcpyHostToDevice, c->stream));
        run(doff.p, nullptr, didx.p);
        VPS_HIP(c, hipMemcpyAsync(idx, didx.p, (size_t)total * 8,
                                  hipMemcpyDeviceToHost, c->stream));
        VPS_HIP(c, hipStreamSynchronize(c->stream));
        for (int64_t i = 0; i < nq; i++) {
            int64_t* b = idx + offsets[i];
            int64_t* e = idx + offsets[i + 1];
            for (int64_t* u = b + 1; u < e; u++) {
                const int64_t v = *u;
                int64_t* w = u - 1;
                while (w >= b && *w > v) {
                    w[1] = w[0];
                    w--;
                }
                w[1] = v;
            }
        }
    }
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    dfree(dq);
    dfree(dcnt);
    dfree(doff);
    dfree(didx);
    return rc;
}

}  // namespace vps
