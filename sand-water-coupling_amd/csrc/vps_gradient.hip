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
#include <cmath>

#include "vps_internal.h"

namespace vps {
namespace {

inline unsigned blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

constexpr int kBox = 10;   // max node candidates per axis (2H/s + 2)

// Boundary part of the update: -sum_b V_b min(n_b.(v_i - v_b), 0) n_b W_h.
template <int D>
__device__ __forceinline__ void boundary_term(const KConst& k, const double4& q,
                                              const float4& vi, const double4* bpos,
                                              const float4* bvel, const float4* bnrm,
                                              const int32_t* bcstart, float& fx,
                                              float& fy, float& fz) {
    for_each_candidate<D>(k, bcstart, bpos, q.x, q.y, q.z, [&](int j, const double4& pj) {
        const double d2 = D == 3 ? sq_dist3(pj.x, pj.y, pj.z, q.x, q.y, q.z)
                                 : sq_dist2(pj.x, pj.y, q.x, q.y);
        if (d2 < k.h2) {
            // W from the select-free spline (no division): r and 1/r from one rsq
            const float d2f = (float)d2;
            const float rinv = d2f > 0.f ? __builtin_amdgcn_rsqf(d2f) : 0.f;
            const float w = spline_w_ab(spline_ab(d2f * rinv, k.inv_h), k.c0h);
            const float4 vb = bvel[j];
            float4 nb = bnrm[j];
            if (nb.x == 0.f && nb.y == 0.f && (D == 2 || nb.z == 0.f)) {
                // a grain: the unit vector from the grain to the particle
                // (none when they coincide)
                nb.x = (float)(q.x - pj.x) * rinv;
                nb.y = (float)(q.y - pj.y) * rinv;
                nb.z = D == 3 ? (float)(q.z - pj.z) * rinv : 0.f;
            }
            const float dn = nb.x * (vi.x - vb.x) + nb.y * (vi.y - vb.y) +
                             (D == 3 ? nb.z * (vi.z - vb.z) : 0.f);
            const float m = dn < 0.f ? dn : 0.f;
            const float sc = vb.w * m * w;
            fx -= sc * nb.x;
            fy -= sc * nb.y;
            if (D == 3) fz -= sc * nb.z;
        }
    });
}


// Virtual part of the update from global memory: the node box around x_i,
// exact FP64 membership, brick map lookups.
template <int D>
__device__ __forceinline__ void virtual_term_global(const KConst& k, const double4& q,
                                                    const int32_t* slot_of_brick,
                                                    const float* vq, float& fx,
                                                    float& fy, float& fz) {
    const double qc[3] = {q.x, q.y, q.z};
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
}

template <int D>
__global__ void __launch_bounds__(256) k_gradient(
    KConst k, int64_t n, const double4* rpos, float4* rvel,
    const int32_t* slot_of_brick, const float* vq, const double4* bpos,
    const float4* bvel, const float4* bnrm, const int32_t* bcstart, SolveGuard guard) {
    const int nblk = (int)((n + 255) / 256);
    const int64_t i = (int64_t)xcd_remap(blockIdx.x, nblk) * 256 + threadIdx.x;
    if (i >= n || !solve_accepted(guard)) return;
    const double4 q = rpos[i];
    float4 vi = rvel[i];
    float fx = 0.f, fy = 0.f, fz = 0.f;
    virtual_term_global<D>(k, q, slot_of_brick, vq, fx, fy, fz);
    boundary_term<D>(k, q, vi, bpos, bvel, bnrm, bcstart, fx, fy, fz);
    vi.x += fx;
    vi.y += fy;
    if (D == 3) vi.z += fz;
    rvel[i] = vi;
}

// ---- LDS-staged 3D gradient ---------------------------------------------
// Work unit: a span of up to span_cells consecutive cells of one cell row
// (the particles of a span are contiguous in the sorted order).  The block
// stages q = V p of the node box that holds every particle's candidates
// (nodes f-3 .. f+4 per axis around f = floor((x - origin) / s): the nodes
// within H = 4s of x), zero off the virtual set and off the lattice, then
// each thread sums its particle's 8 x 8 x 8 candidates from LDS.  Nodes at
// or beyond H get S = 0 from the spline itself; the kernel gradient is
// continuous and zero at H, so the FP32 distance decides membership (a
// node within rounding of H contributes nothing measurable).  A span whose
// box exceeds the stage (particles far outside their cells) takes the
// global per-particle path.
constexpr int kGradThreads = 256;
constexpr int kGradCap = 7168;   // staged nodes (28 KB: five blocks per CU)
constexpr int kGradBoxBricks = 128;   // bricks a staged box may touch

__global__ void __launch_bounds__(256) k_spans(KConst k, int span_cells, int nspan,
                                               const int32_t* cstart, int4* chunks,
                                               int32_t* count) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t nrows = (int64_t)k.cdims[1] * k.cdims[2];
    if (i >= nrows * nspan) return;
    const int64_t row = i / nspan;
    const int sp = (int)(i - row * nspan);
    const int64_t c0 = row * k.cdims[0] + (int64_t)sp * span_cells;
    const int e = (sp + 1) * span_cells < k.cdims[0] ? (sp + 1) * span_cells : k.cdims[0];
    const int64_t c1 = row * k.cdims[0] + e;
    const int a = cstart[c0], b = cstart[c1];
    if (b > a) chunks[atomicAdd(count, 1)] = make_int4(a, b, (int)c0, (int)c1);
}

__device__ __forceinline__ int node_floor(double x, double o, double inv) {
    double f = floor((x - o) * inv);
    f = f < -1.0e9 ? -1.0e9 : (f > 1.0e9 ? 1.0e9 : f);
    return (int)f;
}

// Spans whose box exceeds the stage, from the overflow list that
// k_gradient3s fills: the global-memory path, one thread per particle.
__global__ void __launch_bounds__(kGradThreads) k_gradient_overflow(
    KConst k, const int4* chunks, const int32_t* over, const int32_t* nover,
    const double4* rpos, float4* rvel, const int32_t* slot_of_brick, const float* vq,
    const double4* bpos, const float4* bvel, const float4* bnrm, const int32_t* bcstart,
    SolveGuard guard) {
    if (!solve_accepted(guard)) return;
    const int no = *nover;
    for (int o = blockIdx.x; o < no; o += gridDim.x) {
        const int4 ch = chunks[over[o]];
        for (int p = ch.x + threadIdx.x; p < ch.y; p += kGradThreads) {
            const double4 q = rpos[p];
            const float4 vi = rvel[p];
            float fx = 0.f, fy = 0.f, fz = 0.f;
            virtual_term_global<3>(k, q, slot_of_brick, vq, fx, fy, fz);
            boundary_term<3>(k, q, vi, bpos, bvel, bnrm, bcstart, fx, fy, fz);
            float4 v = vi;
            v.x += fx;
            v.y += fy;
            v.z += fz;
            rvel[p] = v;
        }
    }
}

// CNT (test support): vq holds V_J instead of q_J, and only the number of
// virtual particles whose term enters the sum (V_J > 0, r < H in the FP32
// spline) goes into members[p]; spans beyond the stage are left out.
template <bool CNT = false>
__global__ void __launch_bounds__(kGradThreads, 4) k_gradient3s(
    KConst k, const int4* chunks, int nchunks, const double4* rpos, float4* rvel,
    const int32_t* slot_of_brick, const float* vq, const double4* bpos,
    const float4* bvel, const float4* bnrm, const int32_t* bcstart, int32_t* over,
    int32_t* nover, SolveGuard guard, int32_t* members = nullptr) {
    if (!solve_accepted(guard)) return;   // uniform over the block
    __shared__ float s_q[kGradCap];
    __shared__ int s_bslot[kGradBoxBricks];   // slots of the box's bricks
    __shared__ int s_box[6];
    __shared__ int s_walls;   // boundary particles in the cells around the span
    const int chunk = xcd_remap(blockIdx.x, nchunks);
    const int4 ch = chunks[chunk];
    const int t = threadIdx.x;
    if (t < 3) {
        s_box[t] = 0x7fffffff;
        s_box[t + 3] = -0x7fffffff;
    }
    if (t == 3) s_walls = 0;
    __syncthreads();
    if (t < 9) {
        // the boundary cells that the particles' walks would visit: the
        // span's cells +-1 along x in the 3 x 3 neighbour rows
        const int cdx = k.cdims[0], cdy = k.cdims[1], cdz = k.cdims[2];
        const int row = ch.z / cdx;
        const int cy = row % cdy, cz = row / cdy;
        const int x0 = ch.z - row * cdx, x1 = ch.w - 1 - row * cdx;
        const int yy = cy + t % 3 - 1, zz = cz + t / 3 - 1;
        if (yy >= 0 && zz >= 0 && yy < cdy && zz < cdz) {
            const int64_t rb = ((int64_t)zz * cdy + yy) * cdx;
            const int lo = x0 > 0 ? x0 - 1 : 0, hi = x1 + 1 < cdx ? x1 + 1 : cdx - 1;
            if (bcstart[rb + hi + 1] > bcstart[rb + lo]) s_walls = 1;
        }
    }
    const double inv = 1.0 / k.spacing;
    {
        int mn[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
        int mx[3] = {-0x7fffffff, -0x7fffffff, -0x7fffffff};
        for (int p = ch.x + t; p < ch.y; p += kGradThreads) {
            const double4 q = rpos[p];
            const double qc[3] = {q.x, q.y, q.z};
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const int f = node_floor(qc[a], k.origin[a], inv);
                mn[a] = min(mn[a], f);
                mx[a] = max(mx[a], f);
            }
        }
#pragma unroll
        for (int a = 0; a < 3; a++) {
            mn[a] = -wave_max<int>(-mn[a]);
            mx[a] = wave_max<int>(mx[a]);
        }
        if ((t & 63) == 0) {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                atomicMin(&s_box[a], mn[a]);
                atomicMax(&s_box[a + 3], mx[a]);
            }
        }
    }
    __syncthreads();
    const int lo0 = s_box[0] - 3, lo1 = s_box[1] - 3, lo2 = s_box[2] - 3;
    const int64_t e0 = (int64_t)s_box[3] - s_box[0] + 8;
    const int64_t e1 = (int64_t)s_box[4] - s_box[1] + 8;
    const int64_t e2 = (int64_t)s_box[5] - s_box[2] + 8;
    if (e0 * e1 * e2 > kGradCap) {
        // particles far outside their cells: left to k_gradient_overflow
        if (!CNT && t == 0) over[atomicAdd(nover, 1)] = chunk;
        return;
    }
    // the bricks the box touches (floor division by 8, also below zero)
    const int bx0 = lo0 >> 3, by0 = lo1 >> 3, bz0 = lo2 >> 3;
    const int nbx = ((lo0 + (int)e0 - 1) >> 3) - bx0 + 1;
    const int nby = ((lo1 + (int)e1 - 1) >> 3) - by0 + 1;
    const int nbz = ((lo2 + (int)e2 - 1) >> 3) - bz0 + 1;
    if (nbx * nby * nbz > kGradBoxBricks) {   // (a box of at most 108 bricks fits)
        if (!CNT && t == 0) over[atomicAdd(nover, 1)] = chunk;
        return;
    }
    // their slots (-1 off the virtual set or off the lattice) in one round
    for (int i = t; i < nbx * nby * nbz; i += kGradThreads) {
        const int bx = bx0 + i % nbx, by = by0 + (i / nbx) % nby, bz = bz0 + i / (nbx * nby);
        int sl = -1;
        if (bx >= 0 && by >= 0 && bz >= 0 && bx < k.bdims[0] && by < k.bdims[1] &&
            bz < k.bdims[2])
            sl = slot_of_brick[((int64_t)bz * k.bdims[1] + by) * k.bdims[0] + bx];
        s_bslot[i] = sl;
    }
    __syncthreads();
    {
        // one wave per box row (iy, iz), lanes along x; four rows per pass,
        // their loads issued together (no branch between them: an invalid
        // node reads q[0] and keeps 0)
        constexpr int kW = kGradThreads / 64;
        const int n0 = (int)e0, n1 = (int)e1, nrow = (int)(e1 * e2);
        const int lane = t & 63;
        for (int rw0 = t >> 6; rw0 < nrow; rw0 += 4 * kW) {
            for (int ix = lane; ix < n0; ix += 64) {
                const int gx = lo0 + ix;
                const bool x_in = gx >= 0 && gx < k.ndims[0];
                const int bxi = (gx >> 3) - bx0;
                float v[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int rw = rw0 + u * kW;
                    const int iy = rw % n1, iz = rw / n1;
                    const int gy = lo1 + iy, gz = lo2 + iz;
                    const bool in = x_in && rw < nrow && gy >= 0 && gz >= 0 &&
                                    gy < k.ndims[1] && gz < k.ndims[2];
                    const int bi = in ? (((gz >> 3) - bz0) * nby + ((gy >> 3) - by0)) * nbx + bxi
                                      : 0;
                    const int sl = s_bslot[bi];
                    const bool ok = in && sl >= 0;
                    const int64_t g = ok ? (int64_t)sl * 512 + ((gz & 7) << 6) +
                                               ((gy & 7) << 3) + (gx & 7)
                                         : 0;
                    const float q = vq[g];
                    v[u] = ok ? q : 0.f;
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int rw = rw0 + u * kW;
                    if (rw < nrow) s_q[rw * n0 + ix] = v[u];
                }
            }
        }
    }
    __syncthreads();
    // spline of H (S = -W'(r) / r, q = r / H): with a = max(1 - q, 0) and
    // b = max(1 - 2 q, 0), S = c1 (a^2 - b^2) / r on every piece:
    // c1 (2 - 3 q) / H for q < 1/2, c1 (1 - q)^2 / r up to q = 1, and 0
    // beyond (a = b = 0).  The offsets d are taken in units of H, so that
    // with w = |d|^2 = q^2 and u = 1 / q (one rsq), a = 1 - w u and
    // b = 2 a - 1 (both clamped to [0, 1]: a = 0 beyond q = 1 gives b = 0),
    // and S d = c1 (a^2 - b^2) u (d / H): 9 VALU and the rsq per candidate.
    // The clamps fold into the fmas; c1 is applied once per particle.
    const float c1 = k.c1H;
    const double iHd = 1.0 / k.H;
    for (int p0 = ch.x; p0 < ch.y; p0 += kGradThreads) {
        const int p = p0 + t;
        if (p >= ch.y) break;
        const double4 q = rpos[p];
        const float4 vi = rvel[p];
        float fx = 0.f, fy = 0.f, fz = 0.f;
        int32_t nterms = 0;
        {
            const int f0 = node_floor(q.x, k.origin[0], inv);
            const int f1 = node_floor(q.y, k.origin[1], inv);
            const int f2 = node_floor(q.z, k.origin[2], inv);
            float dxf[8], dx2[8];   // x offsets of the 8 candidates, in units of H
#pragma unroll
            for (int c = 0; c < 8; c++) {
                dxf[c] = (float)(dsub(node_coord(k.origin[0], f0 - 3 + c, k.spacing), q.x) * iHd);
                // + 1e-30: rsq stays finite at a coincident node (whose
                // direction factors are all zero)
                dx2[c] = dxf[c] * dxf[c] + 1e-30f;
            }
            // y and z offsets of candidate 0; candidate c adds c s (FP32, units of H)
            const float dy0 = (float)(dsub(node_coord(k.origin[1], f1 - 3, k.spacing), q.y) * iHd);
            const float dz0 = (float)(dsub(node_coord(k.origin[2], f2 - 3, k.spacing), q.z) * iHd);
            const float sf = (float)(k.spacing * iHd);
            const int n0 = (int)e0, n1 = (int)e1;
            const float* base = s_q + ((f2 - 3 - lo2) * n1 + (f1 - 3 - lo1)) * n0 + (f0 - 3 - lo0);
            // Candidate c of row (cy, cz) is node f - 3 + c: its offset from x
            // lies in (c - 4, c - 3] s.  With d(c) = the smallest |offset| of
            // index c (3, 2, 1, 0, 0, 1, 2, 3) a row holds members only if
            // d(cy)^2 + d(cz)^2 < 16, and only at |c - 3.5| < sqrt(R2) + 0.5,
            // R2 = 16 - d(cy)^2 - d(cz)^2: all 8 for R2 >= 11, c = 1..6 for
            // R2 = 6..8, c = 2..5 for R2 = 3 (the uniform bounds below).  The
            // whole 8 x 8 row nest is unrolled: the bounds are constants, and
            // the LDS reads are scheduled ahead of the candidates that use them.
#pragma unroll
            for (int cz = 0; cz < 8; cz++) {
                const float dzfc = fmaf((float)cz, sf, dz0), dz2c = dzfc * dzfc;
                const int dzm = cz <= 3 ? 3 - cz : cz - 4;
                const float* plane = base + cz * n1 * n0;
#pragma unroll
                for (int cy = 0; cy < 8; cy++) {
                    const int dym = cy <= 3 ? 3 - cy : cy - 4;
                    const int R2 = 16 - dym * dym - dzm * dzm;
                    if (R2 <= 0) continue;   // uniform: no member in the row
                    const int cut = R2 >= 11 ? 0 : (R2 >= 6 ? 1 : 2);
                    const float dyc = fmaf((float)cy, sf, dy0);
                    const float dyz = fmaf(dyc, dyc, dz2c);
                    if (!(dyz < 1.f)) continue;   // this particle: none within H
                    const float* row = plane + cy * n0;
                    float sx = 0.f, ss = 0.f;
#pragma unroll
                    for (int c = 0; c < 8; c++) {
                        if (c < cut || c > 7 - cut) continue;   // uniform
                        const float w = dx2[c] + dyz;
                        const float u = __builtin_amdgcn_rsqf(w);
                        const float a = __builtin_amdgcn_fmed3f(fmaf(-w, u, 1.f), 0.f, 1.f);
                        const float b = __builtin_amdgcn_fmed3f(fmaf(a, 2.f, -1.f), 0.f, 1.f);
                        const float sc = row[c] * u * fmaf(a, a, -(b * b));
                        sx = fmaf(sc, dxf[c], sx);
                        ss += sc;
                        if constexpr (CNT) nterms += (a > 0.f) & (row[c] > 0.f);
                    }
                    fx -= sx;
                    fy -= ss * dyc;
                    fz -= ss * dzfc;
                }
            }
            fx *= c1;
            fy *= c1;
            fz *= c1;
        }
        if constexpr (CNT) {
            members[p] = nterms;
            continue;
        }
        if (s_walls) boundary_term<3>(k, q, vi, bpos, bvel, bnrm, bcstart, fx, fy, fz);
        float4 v = vi;
        v.x += fx;
        v.y += fy;
        v.z += fz;
        rvel[p] = v;
    }
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

int gradient_setup(vps_context* c) {
    c->grad_span_cells = 0;
    if (c->dim != 3) return VPS_OK;
    const KConst& k = c->k;
    // candidates f-3 .. f+4 per axis assume H = 4 s
    if (!(std::fabs(k.H - 4.0 * k.spacing) <= 1e-12 * k.H)) return VPS_OK;
    const double ratio = k.cell_size / k.spacing;
    const int64_t eyz = (int64_t)std::floor(ratio) + 9;   // box extent along y, z
    const double room = (double)kGradCap / (double)(eyz * eyz) - 9.0;
    int span = (int)std::floor(room / ratio);
    if (span < 1) return VPS_OK;
    if (span > k.cdims[0]) span = k.cdims[0];
    c->grad_span_cells = span;
    c->grad_nspan = (k.cdims[0] + span - 1) / span;
    const int64_t nrows = (int64_t)k.cdims[1] * k.cdims[2];
    VPS_TRY(dalloc(c, c->grad_chunks, (size_t)(nrows * c->grad_nspan)));
    VPS_TRY(dalloc(c, c->grad_over, (size_t)(nrows * c->grad_nspan)));
    return VPS_OK;
}

int gradient_spans(vps_context* c) {
    if (c->grad_span_cells <= 0) return VPS_OK;
    const int64_t total = (int64_t)c->k.cdims[1] * c->k.cdims[2] * c->grad_nspan;
    hipLaunchKernelGGL(k_spans, dim3(blocks(total, 256)), dim3(256), 0, c->stream, c->k,
                       c->grad_span_cells, c->grad_nspan, c->cell_start.p,
                       c->grad_chunks.p, (int32_t*)(c->counters.p + 2));
    VPS_HIP(c, hipGetLastError());
    return VPS_OK;
}

// Terms of the staged gradient's virtual walk per real particle (test
// support, sorted order; -1 stays where a span exceeds the stage).
int gradient_member_counts(vps_context* c, int32_t* members) {
    if (c->dim != 3 || c->grad_span_cells <= 0 || c->nslots == 0) return VPS_ERR_INVALID_ARG;
    if (c->grad_nchunks > 0)
        hipLaunchKernelGGL(k_gradient3s<true>, dim3((unsigned)c->grad_nchunks),
                           dim3(kGradThreads), 0, c->stream, c->k, c->grad_chunks.p,
                           (int)c->grad_nchunks, c->rpos[c->cur].p, c->rvel[c->cur].p,
                           c->brick_slot[c->which].p, c->vvol.p, c->bpos.p, c->bvel.p,
                           c->bnrm.p, c->bcell_start.p, (int32_t*)nullptr,
                           (int32_t*)nullptr, SolveGuard{}, members);
    VPS_HIP(c, hipGetLastError());
    return VPS_OK;
}

int pressure_gradient(vps_context* c, SolveGuard g) {
    const int64_t n = c->n;
    const int s = c->cur;
    const int w = c->which;
    if (c->dim == 3 && c->grad_span_cells > 0) {
        if (c->grad_nchunks > 0) {
            int32_t* nover = (int32_t*)(c->counters.p + 3);   // zeroed with the spans
            int32_t* over = c->grad_over.p;
            hipLaunchKernelGGL(k_gradient3s<false>, dim3((unsigned)c->grad_nchunks),
                               dim3(kGradThreads), 0, c->stream, c->k,
                               c->grad_chunks.p, (int)c->grad_nchunks, c->rpos[s].p,
                               c->rvel[s].p, c->brick_slot[w].p, c->vq.p, c->bpos.p,
                               c->bvel.p, c->bnrm.p, c->bcell_start.p, over, nover, g);
            hipLaunchKernelGGL(k_gradient_overflow, dim3(64), dim3(kGradThreads), 0,
                               c->stream, c->k, c->grad_chunks.p, over, nover,
                               c->rpos[s].p, c->rvel[s].p, c->brick_slot[w].p, c->vq.p,
                               c->bpos.p, c->bvel.p, c->bnrm.p, c->bcell_start.p, g);
        }
    } else if (c->dim == 3)
        hipLaunchKernelGGL(k_gradient<3>, dim3(blocks(n, 256)), dim3(256), 0,
                           c->stream, c->k, n, c->rpos[s].p, c->rvel[s].p,
                           c->brick_slot[w].p, c->vq.p, c->bpos.p, c->bvel.p,
                           c->bnrm.p, c->bcell_start.p, g);
    else
        hipLaunchKernelGGL(k_gradient<2>, dim3(blocks(n, 256)), dim3(256), 0,
                           c->stream, c->k, n, c->rpos[s].p, c->rvel[s].p,
                           c->brick_slot[w].p, c->vq.p, c->bpos.p, c->bvel.p,
                           c->bnrm.p, c->bcell_start.p, g);
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
