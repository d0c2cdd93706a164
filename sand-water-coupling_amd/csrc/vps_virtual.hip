// vps_virtual.hip -- virtual particle generation and the fused virtual sums.
//
// Replaces VirtualParticle::generateParticles / computeDensities /
// computeVolumes / calculateVelocity / calculateDiv
// (Cores/Structures/VirtualParticle.cpp:18-62, 136-147, 171-201).
//
// The reference evaluates the weight w_I = sum_j W_h(x_I - x_j) at every node
// of the lattice (6.9M nodes at 1M particles) and keeps the nodes with
// w_I > 1e-6, in x-fastest node order.  Here:
//   1. every real particle flags the 8^D-node bricks its h-ball touches;
//   2. flagged bricks are compacted into slots (ordered by brick id);
//   3. one workgroup per slot evaluates, per node, in ONE walk over the real
//      and boundary neighbours within h: w_I (FP32, with an exact FP64 redo
//      when it lies within 1e-4 relative of the 1e-6 threshold), rho_I,
//      V_I, the smoothed velocity v_I and the divergence D_I (the CG
//      right-hand side).  Folding calculateVelocity/calculateDiv into
//      generation saves a full neighbour walk; project then uses the
//      velocities as of generation, which is also when the reference reads
//      them (nothing changes velocities between generateParticles and
//      project, DualParticleSphLiquid.cpp:118-125);
//   4. the reference order (x-fastest node rank) comes from a lattice bitmask
//      and a popcount scan, so every node knows its index in the reference's
//      virtual list without the list ever being materialised.
#include "vps_internal.h"

namespace vps {
namespace {

inline unsigned blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Marks the bricks within h of each particle and clears the lattice bitmask
// of the previous generation (grid-stride over its words).
template <int D>
__global__ void __launch_bounds__(256) k_mark_bricks(const double4* pos,
                                                     int64_t n, KConst k,
                                                     uint32_t* flag, uint32_t* bits,
                                                     int64_t nwords, int32_t* span_count) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    for (int64_t w = i; w < nwords; w += (int64_t)gridDim.x * 256) bits[w] = 0u;
    if (i == 0) {   // the gradient's span and overflow lists are rebuilt next
        span_count[0] = 0;
        span_count[2] = 0;
    }
    if (i >= n) return;
    const double4 p = pos[i];
    const double pc[3] = {p.x, p.y, p.z};
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < D; a++) {
        const double inv = 1.0 / k.spacing;
        int l = (int)floor((pc[a] - k.h - k.origin[a]) * inv);
        int u = (int)floor((pc[a] + k.h - k.origin[a]) * inv) + 1;
        l = l < 0 ? 0 : l;
        u = u >= k.ndims[a] ? k.ndims[a] - 1 : u;
        if (l > u) return;  // ball outside the lattice
        lo[a] = l / kBrick;
        hi[a] = u / kBrick;
    }
    for (int bz = lo[2]; bz <= hi[2]; bz++)
        for (int by = lo[1]; by <= hi[1]; by++)
            for (int bx = lo[0]; bx <= hi[0]; bx++)
                flag[((int64_t)bz * k.bdims[1] + by) * k.bdims[0] + bx] = 1u;
}

__global__ void __launch_bounds__(256) k_assign_slots(uint32_t* flag,
                                                      int64_t nbricks,
                                                      int32_t* slot_of_brick,
                                                      int32_t* brick_of_slot) {
    const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (b >= nbricks) return;
    // slot_of_brick holds the exclusive prefix of flag on entry; the flag is
    // cleared for the next generation
    const int s = slot_of_brick[b];
    if (flag[b]) {
        brick_of_slot[s] = (int32_t)b;
        flag[b] = 0u;
    } else {
        slot_of_brick[b] = -1;
    }
}

// Real-neighbour sums of one node: w, sum W_H, sum V^2 W, sum V W (the v
// denominator), sum v W / rho and the divergence terms.  dx = q - x_j.
// Every sum is linear in its spline's constant (c0 of W, c1 of S = -W'/r),
// so the terms are summed without it and finish_node applies c0h, c0H and
// c1h once per node (one multiply fewer per term); the boundary sums alike.
struct RealAcc {
    float w = 0.f, sH = 0.f, sV = 0.f, den = 0.f;
    float nvx = 0.f, nvy = 0.f, nvz = 0.f;   // sum v_j W / rho_j
    float gx = 0.f, gy = 0.f, gz = 0.f;      // sum V_j gradK
    float dr = 0.f;                           // sum V_j v_j . gradK
    template <int D>
    __device__ __forceinline__ void add(const KConst& k, double d2, float dx,
                                        float dy, float dz, float4 vj,
                                        float inv_rhoj) {
        const KW kh = kernel_f32(d2, k.h2, k.hf, k.inv_h, 1.f, 1.f);
        const KW kH = kernel_f32(d2, k.H2, k.Hf, k.inv_H, 1.f, 1.f);
        const float Vj = vj.w;
        const float wr = kh.w * inv_rhoj;
        w += kh.w;
        sH += kH.w;
        sV += Vj * Vj * kh.w;
        den += Vj * kh.w;
        nvx += vj.x * wr;
        nvy += vj.y * wr;
        nvz += vj.z * wr;
        if (kh.r > 0.f) {
            // gradK(q - x_j) = -W'(r) (q - x_j) / r
            const float s = -kh.dw / kh.r;
            const float ex = dx * s, ey = dy * s, ez = D == 3 ? dz * s : 0.f;
            gx += Vj * ex;
            gy += Vj * ey;
            gz += Vj * ez;
            dr += Vj * (vj.x * ex + vj.y * ey + vj.z * ez);
        }
    }
    // The same sums from an FP32 squared distance d2f < h^2 (0 < d2f), with
    // the division-free spline.
    template <int D>
    __device__ __forceinline__ void add_fast(const KConst& k, float d2f, float dx,
                                             float dy, float dz, float4 vj,
                                             float inv_rhoj) {
        const float rinv = __builtin_amdgcn_rsqf(d2f);
        const float r = d2f * rinv;
        const SplineAB sp = spline_ab(r, k.inv_h);
        const float wh = spline_w_ab(sp, 1.f);
        const float sh = spline_s_ab(sp, rinv, 1.f);
        const float Vj = vj.w;
        const float wr = wh * inv_rhoj;
        w += wh;
        sH += spline_w_ab(spline_ab(r, k.inv_H), 1.f);
        sV += Vj * Vj * wh;
        den += Vj * wh;
        nvx += vj.x * wr;
        nvy += vj.y * wr;
        nvz += vj.z * wr;
        const float ex = dx * sh, ey = dy * sh, ez = D == 3 ? dz * sh : 0.f;
        gx += Vj * ex;
        gy += Vj * ey;
        gz += Vj * ez;
        dr += Vj * (vj.x * ex + vj.y * ey + vj.z * ez);
    }
};

// Boundary sums of one node: sum V_b W, sum v_b W / rho_b, sum V_b gradK,
// sum V_b v_b . gradK (VirtualParticle.cpp:76-107).
struct BAcc {
    float den = 0.f, bnx = 0.f, bny = 0.f, bnz = 0.f;
    float hx = 0.f, hy = 0.f, hz = 0.f, db = 0.f;
    // boundary particle at pj within h (FP64 test passed): accurate spline
    template <int D>
    __device__ __forceinline__ void add_exact(const KConst& k, double qx, double qy,
                                              double qz, const double4& pj, double d2,
                                              const float4& vj, float rho_b) {
        const KW kh = kernel_f32(d2, k.h2, k.hf, k.inv_h, 1.f, 1.f);
        const float Vj = vj.w;
        const float wr = kh.w / rho_b;
        den += Vj * kh.w;
        bnx += vj.x * wr;
        bny += vj.y * wr;
        bnz += vj.z * wr;
        if (kh.r > 0.f) {
            const float s = -kh.dw / kh.r;
            const float ex = (float)(qx - pj.x) * s;
            const float ey = (float)(qy - pj.y) * s;
            const float ez = D == 3 ? (float)(qz - pj.z) * s : 0.f;
            hx += Vj * ex;
            hy += Vj * ey;
            hz += Vj * ez;
            db += Vj * (vj.x * ex + vj.y * ey + vj.z * ez);
        }
    }
    // the same from an FP32 squared distance 0 < d2f < h^2 and d = q - x_b,
    // division-free (as the real sums)
    __device__ __forceinline__ void add_fast(const KConst& k, float d2f, float dx, float dy,
                                             float dz, const float4& vj, float inv_rho_b) {
        const float rinv = __builtin_amdgcn_rsqf(d2f);
        const SplineAB sp = spline_ab(d2f * rinv, k.inv_h);
        const float wh = spline_w_ab(sp, 1.f);
        const float sh = spline_s_ab(sp, rinv, 1.f);
        const float Vj = vj.w;
        const float wr = wh * inv_rho_b;
        den += Vj * wh;
        bnx += vj.x * wr;
        bny += vj.y * wr;
        bnz += vj.z * wr;
        const float ex = dx * sh, ey = dy * sh, ez = dz * sh;
        hx += Vj * ex;
        hy += Vj * ey;
        hz += Vj * ez;
        db += Vj * (vj.x * ex + vj.y * ey + vj.z * ez);
    }
};

// Everything after the real walk: the exact threshold decision, the boundary
// walk, v_I, D_I and the outputs.
// Boundary walk of one node from the global cell table (FP64 throughout).
template <int D>
struct BoundaryWalkGlobal {
    const int32_t* bcstart;
    const double4* bpos;
    __device__ __forceinline__ void operator()(const KConst& k, double qx, double qy,
                                               double qz, const float4* bvel,
                                               const float4* bnrm, BAcc& b) const {
        for_each_candidate<D>(k, bcstart, bpos, qx, qy, qz, [&](int j, const double4& pj) {
            const double d2 = D == 3 ? sq_dist3(pj.x, pj.y, pj.z, qx, qy, qz)
                                     : sq_dist2(pj.x, pj.y, qx, qy);
            if (d2 < k.h2) b.add_exact<D>(k, qx, qy, qz, pj, d2, bvel[j], bnrm[j].w);
        });
    }
};

template <int D, typename BW>
__device__ __forceinline__ void finish_node(
    const KConst& k, const RealAcc& a, double qx, double qy, double qz,
    int cx, int cy, int cz, int64_t out, const double4* rpos,
    const int32_t* rcstart, const BW& bwalk, const float4* bvel,
    const float4* bnrm, float* vrho, float* vvol, float* vb, uint32_t* bits) {
    // the sums carry no spline constants (see RealAcc)
    const float aw = a.w * k.c0h;
    bool virt = aw > 1e-6f;
    if (fabsf(aw - 1e-6f) <= 1e-10f) {
        // exact FP64 decision near the threshold (VirtualParticle.cpp:56)
        double w64 = 0;
        for_each_candidate<D>(k, rcstart, rpos, qx, qy, qz, [&](int j, const double4& pj) {
            const double d2 = D == 3 ? sq_dist3(pj.x, pj.y, pj.z, qx, qy, qz)
                                     : sq_dist2(pj.x, pj.y, qx, qy);
            if (d2 < k.h2) w64 += kernel_w64(d2, k.inv_h64, k.c0h64);
        });
        virt = w64 > 1e-6;
    }
    if (!virt) {
        vrho[out] = 0.f;
        vvol[out] = 0.f;
        vb[out] = 0.f;
        return;
    }
    // boundary neighbours within h
    BAcc b;
    b.den = a.den;
    bwalk(k, qx, qy, qz, bvel, bnrm, b);
    // v_I = (m sum v_j W/rho_j + m_b sum v_b W/rho_b) / (sum V W + sum V_b W + 1e-6)
    const float inv_den = 1.f / fmaf(b.den, k.c0h, 1e-6f);
    const float mc = k.mass * k.c0h;
    const float ux = (mc * a.nvx + mc * b.bnx) * inv_den;
    const float uy = (mc * a.nvy + mc * b.bny) * inv_den;
    const float uz = (mc * a.nvz + mc * b.bnz) * inv_den;
    const float rho = k.mass * k.c0H * a.sH;
    const float drho = rho - k.rho0;
    const float D_I = k.c1h * (a.dr - (ux * a.gx + uy * a.gy + uz * a.gz) + b.db -
                               (ux * b.hx + uy * b.hy + uz * b.hz)) -
                      k.kappa * (drho > 0.f ? drho : 0.f) / k.rho0;
    vrho[out] = rho;
    vvol[out] = a.sV * k.c0h;
    vb[out] = D_I;
    const int64_t lin = ((int64_t)cz * k.ndims[1] + cy) * k.ndims[0] + cx;
    atomicOr(&bits[lin >> 5], 1u << (lin & 31));
}

// The real walk straight from global memory (2D, and 3D bricks whose
// neighbourhood does not fit the LDS stage).
template <int D>
__device__ __forceinline__ void real_walk_global(const KConst& k,
                                                 const int32_t* rcstart,
                                                 const double4* rpos,
                                                 const float4* rvel,
                                                 const float* rrho, double qx,
                                                 double qy, double qz,
                                                 RealAcc& a) {
    for_each_candidate<D>(k, rcstart, rpos, qx, qy, qz, [&](int j, const double4& pj) {
        const double d2 = D == 3 ? sq_dist3(pj.x, pj.y, pj.z, qx, qy, qz)
                                 : sq_dist2(pj.x, pj.y, qx, qy);
        if (d2 < k.h2)
            a.add<D>(k, d2, (float)(qx - pj.x), (float)(qy - pj.y),
                     D == 3 ? (float)(qz - pj.z) : 0.f, rvel[j], 1.f / rrho[j]);
    });
}

// Fused per-node pass.  One workgroup = one brick; thread = node.
template <int D, int NB>
__global__ void __launch_bounds__(NB) k_virtual_sums(
    KConst k, const int32_t* brick_of_slot, const int32_t* nslots_p,
    const double4* rpos, const float4* rvel, const float* rrho,
    const int32_t* rcstart, const double4* bpos, const float4* bvel,
    const float4* bnrm, const int32_t* bcstart, float* vrho, float* vvol,
    float* vb, uint32_t* bits) {
    const int nslots = *nslots_p;   // the grid may exceed the slot count
    if ((int)blockIdx.x >= nslots) return;
    const int slot = xcd_remap(blockIdx.x, nslots);
    const int brick = brick_of_slot[slot];
    const int t = threadIdx.x;
    const int lx = t & 7, ly = (t >> 3) & 7, lz = D == 3 ? t >> 6 : 0;
    const int bx = brick % k.bdims[0];
    const int by = (brick / k.bdims[0]) % k.bdims[1];
    const int bz = D == 3 ? brick / k.bdims[0] / k.bdims[1] : 0;
    const int cx = bx * kBrick + lx, cy = by * kBrick + ly,
              cz = bz * kBrick + lz;
    const int64_t out = (int64_t)slot * NB + t;
    if (cx >= k.ndims[0] || cy >= k.ndims[1] || (D == 3 && cz >= k.ndims[2])) {
        vrho[out] = 0.f;
        vvol[out] = 0.f;
        vb[out] = 0.f;
        return;
    }
    const double qx = node_coord(k.origin[0], cx, k.spacing);
    const double qy = node_coord(k.origin[1], cy, k.spacing);
    const double qz = D == 3 ? node_coord(k.origin[2], cz, k.spacing) : 0.0;
    RealAcc a;
    real_walk_global<D>(k, rcstart, rpos, rvel, rrho, qx, qy, qz, a);
    finish_node<D>(k, a, qx, qy, qz, cx, cy, cz, out, rpos, rcstart,
                   BoundaryWalkGlobal<D>{bcstart, bpos}, bvel, bnrm, vrho, vvol,
                   vb, bits);
}

// 3D with the brick's real neighbourhood staged in LDS (one workgroup of
// 512 threads per brick, thread = node).  The cells within one cell of the
// brick's nodes form <= kStageRows x-rows of <= kStageCx cells; their
// particles (contiguous per row in the cell-sorted order) are copied once,
// each as its FP32 offset from its own cell corner with 1/rho_j, and every
// node walks its 3 x 3 x 3 cells from LDS in the order of the global walk.
// q - x_j = (q - C_q) - (x_j - C_j) + (C_q - C_j) with C the FP64 cell
// corners, so the FP32 difference is accurate to a few ulp of h.
// Membership is pre-tested in FP32 with a band around h^2 in which the
// exact FP64 test on the original positions decides (also at r = 0 and for
// particles clamped into a border cell), so the neighbour sets stay exact.
// The boundary cell table of the same rows is staged too, so nodes away
// from the walls skip the boundary walk without touching global memory.
constexpr int kStageCap = 4096;
constexpr int kStageRows = 64;
constexpr int kStageCx = 8;

// Boundary walk of one node over the staged rows' boundary cell table, with
// the FP32 pre-test on the boundary's cell offsets (bofs) and the FP64 test
// in the band around h (the boundary set stays exact).
struct BoundaryWalkStaged {
    const int (*bcs)[kStageCx + 1];   // global starts of the staged rows
    const int* meta;                  // clo x,y,z; ncx, ncy
    const double4* bpos;
    const float4* bofs;
    __device__ __forceinline__ void operator()(const KConst& k, double qx, double qy,
                                               double qz, const float4* bvel,
                                               const float4* bnrm, BAcc& b) const {
        const int cx = cell_coord(qx, k.cell_lo[0], k.cell_inv, k.cdims[0]);
        const int cy = cell_coord(qy, k.cell_lo[1], k.cell_inv, k.cdims[1]);
        const int cz = cell_coord(qz, k.cell_lo[2], k.cell_inv, k.cdims[2]);
        const int xlo = max(cx - 1, 0), xhi = min(cx + 1, k.cdims[0] - 1);
        const int i0 = xlo - meta[0], i1 = xhi - meta[0] + 1;
        float tx0 = (float)(qx - cell_corner(k, 0, xlo));
        float tx1 = (float)(qx - cell_corner(k, 0, xlo + 1));
        float tx2 = (float)(qx - cell_corner(k, 0, xlo + 2));
        asm volatile("" : "+v"(tx0), "+v"(tx1), "+v"(tx2));
        for (int z = max(cz - 1, 0); z <= min(cz + 1, k.cdims[2] - 1); z++) {
            const float tz = (float)(qz - cell_corner(k, 2, z));
            for (int y = max(cy - 1, 0); y <= min(cy + 1, k.cdims[1] - 1); y++) {
                const int r = (z - meta[2]) * meta[4] + (y - meta[1]);
                const int jb = bcs[r][i0], je = bcs[r][i1];
                if (jb >= je) continue;
                const float ty = (float)(qy - cell_corner(k, 1, y));
                const int b1 = bcs[r][min(i0 + 1, i1)], b2 = bcs[r][min(i0 + 2, i1)];
                for (int j0 = jb; j0 < je; j0 += 4) {
                  float4 bov[4];
#pragma unroll
                  for (int u = 0; u < 4; u++)
                      bov[u] = j0 + u < je ? bofs[j0 + u] : make_float4(1e30f, 0.f, 0.f, 1.f);
#pragma unroll
                  for (int u = 0; u < 4; u++) {
                    const int j = j0 + u;
                    const float4 bo = bov[u];
                    const float tx = j < b1 ? tx0 : (j < b2 ? tx1 : tx2);
                    const float dx = tx - bo.x, dy = ty - bo.y, dz = tz - bo.z;
                    const float d2f = dx * dx + dy * dy + dz * dz;
                    const Pretest t = pretest(k, d2f, bo.w);
                    if (t.fast) b.add_fast(k, d2f, dx, dy, dz, bvel[j], bo.w);
                    if (t.exact) {
                        const double4 pj = bpos[j];
                        const double d2 = sq_dist3(pj.x, pj.y, pj.z, qx, qy, qz);
                        if (d2 < k.h2) b.add_exact<3>(k, qx, qy, qz, pj, d2, bvel[j], bnrm[j].w);
                    }
                  }
                }
            }
        }
    }
};

__global__ void __launch_bounds__(512) k_virtual_sums3s(
    KConst k, const int32_t* brick_of_slot, const int32_t* nslots_p,
    const double4* rpos, const float4* rvel, const float* rrho, const float* rrinv,
    const float4* rofs, const int32_t* rcstart, const double4* bpos,
    const float4* bofs, const float4* bvel, const float4* bnrm, const int32_t* bcstart,
    float* vrho, float* vvol, float* vb, uint32_t* bits) {
    // x_j - C_clo (NaN: exact test only), y_j - C_j, z_j - C_j as separate
    // arrays (the membership test and the sums read 12 bytes per candidate
    // as three 4-byte loads); x from the row's first cell, y and z from the
    // particle's own cell
    __shared__ float s_px[kStageCap], s_py[kStageCap], s_pz[kStageCap];
    __shared__ int s_cs[kStageRows][kStageCx + 1];  // real cell starts (staged)
    __shared__ int s_bcs[kStageRows][kStageCx + 1]; // boundary cell starts
    __shared__ int s_rg[kStageRows];    // global - staged index
    __shared__ int s_off[kStageRows + 1];
    __shared__ int s_meta[8];   // clo x,y,z; ncx, ncy, ncz; total; ok
    const int nslots = *nslots_p;   // the grid may exceed the slot count
    if ((int)blockIdx.x >= nslots) return;
    const int slot = xcd_remap(blockIdx.x, nslots);
    const int brick = brick_of_slot[slot];
    const int t = threadIdx.x;
    // a wave: the 8 x 8 nodes of one z-plane of the brick
    const int lx = t & 7, ly = (t >> 3) & 7, lz = t >> 6;
    const int bx = brick % k.bdims[0];
    const int by = (brick / k.bdims[0]) % k.bdims[1];
    const int bz = brick / k.bdims[0] / k.bdims[1];
    const int n0[3] = {bx * kBrick, by * kBrick, bz * kBrick};
    if (t == 0) {
        int ok = 1, nrows = 1;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const int last = min(n0[a] + kBrick - 1, k.ndims[a] - 1);
            const int clo = max(cell_coord(node_coord(k.origin[a], n0[a], k.spacing),
                                           k.cell_lo[a], k.cell_inv, k.cdims[a]) - 1, 0);
            const int chi = min(cell_coord(node_coord(k.origin[a], last, k.spacing),
                                           k.cell_lo[a], k.cell_inv, k.cdims[a]) + 1,
                                k.cdims[a] - 1);
            s_meta[a] = clo;
            s_meta[3 + a] = chi - clo + 1;
            if (a > 0) nrows *= chi - clo + 1;
        }
        if (s_meta[3] > kStageCx || nrows > kStageRows) ok = 0;
        s_meta[7] = ok;
    }
    __syncthreads();
    const int ncx = s_meta[3], ncy = s_meta[4], ncz = s_meta[5];
    const int nrows = ncy * ncz;
    if (s_meta[7] && t < nrows) {
        const int y = s_meta[1] + t % ncy, z = s_meta[2] + t / ncy;
        const int64_t grow = ((int64_t)z * k.cdims[1] + y) * k.cdims[0] + s_meta[0];
        for (int i = 0; i <= ncx; i++) {
            s_cs[t][i] = rcstart[grow + i];
            s_bcs[t][i] = bcstart[grow + i];
        }
    }
    __syncthreads();
    if (t == 0 && s_meta[7]) {
        int acc = 0;
        for (int r = 0; r < nrows; r++) {
            s_off[r] = acc;
            acc += s_cs[r][ncx] - s_cs[r][0];
        }
        s_off[nrows] = acc;
        s_meta[6] = acc;
        if (acc > kStageCap) s_meta[7] = 0;
    }
    __syncthreads();
    const bool staged = s_meta[7] != 0;
    if (staged) {
        if (t < nrows) {
            const int g0 = s_cs[t][0];
            s_rg[t] = g0 - s_off[t];
            for (int i = 0; i <= ncx; i++) s_cs[t][i] += s_off[t] - g0;
        }
        __syncthreads();
        const int total = s_meta[6];
        for (int e0 = 0; e0 < total; e0 += 4 * 512) {
            float4 pj[4];
            int ci[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int e = e0 + u * 512 + t;
                ci[u] = 0;
                if (e < total) {
                    // row of staged element e: last r with s_off[r] <= e
                    int lo = 0, hi = nrows - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (s_off[mid] <= e) lo = mid;
                        else hi = mid - 1;
                    }
                    pj[u] = rofs[e + s_rg[lo]];
                    // its cell within the row (x = clo + ci)
                    for (int i = 1; i < ncx; i++) ci[u] += s_cs[lo][i] <= e;
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int e = e0 + u * 512 + t;
                if (e < total) {
                    // x relative to the corner of the row's first cell
                    // (C_(clo+ci) - C_clo = ci cells), rounded once from
                    // FP64; a particle marked for the exact test only gets
                    // x = NaN, which no FP32 test passes
                    const float xr = pj[u].w < 0.f
                                         ? __builtin_nanf("")
                                         : (float)((double)pj[u].x + (double)ci[u] * k.cell_size);
                    s_px[e] = xr;
                    s_py[e] = pj[u].y;
                    s_pz[e] = pj[u].z;
                }
            }
        }
        __syncthreads();
    }
    const int cx = n0[0] + lx, cy = n0[1] + ly, cz = n0[2] + lz;
    const int64_t out = (int64_t)slot * 512 + t;
    if (cx >= k.ndims[0] || cy >= k.ndims[1] || cz >= k.ndims[2]) {
        vrho[out] = 0.f;
        vvol[out] = 0.f;
        vb[out] = 0.f;
        return;
    }
    const double qx = node_coord(k.origin[0], cx, k.spacing);
    const double qy = node_coord(k.origin[1], cy, k.spacing);
    const double qz = node_coord(k.origin[2], cz, k.spacing);
    RealAcc a;
    if (!staged) {
        real_walk_global<3>(k, rcstart, rpos, rvel, rrho, qx, qy, qz, a);
        finish_node<3>(k, a, qx, qy, qz, cx, cy, cz, out, rpos, rcstart,
                       BoundaryWalkGlobal<3>{bcstart, bpos}, bvel, bnrm, vrho,
                       vvol, vb, bits);
        return;
    }
    // q - C_c per axis for the candidate cell c, rounded once from FP64
    // (the staged x_j - C_j complete the difference q - x_j)
    const int cc0 = cell_coord(qx, k.cell_lo[0], k.cell_inv, k.cdims[0]);
    const int cc1 = cell_coord(qy, k.cell_lo[1], k.cell_inv, k.cdims[1]);
    const int cc2 = cell_coord(qz, k.cell_lo[2], k.cell_inv, k.cdims[2]);
    const int xlo = max(cc0 - 1, 0), xhi = min(cc0 + 1, k.cdims[0] - 1);
    const int zlo = max(cc2 - 1, 0), zhi = min(cc2 + 1, k.cdims[2] - 1);
    const int ylo = max(cc1 - 1, 0), yhi = min(cc1 + 1, k.cdims[1] - 1);
    // The (up to) three cells xlo .. xhi of a row are one contiguous staged
    // range.  The membership test and a member's sums take x relative to the
    // same corner C_clo for all of them (one offset per node, rounded once
    // from FP64: within 2e-7 h for offsets of up to three cells, inside the
    // FP64 band for the test and far below the tolerance of the sums).
    const float tx = (float)(qx - cell_corner(k, 0, s_meta[0]));
    float tx0 = (float)(qx - cell_corner(k, 0, xlo));
    float tx1 = (float)(qx - cell_corner(k, 0, xlo + 1));
    float tx2 = (float)(qx - cell_corner(k, 0, xlo + 2));
    asm volatile("" : "+v"(tx0), "+v"(tx1), "+v"(tx2));
    const int i0 = xlo - s_meta[0], i1 = xhi - s_meta[0];
    // h2_min < d2f < h2f_lo as one unsigned test on the bits (NaN fails).
    // Below h2_min (the band's width, far above the error of d2f near 0)
    // the exact test decides, so a fast member's accurate d2f is > 0.
    const uint32_t lo_min = __float_as_uint(k.h2f_hi - k.h2f_lo);
    const uint32_t lo_span = __float_as_uint(k.h2f_lo) - lo_min - 1u;
    // Cells whose nearest point lies beyond h (1 - 2e-6) are skipped: a
    // whole row by its y and z distance, the end cells of a row by their x
    // distance.  Only interior cells are culled (a border cell may hold
    // particles clamped into it from outside its extent).  The margin is
    // twenty times the FP32 error of these distances; a member it drops lies
    // within 2.1e-6 h of the cutoff, where W and W' are below 1e-17 and
    // 5e-12 of their peaks, far below FP32 resolution of the sums.  On a
    // lattice aligned with the cell grid (a node every half cell) it skips
    // the cells behind a node on a cell face; the nodes of a wave share
    // their z, so the rows behind them in z are skipped by the whole wave.
    const float csf = (float)k.cell_size;
    const float h2c = (float)k.h2 * (1.f - 4e-6f);
    auto gap = [csf](float t) { return fmaxf(fmaxf(-t, t - csf), 0.f); };   // q - C -> distance
    const float txh = xhi - xlo == 2 ? tx2 : (xhi - xlo == 1 ? tx1 : tx0);
    const bool cut_l = xlo < xhi && xlo > 0, cut_r = xhi > xlo && xhi < k.cdims[0] - 1;
    const float gl = gap(tx0), gr = gap(txh);
    const float dl2 = cut_l ? gl * gl : 1e30f, dr2 = cut_r ? gr * gr : 1e30f;
    for (int z = zlo; z <= zhi; z++) {
        const float tz = (float)(qz - cell_corner(k, 2, z));
        const float gz = z > 0 && z < k.cdims[2] - 1 ? gap(tz) : 0.f;
        for (int y = ylo; y <= yhi; y++) {
            const float ty = (float)(qy - cell_corner(k, 1, y));
            const float gy = y > 0 && y < k.cdims[1] - 1 ? gap(ty) : 0.f;
            const float rem = h2c - gy * gy - gz * gz;
            if (!(rem > 0.f)) continue;
            const int r = (z - s_meta[2]) * ncy + (y - s_meta[1]);
            const int rg = s_rg[r];
            const int jb = s_cs[r][i0 + (dl2 >= rem)], je = s_cs[r][i1 + 1 - (dr2 >= rem)];
            // Two passes per chunk of 32 candidates: the cheap FP32 test
            // sets one bit per member, then each lane visits only its own
            // members (a lane has ~15% of its candidates within h, so the
            // wave no longer runs the accumulation for every candidate).
            // Candidate c0 + i sits at bit i, so the lowest set bit is the
            // next in order.
            for (int c0 = jb; c0 < je; c0 += 32) {
                const int c1 = min(c0 + 32, je);
                uint32_t mfast = 0u, mexact = 0u;
                for (int j = c0; j < c1; j++) {
                    const float dx = tx - s_px[j], dy = ty - s_py[j], dz = tz - s_pz[j];
                    const float d2f = dx * dx + dy * dy + dz * dz;
                    const bool fast = __float_as_uint(d2f) - lo_min - 1u < lo_span;
                    const bool exact = !(d2f >= k.h2f_hi) & !fast;   // NaN: exact
                    mfast |= (uint32_t)fast << (j - c0);
                    mexact |= (uint32_t)exact << (j - c0);
                }
                while (mfast) {
                    const int j = c0 + __builtin_ctz(mfast);
                    mfast &= mfast - 1u;
                    const float dx = tx - s_px[j], dy = ty - s_py[j], dz = tz - s_pz[j];
                    const float d2f = dx * dx + dy * dy + dz * dz;
                    a.add_fast<3>(k, d2f, dx, dy, dz, rvel[j + rg], rrinv[j + rg]);
                }
                while (mexact) {
                    // the reference's FP64 test and the accurate spline
                    const int j = c0 + __builtin_ctz(mexact);
                    mexact &= mexact - 1u;
                    const int g = j + rg;
                    const double4 pj = rpos[g];
                    const double d2 = sq_dist3(pj.x, pj.y, pj.z, qx, qy, qz);
                    if (d2 < k.h2)
                        a.add<3>(k, d2, (float)(qx - pj.x), (float)(qy - pj.y),
                                 (float)(qz - pj.z), rvel[g], rrinv[g]);
                }
            }
        }
    }
    finish_node<3>(k, a, qx, qy, qz, cx, cy, cz, out, rpos, rcstart,
                   BoundaryWalkStaged{s_bcs, s_meta, bpos, bofs}, bvel, bnrm, vrho,
                   vvol, vb, bits);
}

// Slots of the 3^D neighbour bricks of every slot (zero brick if absent).
template <int D>
__global__ void __launch_bounds__(256) k_nbr_slots(KConst k,
                                                   const int32_t* brick_of_slot,
                                                   const int32_t* slot_of_brick,
                                                   const int32_t* nslots_p, int32_t zero,
                                                   int32_t* nbr) {
    constexpr int NN = D == 3 ? 27 : 9;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)*nslots_p * NN) return;
    const int64_t slot = i / NN;
    const int t = (int)(i - slot * NN);
    const int brick = brick_of_slot[slot];
    const int bx = brick % k.bdims[0];
    const int by = (brick / k.bdims[0]) % k.bdims[1];
    const int bz = D == 3 ? brick / k.bdims[0] / k.bdims[1] : 0;
    const int nx = bx + (t % 3) - 1, ny = by + (t / 3) % 3 - 1,
              nz = D == 3 ? bz + t / 9 - 1 : 0;
    int s = zero;
    if (nx >= 0 && ny >= 0 && nz >= 0 && nx < k.bdims[0] && ny < k.bdims[1] &&
        nz < k.bdims[2]) {
        const int sl = slot_of_brick[((int64_t)nz * k.bdims[1] + ny) * k.bdims[0] + nx];
        if (sl >= 0) s = sl;
    }
    nbr[slot * 27 + t] = s;
}

// Column order of the 3D slots, for the CG iteration's walk up z-runs of
// bricks.  A run starts at a slot whose dz = -1 neighbour (index 4 of the
// 27) is the zero brick and continues through the dz = +1 neighbours (index
// 22); run[s] is the run's length at its first slot and 0 elsewhere.
__global__ void __launch_bounds__(256) k_col_runs(const int32_t* nbr, const int32_t* nslots_p,
                                                  int64_t grid, int32_t zero, int32_t* run) {
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= grid) return;
    // the grid may exceed the slot count; a count beyond the grid (the
    // neighbour table is then incomplete) leaves the order to the rerun on
    // the exact grid
    const int32_t nsl = *nslots_p;
    if (s >= nsl || nsl > grid) {
        run[s] = 0;
        return;
    }
    int32_t len = 0;
    if (nbr[s * 27 + 4] == zero) {
        int32_t b = (int32_t)s;
        len = 1;
        while ((b = nbr[(int64_t)b * 27 + 22]) != zero) len++;
    }
    run[s] = len;
}

// order[off[first slot of the run] + height in the run] = slot
__global__ void __launch_bounds__(256) k_col_place(const int32_t* nbr, const int32_t* nslots_p,
                                                   int64_t grid, int32_t zero,
                                                   const int32_t* off, int32_t* order) {
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int32_t nsl = *nslots_p;
    if (s >= nsl || nsl > grid) return;
    int32_t b = (int32_t)s, d = 0, n;
    while ((n = nbr[(int64_t)b * 27 + 4]) != zero) {
        b = n;
        d++;
    }
    order[off[b] + d] = (int32_t)s;
}

__global__ void __launch_bounds__(256) k_popc(const uint32_t* bits,
                                              int64_t nwords, int32_t* cnt) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < nwords) cnt[i] = __popc(bits[i]);
}

template <int D, int NB>
__global__ void __launch_bounds__(NB) k_node_rank(KConst k,
                                                  const int32_t* brick_of_slot,
                                                  const int32_t* nslots_p,
                                                  const float* vvol,
                                                  const uint32_t* bits,
                                                  const int32_t* prefix,
                                                  int32_t* rank) {
    const int slot = blockIdx.x;
    if (slot >= *nslots_p) return;
    const int64_t out = (int64_t)slot * NB + threadIdx.x;
    if (!(vvol[out] > 0.f)) {
        rank[out] = -1;
        return;
    }
    const int brick = brick_of_slot[slot];
    const int t = threadIdx.x;
    const int bx = brick % k.bdims[0];
    const int by = (brick / k.bdims[0]) % k.bdims[1];
    const int bz = D == 3 ? brick / k.bdims[0] / k.bdims[1] : 0;
    const int cx = bx * kBrick + (t & 7), cy = by * kBrick + ((t >> 3) & 7),
              cz = D == 3 ? bz * kBrick + (t >> 6) : 0;
    const int64_t lin = ((int64_t)cz * k.ndims[1] + cy) * k.ndims[0] + cx;
    const uint32_t word = bits[lin >> 5];
    const uint32_t below = word & ((1u << (lin & 31)) - 1u);
    rank[out] = prefix[lin >> 5] + __popc(below);
}

// CG initial guess for every node of the active bricks (0 off the set).
template <int D, int NB>
__global__ void __launch_bounds__(NB) k_warm_start(
    int mode, KConst k, const int32_t* brick_of_slot, const float* vvol,
    const int32_t* rank, const float* p_rank_prev, int64_t nprev,
    const int32_t* prev_slot_of_brick, const float* x_prev, float* x0) {
    const int slot = blockIdx.x;
    const int64_t out = (int64_t)slot * NB + threadIdx.x;
    float v = 0.f;
    if (vvol[out] > 0.f) {
        if (mode == VPS_WARM_INDEX_CARRY) {
            const int r = rank[out];
            if (r < nprev) v = p_rank_prev[r];
        } else if (mode == VPS_WARM_NODE_KEYED || mode == VPS_WARM_LEARNED) {
            if (prev_slot_of_brick) {
                const int ps = prev_slot_of_brick[brick_of_slot[slot]];
                if (ps >= 0) v = x_prev[(int64_t)ps * NB + threadIdx.x];
            }
        }
    }
    x0[out] = v;
}

// After the solve: pressures by reference rank (index-carry warm start of
// the next step) and q = V p for the pressure gradient.
__global__ void __launch_bounds__(256) k_after_solve(int64_t nnode,
                                                     const float* vvol,
                                                     const float* x,
                                                     const int32_t* rank,
                                                     float* p_rank, float* q,
                                                     SolveGuard guard) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nnode || !solve_accepted(guard)) return;
    const float V = vvol[i];
    const float p = x[i];
    q[i] = V * p;
    if (V > 0.f) p_rank[rank[i]] = p;
}

}  // namespace

// Slot-indexed passes of one generation on a grid of `grid` blocks (slots);
// blocks at or beyond the device-side slot count return at once.
static int generate_slots(vps_context* c, int w, int64_t grid) {
    const int D = c->dim;
    const int rs = c->cur;
    const int32_t* nsl = (const int32_t*)c->counters.p;   // slot count
    int64_t* dcnt = c->counters.p;
    if (D == 3)
        hipLaunchKernelGGL(k_virtual_sums3s, dim3((unsigned)grid),
                           dim3(512), 0, c->stream, c->k, c->slot_brick[w].p,
                           nsl, c->rpos[rs].p, c->rvel[rs].p, c->rrho.p, c->rrinv.p, c->rofs.p,
                           c->cell_start.p, c->bpos.p, c->bofs.p, c->bvel.p, c->bnrm.p,
                           c->bcell_start.p, c->vrho.p, c->vvol.p, c->vb.p,
                           c->bits.p);
    else
        hipLaunchKernelGGL((k_virtual_sums<2, 64>), dim3((unsigned)grid),
                           dim3(64), 0, c->stream, c->k, c->slot_brick[w].p,
                           nsl, c->rpos[rs].p, c->rvel[rs].p, c->rrho.p,
                           c->cell_start.p, c->bpos.p, c->bvel.p, c->bnrm.p,
                           c->bcell_start.p, c->vrho.p, c->vvol.p, c->vb.p,
                           c->bits.p);
    if (D == 3) {
        hipLaunchKernelGGL(k_nbr_slots<3>, dim3(blocks(grid * 27, 256)), dim3(256),
                           0, c->stream, c->k, c->slot_brick[w].p, c->brick_slot[w].p,
                           nsl, (int32_t)c->nbricks, c->nbr_slot.p);
        // the column order, on the grid (before the read-back below, so
        // that it runs while the host waits)
        hipLaunchKernelGGL(k_col_runs, dim3(blocks(grid, 256)), dim3(256), 0, c->stream,
                           c->nbr_slot.p, nsl, grid, (int32_t)c->nbricks, c->col_run.p);
        VPS_TRY(scan_exclusive_i32(c, c->col_run.p, c->col_run.p, grid, nullptr));
        hipLaunchKernelGGL(k_col_place, dim3(blocks(grid, 256)), dim3(256), 0, c->stream,
                           c->nbr_slot.p, nsl, grid, (int32_t)c->nbricks, c->col_run.p,
                           c->col_order.p);
    } else
        hipLaunchKernelGGL(k_nbr_slots<2>, dim3(blocks(grid * 9, 256)), dim3(256),
                           0, c->stream, c->k, c->slot_brick[w].p, c->brick_slot[w].p,
                           nsl, (int32_t)c->nbricks, c->nbr_slot.p);
    hipLaunchKernelGGL(k_popc, dim3(blocks(c->nwords, 256)), dim3(256), 0,
                       c->stream, c->bits.p, c->nwords,
                       (int32_t*)c->word_prefix.p);
    VPS_TRY(scan_exclusive_i32(c, (const int32_t*)c->word_prefix.p,
                               (int32_t*)c->word_prefix.p, c->nwords,
                               (int32_t*)(dcnt + 1)));
    if (D == 3)
        hipLaunchKernelGGL((k_node_rank<3, 512>), dim3((unsigned)grid),
                           dim3(512), 0, c->stream, c->k, c->slot_brick[w].p,
                           nsl, c->vvol.p, c->bits.p,
                           (const int32_t*)c->word_prefix.p, c->vrank.p);
    else
        hipLaunchKernelGGL((k_node_rank<2, 64>), dim3((unsigned)grid),
                           dim3(64), 0, c->stream, c->k, c->slot_brick[w].p,
                           nsl, c->vvol.p, c->bits.p,
                           (const int32_t*)c->word_prefix.p, c->vrank.p);
    VPS_HIP(c, hipGetLastError());
    // slot, virtual and gradient-span counts in one read-back
    VPS_HIP(c, hipMemcpyAsync(c->host_counters, dcnt, 24, hipMemcpyDeviceToHost,
                              c->stream));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    return VPS_OK;
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

int virtual_generate(vps_context* c) {
    // The new slot map goes into the buffer that does NOT hold the last solved
    // pressures, so both warm starts can still read them.
    const int w = c->last_solve_buf >= 0 ? 1 - c->last_solve_buf : 0;
    const int D = c->dim;
    const int s = c->cur;
    // brick flags are zero on entry (cleared by k_assign_slots)
    const unsigned mb = blocks(c->n, 256) > 0 ? blocks(c->n, 256) : 1;
    if (D == 3)
        hipLaunchKernelGGL(k_mark_bricks<3>, dim3(mb), dim3(256),
                           0, c->stream, c->rpos[s].p, c->n, c->k,
                           c->brick_flag.p, c->bits.p, c->nwords,
                           (int32_t*)(c->counters.p + 2));
    else
        hipLaunchKernelGGL(k_mark_bricks<2>, dim3(mb), dim3(256),
                           0, c->stream, c->rpos[s].p, c->n, c->k,
                           c->brick_flag.p, c->bits.p, c->nwords,
                           (int32_t*)(c->counters.p + 2));
    VPS_TRY(gradient_spans(c));
    int64_t* dcnt = c->counters.p;
    VPS_TRY(scan_exclusive_i32(c, (const int32_t*)c->brick_flag.p,
                               c->brick_slot[w].p, c->nbricks,
                               (int32_t*)dcnt));
    hipLaunchKernelGGL(k_assign_slots, dim3(blocks(c->nbricks, 256)), dim3(256),
                       0, c->stream, c->brick_flag.p, c->nbricks,
                       c->brick_slot[w].p, c->slot_brick[w].p);
    // The slot count stays on the device: the slot passes run on a grid
    // sized from the last generation's count (all bricks the first time),
    // and again on the exact grid in the rare step where it grew past that.
    int64_t grid = c->nbricks;
    if (c->nslots > 0) {
        grid = c->nslots + c->nslots / 8 + 64;
        if (grid > c->nbricks) grid = c->nbricks;
    }
    VPS_TRY(generate_slots(c, w, grid));
    int64_t nslots = *(int32_t*)c->host_counters;
    if (nslots > grid) {
        hipLaunchKernelGGL(k_fill_u32, dim3(blocks(c->nwords, 256)), dim3(256), 0,
                           c->stream, c->bits.p, c->nwords, 0u);
        VPS_TRY(generate_slots(c, w, nslots));
        nslots = *(int32_t*)c->host_counters;
    }
    c->nslots = nslots;
    c->which = w;
    c->nvirt = nslots ? *(int32_t*)(c->host_counters + 1) : 0;
    c->grad_nchunks = *(int32_t*)(c->host_counters + 2);
    return VPS_OK;
}

int warm_start(vps_context* c, int mode, float* x0) {
    const int w = c->which;
    const int lb = c->last_solve_buf;
    const int32_t* prev_map = lb >= 0 ? c->brick_slot[lb].p : nullptr;
    const float* p_prev = lb >= 0 ? c->p_by_rank[lb].p : nullptr;
    const int64_t nprev = lb >= 0 ? c->last_solve_nvirt : 0;
    const float* x_prev = lb >= 0 ? c->vx[lb].p : nullptr;
    if (c->dim == 3)
        hipLaunchKernelGGL((k_warm_start<3, 512>), dim3((unsigned)c->nslots),
                           dim3(512), 0, c->stream, mode, c->k,
                           c->slot_brick[w].p, c->vvol.p, c->vrank.p, p_prev,
                           nprev, prev_map, x_prev, x0);
    else
        hipLaunchKernelGGL((k_warm_start<2, 64>), dim3((unsigned)c->nslots),
                           dim3(64), 0, c->stream, mode, c->k,
                           c->slot_brick[w].p, c->vvol.p, c->vrank.p, p_prev,
                           nprev, prev_map, x_prev, x0);
    VPS_HIP(c, hipGetLastError());
    return VPS_OK;
}

int after_solve(vps_context* c, SolveGuard g) {
    const int w = c->which;
    const int64_t nn = c->nslots * c->brick_nodes;
    if (nn == 0) return VPS_OK;
    hipLaunchKernelGGL(k_after_solve, dim3(blocks(nn, 256)), dim3(256), 0,
                       c->stream, nn, c->vvol.p, c->vx[w].p, c->vrank.p,
                       c->p_by_rank[w].p, c->vq.p, g);
    VPS_HIP(c, hipGetLastError());
    if (g.mode == 0) after_solve_commit(c);
    return VPS_OK;
}

void after_solve_commit(vps_context* c) {
    if (c->nslots * c->brick_nodes == 0) return;
    c->last_solve_buf = c->which;
    c->last_solve_nvirt = c->nvirt;
}

}  // namespace vps
