// vps_cells.hip -- uniform-grid counting sort and real/boundary densities.
//
// Replaces HashGridSearcher::reset (ParticlesNearbySearcher.cpp:10-18) and
// SmoothedParticles::computeDensities / computeVolumes
// (SmoothedParticles.cpp:25-36).  The reference hashes cells into 2N
// buckets with OR-combined primes and walks 27 sorted bucket keys per query;
// here particles are counting-sorted by a dense x-fastest cell id (cell size
// h (1 + 1e-12)), so the 3 cells of an x-row are one contiguous range and a
// query reads 9 ranges (3 in 2D).  Membership is the reference's exact FP64
// test (x_j - q).squaredNorm() < h^2 (.cpp:32), so the neighbour sets are
// identical; only the visiting order differs.  Within a cell particles are
// ordered by caller index, so every sum is deterministic.
#include "vps_internal.h"

namespace vps {

namespace {

// Cell id and a slot inside the cell.  Particles arrive in last step's cell
// order, so a wave holds runs of equal cells: one atomic per run (its head
// lane adds the run length), lanes take consecutive slots of their run.
template <int D>
__global__ void __launch_bounds__(256) k_cell_count(const double4* pos,
                                                    int64_t n, KConst k,
                                                    int32_t* count,
                                                    int32_t* cell_of,
                                                    int32_t* slot_of) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool valid = i < n;
    int c = -1;
    if (valid) {
        const double4 p = pos[i];
        const int cx = cell_coord(p.x, k.cell_lo[0], k.cell_inv, k.cdims[0]);
        const int cy = cell_coord(p.y, k.cell_lo[1], k.cell_inv, k.cdims[1]);
        const int cz = D == 3 ? cell_coord(p.z, k.cell_lo[2], k.cell_inv, k.cdims[2]) : 0;
        c = (cz * k.cdims[1] + cy) * k.cdims[0] + cx;
        cell_of[i] = c;
    }
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(c, 1, 64);
    const bool head = valid && (lane == 0 || prev != c);
    const unsigned long long heads = __ballot(head);
    const unsigned long long vmask = __ballot(valid);
    const unsigned long long upto =
        lane == 63 ? ~0ull : ((2ull << lane) - 1ull);   // lanes 0..lane
    const int leader = 63 - __clzll(heads & upto);
    int base = 0;
    if (head) {
        const unsigned long long after = heads & ~upto;
        const int next = after ? __ffsll((long long)after) - 1 : 64;
        const int nvalid = __popcll(vmask);
        const int end = next < nvalid ? next : nvalid;
        base = atomicAdd(&count[c], end - lane);
    }
    base = __shfl(base, leader < 0 ? 0 : leader, 64);
    if (valid) slot_of[i] = base + (lane - leader);
}

__global__ void __launch_bounds__(256) k_cell_scatter(const int32_t* cell_of,
                                                      const int32_t* slot_of,
                                                      const int32_t* cstart,
                                                      int64_t n, int32_t* perm) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    perm[cstart[cell_of[i]] + slot_of[i]] = (int32_t)i;
}

// Orders each cell's entries by caller id (atomics gave an arbitrary order).
__global__ void __launch_bounds__(256) k_cell_order(const int32_t* cstart,
                                                    int64_t ncells,
                                                    int32_t* perm,
                                                    const int32_t* id) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= ncells) return;
    const int b = cstart[c], e = cstart[c + 1];
    for (int t = b + 1; t < e; t++) {
        const int v = perm[t];
        const int key = id[v];
        int u = t - 1;
        while (u >= b && id[perm[u]] > key) {
            perm[u + 1] = perm[u];
            u--;
        }
        perm[u + 1] = v;
    }
}

template <typename A, typename B>
__global__ void __launch_bounds__(256) k_gather3(const int32_t* perm, int64_t n,
                                                 const double4* p_in,
                                                 double4* p_out, const A* a_in,
                                                 A* a_out, const B* b_in,
                                                 B* b_out, const int32_t* id_in,
                                                 int32_t* id_out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int s = perm[i];
    p_out[i] = p_in[s];
    a_out[i] = a_in[s];
    if (b_in) b_out[i] = b_in[s];
    id_out[i] = id_in[s];
}

// rho_i = m * sum_j W_h(|x_i - x_j|) over j with FP64 |x_j - x_i|^2 < h^2
// (SmoothedParticles.cpp:25-31 via getNeighborWeight, .cpp:48-54).
template <int D>
__global__ void __launch_bounds__(256) k_density(const double4* pos,
                                                 const int32_t* cstart,
                                                 int64_t n, KConst k,
                                                 float* rho, float4* vel,
                                                 int set_volume) {
    const int nb = (int)((n + 255) / 256);
    const int64_t i = (int64_t)xcd_remap(blockIdx.x, nb) * 256 + threadIdx.x;
    if (i >= n) return;
    const double4 q = pos[i];
    float sum = 0.f;
    for_each_candidate<D>(k, cstart, pos, q.x, q.y, q.z, [&](int j, const double4& pj) {
        const double d2 = D == 3 ? sq_dist3(pj.x, pj.y, pj.z, q.x, q.y, q.z)
                                 : sq_dist2(pj.x, pj.y, q.x, q.y);
        if (d2 < k.h2) {
            const KW w = kernel_f32(d2, k.h2, k.hf, k.inv_h, k.c0h, k.c1h);
            sum += w.w;
        }
    });
    const float r = k.mass * sum;
    rho[i] = r;
    if (set_volume) vel[i].w = k.mass / r;
}

// 3D density from the FP32 cell-corner offsets (16 B per candidate instead
// of the FP64 position): x_i - x_j = (x_i - C_i) - (x_j - C_j) + (C_i - C_j)
// per axis, with C the FP64 cell corners.  W_h is continuous and zero at h,
// so the FP32 distance decides which terms enter (a pair within rounding of
// h adds nothing measurable); particles clamped into a border cell (ofs.w <
// 0) take the FP64 distance.
// BND: the boundary density (own searcher, DualParticleSphLiquid.cpp:96-97)
// into nrm.w, V_b = m / rho_b into vel.w; otherwise rho and 1 / rho, and
// V = m / rho into vel.w when set_volume.
// Off: the type of the candidates' byte offsets (uint32_t when the set's
// offsets fit, so the loads take a scalar base and a 32-bit lane offset).
// CNT (test support): only the number of terms that enter the sum (a > 0: r < h,
// where the spline is nonzero) into members[i].
template <bool BND, typename Off, bool CNT = false>
__global__ void __launch_bounds__(256) k_density3f(const double4* pos, const float4* ofs,
                                                   const int32_t* cstart, int64_t n,
                                                   KConst k, float* rho, float* rinv,
                                                   float4* nrm, float4* vel, int set_volume,
                                                   int32_t* members = nullptr) {
    const int nb = (int)((n + 255) / 256);
    const int64_t i = (int64_t)xcd_remap(blockIdx.x, nb) * 256 + threadIdx.x;
    if (i >= n) return;
    const double4 q = pos[i];
    const int cc0 = cell_coord(q.x, k.cell_lo[0], k.cell_inv, k.cdims[0]);
    const int cc1 = cell_coord(q.y, k.cell_lo[1], k.cell_inv, k.cdims[1]);
    const int cc2 = cell_coord(q.z, k.cell_lo[2], k.cell_inv, k.cdims[2]);
    const int x0 = cc0 > 0 ? cc0 - 1 : 0;
    const int x1 = cc0 + 1 < k.cdims[0] ? cc0 + 1 : k.cdims[0] - 1;
    const int zlo = cc2 > 0 ? cc2 - 1 : 0, zhi = cc2 + 1 < k.cdims[2] ? cc2 + 1 : k.cdims[2] - 1;
    const int ylo = cc1 > 0 ? cc1 - 1 : 0, yhi = cc1 + 1 < k.cdims[1] ? cc1 + 1 : k.cdims[1] - 1;
    // q - C_x0, rounded once from FP64 (the staged x_j - C_j complete
    // q - x_j).  A candidate in cell x0 + d (d = 0, 1, 2; the offset's w is
    // its cell x + 1) takes q - C_x0 - d cells, formed by one fma: exact
    // when q - C_x0 lies within [1, 2] cells (Sterbenz), else rounded once.
    const float tx0 = (float)(q.x - cell_corner(k, 0, x0));
    const float csf = (float)k.cell_size;
    const float x0w = (float)(x0 + 1);
    // a particle clamped into a border cell takes the FP64 distance throughout
    const float far = 2.f * (float)k.cell_size;
    const float qo0 = (float)(q.x - cell_corner(k, 0, cc0));
    const float qo1 = (float)(q.y - cell_corner(k, 1, cc1));
    const float qo2 = (float)(q.z - cell_corner(k, 2, cc2));
    const bool qfar = !(fabsf(qo0) <= far && fabsf(qo1) <= far && fabsf(qo2) <= far);
    // W / c0 = 2 max(1 - x, 0)^3 - max(1 - 2x, 0)^3: the two pieces of the
    // cubic spline in one expression (for x < 1/2 it expands to
    // 6 x^2 (x - 1) + 1), zero from x = 1 on, so no support test is needed.
    // The two cubes are summed apart (two multiplies and two fmas per
    // candidate) and combined, with c0 m, once per particle.  One loop per
    // row of three cells (the lanes of a wave sit in different cells:
    // per-cell loops would run the longest cell of every lane in turn).
    const float ih = k.inv_h, ih2 = 2.f * k.inv_h;
    float suma = 0.f, sumc = 0.f;
    int32_t nterms = 0;
    for (int z = zlo; z <= zhi; z++) {
        const float tz = (float)(q.z - cell_corner(k, 2, z));
        for (int y = ylo; y <= yhi; y++) {
            const float ty = (float)(q.y - cell_corner(k, 1, y));
            const int64_t row = ((int64_t)z * k.cdims[1] + y) * k.cdims[0];
            const int b = cstart[row + x0], e = cstart[row + x1 + 1];
            for (int j = b; j < e; j++) {
                const Off o = (Off)j * 16u;
                const float4 p = *(const float4*)((const char*)ofs + o);
                const float tx = fmaf(x0w - p.w, csf, tx0);
                const float dx = tx - p.x, dy = ty - p.y, dz = tz - p.z;
                float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                if (p.w < 0.f || qfar) {   // rare: FP64 distance
                    const double4 pj = *(const double4*)((const char*)pos + (uint64_t)o * 2u);
                    d2 = (float)sq_dist3(pj.x, pj.y, pj.z, q.x, q.y, q.z);
                }
                // v_sqrt_f32 (1 ulp): W is rounded to FP32 anyway
                const float r = __builtin_amdgcn_sqrtf(d2);
                // max(., 0) of values <= 1 as the clamp to [0, 1], which
                // folds into the fma
                const float a = __builtin_amdgcn_fmed3f(fmaf(-ih, r, 1.f), 0.f, 1.f);
                const float c = __builtin_amdgcn_fmed3f(fmaf(-ih2, r, 1.f), 0.f, 1.f);
                suma = fmaf(a * a, a, suma);
                sumc = fmaf(c * c, c, sumc);
                if constexpr (CNT) nterms += a > 0.f;
            }
        }
    }
    if constexpr (CNT) {
        members[i] = nterms;
        return;
    }
    const float r = k.mass * k.c0h * fmaf(2.f, suma, -sumc);
    if constexpr (BND) {
        nrm[i].w = r;
        vel[i].w = k.mass / r;
    } else {
        rho[i] = r;
        rinv[i] = 1.f / r;
        if (set_volume) vel[i].w = k.mass / r;
    }
}

// FP32 offset of every (sorted) particle from its cell corner; w = its cell
// x + 1, negated when the particle is clamped into a border cell (|offset|
// > 2 cells).  The walks read the sign; the density also the cell.
__global__ void __launch_bounds__(256) k_cell_offsets3(const double4* pos,
                                                       int64_t n, KConst k,
                                                       float4* ofs) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double4 p = pos[i];
    const double pc[3] = {p.x, p.y, p.z};
    const float far = 2.f * (float)k.cell_size;
    float of[3];
    bool is_far = false;
    int cx = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int c = cell_coord(pc[a], k.cell_lo[a], k.cell_inv, k.cdims[a]);
        if (a == 0) cx = c;
        of[a] = (float)(pc[a] - cell_corner(k, a, c));
        is_far |= !(fabsf(of[a]) <= far);
    }
    const float w = (float)(cx + 1);
    ofs[i] = make_float4(of[0], of[1], of[2], is_far ? -w : w);
}

// Boundary density from boundary-only neighbours (the boundary set has its
// own searcher, DualParticleSphLiquid.cpp:96-97); V_b = m / rho_b.
template <int D>
__global__ void __launch_bounds__(256) k_bdensity(const double4* pos,
                                                  const int32_t* cstart,
                                                  int64_t n, KConst k,
                                                  float4* nrm, float4* vel) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double4 q = pos[i];
    float sum = 0.f;
    for_each_candidate<D>(k, cstart, pos, q.x, q.y, q.z, [&](int j, const double4& pj) {
        const double d2 = D == 3 ? sq_dist3(pj.x, pj.y, pj.z, q.x, q.y, q.z)
                                 : sq_dist2(pj.x, pj.y, q.x, q.y);
        if (d2 < k.h2) {
            const KW w = kernel_f32(d2, k.h2, k.hf, k.inv_h, k.c0h, k.c1h);
            sum += w.w;
        }
    });
    const float r = k.mass * sum;
    nrm[i].w = r;
    vel[i].w = k.mass / r;
}

// FP32 offsets of the (sorted) boundary particles from their cell corners,
// w = +1/rho_b, or -1/rho_b when clamped into a border cell (as the real
// offsets); recomputed with every boundary density.
__global__ void __launch_bounds__(256) k_boundary_offsets3(const double4* pos,
                                                           const float4* nrm, int64_t n,
                                                           KConst k, float4* ofs) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double4 p = pos[i];
    const double pc[3] = {p.x, p.y, p.z};
    const float far = 2.f * (float)k.cell_size;
    float of[3];
    bool is_far = false;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int c = cell_coord(pc[a], k.cell_lo[a], k.cell_inv, k.cdims[a]);
        of[a] = (float)(pc[a] - cell_corner(k, a, c));
        is_far |= !(fabsf(of[a]) <= far);
    }
    const float ir = 1.f / nrm[i].w;
    ofs[i] = make_float4(of[0], of[1], of[2], is_far ? -ir : ir);
}

inline unsigned blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

// Counting sort of a particle set into the cell grid.  perm/cell_of/slot_of
// are scratch of length n; cstart gets ncells + 1 entries.
template <typename A, typename B>
static int sort_set(vps_context* c, int64_t n, const double4* pin,
                    double4* pout, const A* ain, A* aout, const B* bin,
                    B* bout, const int32_t* idin, int32_t* idout,
                    int32_t* cstart) {
    const int64_t C = c->ncells;
    // cell_count is zero on entry: cleared at creation and by every scan of it
    if (c->dim == 3)
        hipLaunchKernelGGL(k_cell_count<3>, dim3(blocks(n, 256)), dim3(256), 0,
                           c->stream, pin, n, c->k, c->cell_count.p,
                           c->cell_of.p, c->slot_of.p);
    else
        hipLaunchKernelGGL(k_cell_count<2>, dim3(blocks(n, 256)), dim3(256), 0,
                           c->stream, pin, n, c->k, c->cell_count.p,
                           c->cell_of.p, c->slot_of.p);
    VPS_TRY(scan_exclusive_i32(c, c->cell_count.p, cstart, C, cstart + C,
                               c->cell_count.p));
    hipLaunchKernelGGL(k_cell_scatter, dim3(blocks(n, 256)), dim3(256), 0,
                       c->stream, c->cell_of.p, c->slot_of.p, cstart, n,
                       c->perm.p);
    hipLaunchKernelGGL(k_cell_order, dim3(blocks(C, 256)), dim3(256), 0,
                       c->stream, cstart, C, c->perm.p, idin);
    hipLaunchKernelGGL((k_gather3<A, B>), dim3(blocks(n, 256)), dim3(256), 0,
                       c->stream, c->perm.p, n, pin, pout, ain, aout, bin,
                       bout, idin, idout);
    VPS_HIP(c, hipGetLastError());
    return VPS_OK;
}

int cells_sort_real(vps_context* c) {
    const int s = c->cur, d = 1 - c->cur;
    VPS_TRY((sort_set<float4, float4>(c, c->n, c->rpos[s].p, c->rpos[d].p,
                                      c->rvel[s].p, c->rvel[d].p, nullptr,
                                      nullptr, c->rid[s].p, c->rid[d].p,
                                      c->cell_start.p)));
    c->cur = d;
    return VPS_OK;
}

int cells_sort_boundary(vps_context* c) {
    // sort into the spare buffers, then swap them in (both stay allocated:
    // a moving boundary is re-sorted every step)
    const int64_t nb = c->nb;
    VPS_TRY(dalloc(c, c->bpos_spare, nb));
    VPS_TRY(dalloc(c, c->bvel_spare, nb));
    VPS_TRY(dalloc(c, c->bnrm_spare, nb));
    VPS_TRY(dalloc(c, c->bid_spare, nb));
    VPS_TRY(dalloc(c, c->cell_of, nb > c->n ? nb : c->n));
    VPS_TRY(dalloc(c, c->slot_of, nb > c->n ? nb : c->n));
    VPS_TRY(dalloc(c, c->perm, nb > c->n ? nb : c->n));
    VPS_TRY((sort_set<float4, float4>(c, nb, c->bpos.p, c->bpos_spare.p, c->bvel.p,
                                      c->bvel_spare.p, c->bnrm.p, c->bnrm_spare.p,
                                      c->bid.p, c->bid_spare.p, c->bcell_start.p)));
    std::swap(c->bpos, c->bpos_spare);
    std::swap(c->bvel, c->bvel_spare);
    std::swap(c->bnrm, c->bnrm_spare);
    std::swap(c->bid, c->bid_spare);
    return VPS_OK;
}

int real_density(vps_context* c, bool set_volume) {
    const int64_t n = c->n;
    const int s = c->cur;
    if (c->dim == 3) {
        hipLaunchKernelGGL(k_cell_offsets3, dim3(blocks(n, 256)), dim3(256), 0,
                           c->stream, c->rpos[s].p, n, c->k, c->rofs.p);
        if (n * 16 <= (int64_t)UINT32_MAX)
            hipLaunchKernelGGL((k_density3f<false, uint32_t>), dim3(blocks(n, 256)), dim3(256),
                               0, c->stream, c->rpos[s].p, c->rofs.p, c->cell_start.p, n,
                               c->k, c->rrho.p, c->rrinv.p, (float4*)nullptr, c->rvel[s].p,
                               set_volume ? 1 : 0);
        else
            hipLaunchKernelGGL((k_density3f<false, uint64_t>), dim3(blocks(n, 256)), dim3(256),
                               0, c->stream, c->rpos[s].p, c->rofs.p, c->cell_start.p, n,
                               c->k, c->rrho.p, c->rrinv.p, (float4*)nullptr, c->rvel[s].p,
                               set_volume ? 1 : 0);
    } else
        hipLaunchKernelGGL(k_density<2>, dim3(blocks(n, 256)), dim3(256), 0,
                           c->stream, c->rpos[s].p, c->cell_start.p, n, c->k,
                           c->rrho.p, c->rvel[s].p, set_volume ? 1 : 0);
    VPS_HIP(c, hipGetLastError());
    return VPS_OK;
}

// FP32 cell offsets of the (sorted) particles (3D).
int real_offsets(vps_context* c) {
    const int64_t n = c->n;
    hipLaunchKernelGGL(k_cell_offsets3, dim3(blocks(n, 256)), dim3(256), 0, c->stream,
                       c->rpos[c->cur].p, n, c->k, c->rofs.p);
    VPS_HIP(c, hipGetLastError());
    return VPS_OK;
}

// Terms of the FP32 density walks (test support): set 0 the real density
// (real particles), 1 the boundary density (boundary particles, from offsets
// recomputed with the cell index); members in the sorted order.
int density_member_counts(vps_context* c, int which, int32_t* members) {
    if (c->dim != 3 || (which != 0 && which != 1)) return VPS_ERR_INVALID_ARG;
    const int64_t n = which == 0 ? c->n : c->nb;
    if (n == 0) return VPS_OK;
    if (which == 0) {
        const int s = c->cur;
        hipLaunchKernelGGL(k_cell_offsets3, dim3(blocks(n, 256)), dim3(256), 0, c->stream,
                           c->rpos[s].p, n, c->k, c->rofs.p);
        hipLaunchKernelGGL((k_density3f<false, uint64_t, true>), dim3(blocks(n, 256)),
                           dim3(256), 0, c->stream, c->rpos[s].p, c->rofs.p, c->cell_start.p,
                           n, c->k, (float*)nullptr, (float*)nullptr, (float4*)nullptr,
                           (float4*)nullptr, 0, members);
    } else {
        DBuf<float4> ofs;
        VPS_TRY(dalloc(c, ofs, (size_t)n));
        hipLaunchKernelGGL(k_cell_offsets3, dim3(blocks(n, 256)), dim3(256), 0, c->stream,
                           c->bpos.p, n, c->k, ofs.p);
        hipLaunchKernelGGL((k_density3f<true, uint64_t, true>), dim3(blocks(n, 256)),
                           dim3(256), 0, c->stream, c->bpos.p, ofs.p, c->bcell_start.p, n,
                           c->k, (float*)nullptr, (float*)nullptr, (float4*)nullptr,
                           (float4*)nullptr, 0, members);
        const hipError_t e = hipStreamSynchronize(c->stream);
        dfree(ofs);
        VPS_HIP(c, e);
    }
    VPS_HIP(c, hipGetLastError());
    return VPS_OK;
}

int boundary_density(vps_context* c) {
    const int64_t nb = c->nb;
    if (c->dim == 3)
    {
        // FP32 cell offsets first (w = +-1), the density walk as for the real
        // particles, then the offsets with w = +-1 / rho_b for the virtual sums
        VPS_TRY(dalloc(c, c->bofs, (size_t)nb));
        hipLaunchKernelGGL(k_cell_offsets3, dim3(blocks(nb, 256)), dim3(256), 0,
                           c->stream, c->bpos.p, nb, c->k, c->bofs.p);
        if (nb * 16 <= (int64_t)UINT32_MAX)
            hipLaunchKernelGGL((k_density3f<true, uint32_t>), dim3(blocks(nb, 256)), dim3(256),
                               0, c->stream, c->bpos.p, c->bofs.p, c->bcell_start.p, nb, c->k,
                               (float*)nullptr, (float*)nullptr, c->bnrm.p, c->bvel.p, 0);
        else
            hipLaunchKernelGGL((k_density3f<true, uint64_t>), dim3(blocks(nb, 256)), dim3(256),
                               0, c->stream, c->bpos.p, c->bofs.p, c->bcell_start.p, nb, c->k,
                               (float*)nullptr, (float*)nullptr, c->bnrm.p, c->bvel.p, 0);
        hipLaunchKernelGGL(k_boundary_offsets3, dim3(blocks(nb, 256)), dim3(256), 0,
                           c->stream, c->bpos.p, c->bnrm.p, nb, c->k, c->bofs.p);
    }
    else
        hipLaunchKernelGGL(k_bdensity<2>, dim3(blocks(nb, 256)), dim3(256), 0,
                           c->stream, c->bpos.p, c->bcell_start.p, nb, c->k,
                           c->bnrm.p, c->bvel.p);
    VPS_HIP(c, hipGetLastError());
    return VPS_OK;
}

}  // namespace vps
