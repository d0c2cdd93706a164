// vps_stencil.hip -- matrix-free virtual Laplacian, its diagonal, and CG.
//
// Replaces VirtualParticle::calculateCoef / setAlpha
// (Cores/Structures/VirtualParticle.cpp:64-73, 149-169) and
// IterativeSolver::solve -> Eigen ConjugateGradient
// (Cores/Solvers/IterativeSolver.h:19-40).
//
// Virtual particles sit on lattice nodes, so for J in N_H(I) the distance
// r_IJ depends only on the integer offset; with H = 4 spacings the
// neighbourhood is the 251-offset ball |off|^2 < 16 in 3D (45 in 2D; the
// |off|^2 = 16 ties carry W'(H) = 0).  The assembled matrix
//     A_IJ = 2 V_J W'_H(r) / (r + 1e-6)       (J != I)
//     A_II = max(s_I, alpha0),   s_I = -sum_J A_IJ
// is never stored: with cc[k2] = 2 W'_H(sqrt(k2) s) / (sqrt(k2) s + 1e-6)
//     (A p)_I   = A_II p_I + sum_off cc V_J p_J         (VPS_OP_ROW)
//     (A^T p)_I = A_II p_I + V_I sum_off cc p_J         (VPS_OP_TRANSPOSE)
// and A^T is what Eigen's CG runs (column-major matrix, Lower|Upper).
//
// Kernel shape (3D): one 64-lane wave per 8x8x8 brick.  The 14^3 halo tile
// (the brick plus 3 nodes on each side, read from the 27 neighbouring
// bricks) is staged in LDS with unconditional loads: a constant table maps
// each tile element to (neighbour brick, local node), and absent bricks are
// redirected to an all-zero brick kept at the end of every per-node array.
// Lane (x, y) owns the z column of 8 outputs.  Each LDS plane is reduced
// once into four plane sums P_k (k = |dz|) using the 8-fold symmetry of cc
// (the (+-a, +-b) points of a class are added first and multiplied once per
// k), and a 7-slot sliding window turns plane sums into outputs.
//
// CG (Eigen's algorithm, identity preconditioner, recursive residual):
//   init      r = b - M x0, p = r, b.b, r.r     (stencil launch + 1-block
//                                                finish)
//   iterate   SpMV(k):   rr_{k-1} -> converged? beta; p = r + beta p_old
//                        (fused into the halo load, p double-buffered);
//                        t = M p; partial p.t per block
//             update(k): alpha = rr_{k-1} / p.t; x += alpha p;
//                        r -= alpha t; partial r.r per block
// 3D: one launch per iteration (k_cg_iter3, persistent blocks) forms the
// sums of the previous launch's partials, applies the update and the SpMV.
// Every grid-wide sum is formed by the CONSUMING kernel: its blocks re-sum
// the producer's per-block partials in a fixed order (2D: a ring of three
// iterations; 3D: two alternating sets).  No device-scope fence or atomic is issued inside the loop --
// on a multi-XCD part an agent-scope release writes back the L2 -- and the
// result is bitwise reproducible.  FP32 vectors, FP64 dots and norms.  At
// convergence the FP64 true residual ||b - M x|| is checked; if it is above
// the tolerance CG restarts from the true residual (residual replacement).
#include <cfloat>
#include <algorithm>
#include <cmath>

#include <chrono>

#include "vps_internal.h"

namespace vps {

struct CoefF {
    float c[16];
};
struct CoefD {
    double c[16];
};

namespace {

typedef float f32x4_t __attribute__((ext_vector_type(4)));
// Streaming (nontemporal) 16-byte store.
__device__ __forceinline__ void store_nt4(float* p, float a, float b, float c,
                                          float d) {
    f32x4_t v = {a, b, c, d};
    __builtin_nontemporal_store(v, (f32x4_t*)p);
}

constexpr int kT = 14;             // halo tile side (8 + 2 * 3)
constexpr int kT2 = kT * kT;
constexpr int kHaloIters2 = (kT2 + 63) / 64;   // 4
constexpr int kUpdBlocks = 256;                // grid of the CG update
constexpr int kUpdThreads = 1024;
constexpr int kIterBlocks = 1024;                // persistent blocks of the 3D iteration

// 2D tile element e -> (neighbour brick 0..8) << 9 | local node; elements
// past the tile map to the zero brick slot 9.  (3D decodes arithmetically.)
struct HaloTab2 {
    uint16_t v[kHaloIters2 * 64];
};
constexpr HaloTab2 make_halo2() {
    HaloTab2 t{};
    for (int e = 0; e < kHaloIters2 * 64; e++) {
        if (e >= kT2) {
            t.v[e] = (uint16_t)(9 << 9);
            continue;
        }
        const int ux = e % kT + 5, uy = e / kT + 5;
        const int nb = (uy >> 3) * 3 + (ux >> 3);
        const int loc = (ux & 7) + 8 * (uy & 7);
        t.v[e] = (uint16_t)((nb << 9) | loc);
    }
    return t;
}
__constant__ HaloTab2 kHalo2 = make_halo2();

enum Kind {
    K_DIAG = 0,        // s = -sum cc V_J; diag = max(s, alpha0)   (double)
    K_APPLY = 1,       // out = M in                                 (float)
    K_CG_INIT = 2,     // r = b - M x0; p = r; partial b.b, r.r     (float)
    K_SPMV_FIRST = 3,  // t = M p; partial p.t                       (float)
    K_SPMV_FUSED = 4,  // p = r + beta p_old; t = M p; partial p.t   (float)
    K_TRUE_RES = 5     // rt = b - M x (double); partial rt.rt      (double)
    // (the 3D CG iteration is k_cg_iter3)
};

struct StencilArgs {
    KConst k;
    const int32_t* slot_of_brick;
    const int32_t* brick_of_slot;
    const int32_t* nbr_slot;   // [slot][27] neighbour slots (3D)
    int64_t nslots;
    int64_t zero_slot;     // index of the all-zero brick
    const float* vvol;
    const float* vdiag;
    const double* vdiag64;
    const float* vb;
    const float* in;       // p / x / x0 / in (p_old when fused)
    const double* in_base; // K_TRUE_RES: FP64 base of x (x = in_base + in)
    const float* in2;      // r for the fused update (r_old: 3D iteration)
    const float* in3;      // t_old (3D iteration)
    float* out;            // t / r / out
    float* out2;           // p_new (fused) or p (init)
    float* out3;           // r_new (3D iteration)
    float* xio;            // x, updated in place (3D iteration)
    double* out_d;         // diag64
    float* out_f;          // diag32
    float* out_s;          // unclamped row sum s (feature of the learned start)
    double alpha0;
    CgState* st;
    double* part0;         // this kernel's per-block partials (sum 0)
    double* part1;         // per-block partials (sum 1: init, FIRST, UPD)
    double* part2;         // per-block partials (sum 2: FIRST, UPD)
    double* part3;         // per-block partials (sum 3: FIRST, UPD)
    const double* rr1;     // update partials of iteration kk-1
    const double* rr2;     // update partials of iteration kk-2
    int kk;                // iteration index since the (re)start
    int iter_base;         // Eigen iteration count at the (re)start
    int64_t max_iter;
    int spec;              // K_TRUE_RES: run only if the solve has converged
};

// Sum of n per-block partials, identical in every block and lane: lane l
// adds partials l, l + 64, ... in order, then a fixed DPP tree.
__device__ __forceinline__ double sum_partials64(const double* p, int n) {
    const int lane = threadIdx.x & 63;
    double s = 0;
    for (int i = lane; i < n; i += 64) s += p[i];
    return wave_sum_all(s);
}

template <int BS>
__device__ __forceinline__ double block_sum(double v, double* sbuf) {
    v = wave_sum_all(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sbuf[wv] = v;
    __syncthreads();
    double t = 0;
#pragma unroll
    for (int i = 0; i < BS / 64; i++) t += sbuf[i];
    return t;
}

// Plane sums P_k, k = 0..3, at the LDS point pl (row stride RS).  The
// in-plane points are first summed by squared radius m = a^2 + b^2 (the nine
// values 0, 1, 2, 4, 5, 8, 9, 10, 13 of the disk m < 16), then
// P_k = sum_m cc[m + k^2] U_m over m + k^2 < 16 (k^2 = 0 excluded).
template <int RS, typename T, typename C>
__device__ __forceinline__ void plane4(const T* pl, const C& cc, T P[4]) {
    auto at = [&](int a, int b) { return pl[a + b * RS]; };
    // the 4 points (+-a, 0), (0, +-a)
    auto axis4 = [&](int a) { return (at(a, 0) + at(-a, 0)) + (at(0, a) + at(0, -a)); };
    // the 4 points (+-a, +-a)
    auto diag4 = [&](int a) { return (at(a, a) + at(-a, a)) + (at(a, -a) + at(-a, -a)); };
    // the 8 points (+-a, +-b), (+-b, +-a), a != b
    auto ring8 = [&](int a, int b) {
        return ((at(a, b) + at(-a, b)) + (at(a, -b) + at(-a, -b))) +
               ((at(b, a) + at(-b, a)) + (at(b, -a) + at(-b, -a)));
    };
    const T U0 = pl[0], U1 = axis4(1), U2 = diag4(1), U4 = axis4(2), U5 = ring8(1, 2),
            U8 = diag4(2), U9 = axis4(3), U10 = ring8(1, 3), U13 = ring8(2, 3);
    const T c1 = (T)cc.c[1], c2 = (T)cc.c[2], c3 = (T)cc.c[3], c4 = (T)cc.c[4],
            c5 = (T)cc.c[5], c6 = (T)cc.c[6], c8 = (T)cc.c[8], c9 = (T)cc.c[9],
            c10 = (T)cc.c[10], c11 = (T)cc.c[11], c12 = (T)cc.c[12], c13 = (T)cc.c[13],
            c14 = (T)cc.c[14];
    P[0] = c1 * U1 + c2 * U2 + c4 * U4 + c5 * U5 + c8 * U8 + c9 * U9 + c10 * U10 +
           c13 * U13;
    P[1] = c1 * U0 + c2 * U1 + c3 * U2 + c5 * U4 + c6 * U5 + c9 * U8 + c10 * U9 +
           c11 * U10 + c14 * U13;
    P[2] = c4 * U0 + c5 * U1 + c6 * U2 + c8 * U4 + c9 * U5 + c12 * U8 + c13 * U9 +
           c14 * U10;
    P[3] = c9 * U0 + c10 * U1 + c11 * U2 + c13 * U4 + c14 * U5;
}

// 2D: in-plane sum only.
template <typename T, typename C>
__device__ __forceinline__ T plane1(const T* pl, const C& cc) {
    T P = T(0);
#pragma unroll
    for (int a = 0; a <= 3; a++) {
#pragma unroll
        for (int b = 0; b <= 3; b++) {
            const int k2 = a * a + b * b;
            if (k2 >= 16 || k2 == 0) continue;
            T u;
            if (a == 0) u = pl[b * kT] + pl[-b * kT];
            else if (b == 0) u = pl[a] + pl[-a];
            else
                u = (pl[a + b * kT] + pl[-a + b * kT]) +
                    (pl[a - b * kT] + pl[-a - b * kT]);
            P += (T)cc.c[k2] * u;
        }
    }
    return P;
}

// Prologue of the CG SpMV: convergence test of iteration kk-1, beta.
// Returns false when the solve is over (block 0 records why).
__device__ __forceinline__ bool spmv_prologue(const StencilArgs& a,
                                              float* beta) {
    CgState* st = a.st;
    if (st->done) return false;
    *beta = 0.f;
    if (a.kk == 0) return true;
    const double r1 = sum_partials64(a.rr1, kUpdBlocks);
    const double r2 = a.kk >= 2 ? sum_partials64(a.rr2, kUpdBlocks) : st->abs0;
    const int k = a.iter_base + a.kk;
    int done = 0;
    if (!isfinite(r1)) done = 3;
    else if (r1 < st->thr) done = 1;
    else if (k >= a.max_iter) done = 2;
    if (done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->rr = r1;
            st->iter = done == 1 ? k - 1 : k;
            st->done = done;
        }
        return false;
    }
    *beta = (float)(r1 / r2);
    return true;
}

template <int KIND, int OP, typename T>
__device__ __forceinline__ T load_input(const StencilArgs& a, uint32_t g,
                                        float beta, float* pnew_out) {
    if constexpr (KIND == K_DIAG) {
        return (T)a.vvol[g];
    } else {
        float p;
        if constexpr (KIND == K_SPMV_FUSED) p = a.in2[g] + beta * a.in[g];
        else p = a.in[g];
        *pnew_out = p;
        if constexpr (OP == VPS_OP_ROW) return (T)a.vvol[g] * (T)p;
        else return (T)p;
    }
}

// 3D: one workgroup of 4 waves per 8^3 brick.  The radius-3 halo is loaded
// as 16-byte half rows of the neighbour bricks (x-fastest 8^3 layout): every
// (y, z) line of the tile is the upper half row of the left brick, the
// center row and the lower half row of the right brick (x = -4 .. 11), so
// loads are float4 and contiguous within each brick.  The LDS tile keeps
// rows of 24 floats (x = -8 .. 15, stride chosen so that the 8 x 8 columns
// of a wave fall in distinct banks).  The 14 plane sums P_0..P_3 of every
// (x, y) column are computed once (plane zp by wave zp mod 4) into LDS, and
// wave w then forms the outputs of planes z = 2w, 2w + 1 as
// P_3(z) + P_2(z+1) + P_1(z+2) + P_0(z+3) + P_1(z+4) + P_2(z+5) + P_3(z+6)
// (tile planes, in that order).
constexpr int kS3Threads = 256;
constexpr int kRS = 24;                     // tile row stride (floats)
constexpr int kPS = kRS * kT;               // tile plane stride
constexpr int kUnits = kT * kT * 4;         // 16-byte loads per tile: 784
constexpr int kUnitIters = (kUnits + kS3Threads - 1) / kS3Threads;   // 4
static_assert(kS3Threads * (kUnitIters - 1) <= kUnits, "full unit rounds");
constexpr int kChainUnit0 = 6 * kT * 4;         // first unit of tile plane 6: 336
constexpr int kOwnX = 11 * 64;                  // own nodes of tile planes 3..13

template <int KIND, int OP, typename T, typename C>
__global__ void __launch_bounds__(kS3Threads) k_stencil3(StencilArgs a, C cc) {
    __shared__ T tile[kT * kPS];
    // plane sums P_k of plane zp, column l; reused for the first SpMV's
    // block partials after the outputs (8-byte aligned for that)
    __shared__ __align__(16) T psum[kT][4][64];
    __shared__ float own[512];      // raw own-brick input (p / x)
    // K_TRUE_RES: the own brick's x = base + correction in FP64
    __shared__ double own64[KIND == K_TRUE_RES ? 512 : 1];
    __shared__ uint32_t sbase[27];
    __shared__ double sred[4][4];
    __shared__ int s_done;
    const int slot = xcd_remap(blockIdx.x, (int)a.nslots);
    const int t = threadIdx.x;
    const int l = t & 63, wv = t >> 6;
    if (t < 27) sbase[t] = (uint32_t)a.nbr_slot[(int64_t)slot * 27 + t] * 512u;
    if constexpr (KIND == K_SPMV_FIRST || KIND == K_SPMV_FUSED || KIND == K_TRUE_RES) {
        // a solve that has ended skips the halo loads (launches past the
        // convergence of a batch); a speculative true-residual check runs
        // only once the recursive residual has converged
        if (t == 32) s_done = a.st->done;
    }
    const int x = l & 7, y = l >> 3;
    const int64_t base = (int64_t)slot * 512;
    // epilogue operands of this lane's 2 outputs, fetched early
    float Vz[2], Dz[2], Bz[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int64_t g = base + x + 8 * y + 64 * (2 * wv + q);
        Vz[q] = a.vvol[g];
        Dz[q] = (KIND != K_DIAG && KIND != K_TRUE_RES) ? a.vdiag[g] : 0.f;
        Bz[q] = (KIND == K_CG_INIT || KIND == K_TRUE_RES) ? a.vb[g] : 0.f;
    }
    __syncthreads();
    if constexpr (KIND == K_SPMV_FIRST || KIND == K_SPMV_FUSED) {
        if (s_done) return;   // uniform over the block
    }
    if constexpr (KIND == K_TRUE_RES) {
        if (a.spec && s_done != 1) return;   // uniform over the block
    }
    // Halo load: unit u = (line (yt, zt), part k): k = 0 the upper half row of
    // the left brick (x = -4..-1), 1 and 2 the center row, 3 the lower half
    // row of the right brick (x = 8..11).  The raw loads do not depend on
    // beta, so they are issued before the prologue's reductions.
    float4 ra[kUnitIters], rb[kUnitIters], rv[kUnitIters];
    double4 rx[KIND == K_TRUE_RES ? kUnitIters : 1];
#pragma unroll
    for (int i = 0; i < kUnitIters; i++) {
        const int u = t + kS3Threads * i;
        ra[i] = rb[i] = rv[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        if constexpr (KIND == K_TRUE_RES) rx[i] = make_double4(0.0, 0.0, 0.0, 0.0);
        // units of the first kUnitIters - 1 rounds always exist (t < 256)
        if (i < kUnitIters - 1 || u < kUnits) {
            const int k = u & 3, line = u >> 2;
            const int yt = line % kT, zt = line / kT;
            const int yy = yt + 5, zz = zt + 5;   // + 8 - 3: 0..23
            const int bxo = k == 0 ? 0 : (k == 3 ? 2 : 1);
            const int nb = (zz >> 3) * 9 + (yy >> 3) * 3 + bxo;
            const int xl = (k == 0 || k == 2) ? 4 : 0;
            const uint32_t g = sbase[nb] + (uint32_t)(xl + 8 * ((yy & 7) + 8 * (zz & 7)));
            if constexpr (KIND == K_DIAG) {
                rv[i] = *(const float4*)(a.vvol + g);
            } else {
                ra[i] = *(const float4*)(a.in + g);
                if constexpr (KIND == K_SPMV_FUSED) rb[i] = *(const float4*)(a.in2 + g);
                if constexpr (KIND == K_TRUE_RES) rx[i] = *(const double4*)(a.in_base + g);
                if constexpr (OP == VPS_OP_ROW) rv[i] = *(const float4*)(a.vvol + g);
            }
        }
    }
    float beta = 0.f;
    if constexpr (KIND == K_SPMV_FIRST || KIND == K_SPMV_FUSED) {
        if (!spmv_prologue(a, &beta)) return;   // uniform over the grid
    }
#pragma unroll
    for (int i = 0; i < kUnitIters; i++) {
        const int u = t + kS3Threads * i;
        if (i < kUnitIters - 1 || u < kUnits) {
            const int k = u & 3, line = u >> 2;
            const int yt = line % kT, zt = line / kT;
            T v[4];
            float pn[4];
            if constexpr (KIND == K_DIAG) {
                v[0] = (T)rv[i].x; v[1] = (T)rv[i].y; v[2] = (T)rv[i].z; v[3] = (T)rv[i].w;
                pn[0] = pn[1] = pn[2] = pn[3] = 0.f;
            } else {
                float4 p = ra[i];
                if constexpr (KIND == K_SPMV_FUSED)
                    p = make_float4(rb[i].x + beta * ra[i].x, rb[i].y + beta * ra[i].y,
                                    rb[i].z + beta * ra[i].z, rb[i].w + beta * ra[i].w);
                pn[0] = p.x; pn[1] = p.y; pn[2] = p.z; pn[3] = p.w;
                T xv[4] = {(T)p.x, (T)p.y, (T)p.z, (T)p.w};
                if constexpr (KIND == K_TRUE_RES) {
                    xv[0] += rx[i].x; xv[1] += rx[i].y; xv[2] += rx[i].z; xv[3] += rx[i].w;
                }
                if constexpr (OP == VPS_OP_ROW) {
                    v[0] = (T)rv[i].x * xv[0]; v[1] = (T)rv[i].y * xv[1];
                    v[2] = (T)rv[i].z * xv[2]; v[3] = (T)rv[i].w * xv[3];
                } else {
                    v[0] = xv[0]; v[1] = xv[1]; v[2] = xv[2]; v[3] = xv[3];
                }
                if constexpr (KIND == K_TRUE_RES) {
                    if ((k == 1 || k == 2) && yt >= 3 && yt < 11 && zt >= 3 && zt < 11) {
                        const int loc = 4 * (k - 1) + 8 * ((yt - 3) + 8 * (zt - 3));
#pragma unroll
                        for (int c = 0; c < 4; c++) own64[loc + c] = (double)xv[c];
                    }
                }
            }
            T* dst = &tile[zt * kPS + yt * kRS + 4 + 4 * k];
#pragma unroll
            for (int c = 0; c < 4; c++) dst[c] = v[c];
            if constexpr (KIND != K_DIAG) {
                if ((k == 1 || k == 2) && yt >= 3 && yt < 11 && zt >= 3 && zt < 11) {
                    const int loc = 4 * (k - 1) + 8 * ((yt - 3) + 8 * (zt - 3));
#pragma unroll
                    for (int c = 0; c < 4; c++) own[loc + c] = pn[c];
                    if constexpr (KIND == K_SPMV_FUSED)
                        store_nt4(a.out2 + base + loc, pn[0], pn[1], pn[2], pn[3]);
                }
            }
        }
    }
    __syncthreads();
    // plane sums, each plane once; column (x, y) sits at tile x = x + 8
    for (int zp = wv; zp < kT; zp += 4) {
        T P[4];
        plane4<kRS, T>(&tile[zp * kPS + (y + 3) * kRS + (x + 8)], cc, P);
#pragma unroll
        for (int q = 0; q < 4; q++) psum[zp][q][l] = P[q];
    }
    __syncthreads();
    double red0 = 0, red1 = 0, red2 = 0, red3 = 0;
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int z = 2 * wv + q;
        T acc = T(0);
        acc += psum[z][3][l];
        acc += psum[z + 1][2][l];
        acc += psum[z + 2][1][l];
        acc += psum[z + 3][0][l];
        acc += psum[z + 4][1][l];
        acc += psum[z + 5][2][l];
        acc += psum[z + 6][3][l];
        const int loc = x + 8 * y + 64 * z;
        const int64_t g = base + loc;
        const float V = Vz[q], Dg = Dz[q], Bg = Bz[q];
        const bool on = V > 0.f;
        if constexpr (KIND == K_DIAG) {
            const double sd = -(double)acc;
            const double d = on ? (sd < a.alpha0 ? a.alpha0 : sd) : 0.0;
            a.out_d[g] = d;
            a.out_f[g] = (float)d;
            a.out_s[g] = on ? (float)sd : 0.f;
        } else if constexpr (KIND == K_TRUE_RES) {
            const double xi = own64[loc];
            const double st = OP == VPS_OP_ROW ? (double)acc : (double)V * (double)acc;
            const double mx = a.vdiag64[g] * xi + st;
            const double rt = on ? (double)Bg - mx : 0.0;
            a.out[g] = (float)rt;
            red0 += rt * rt;
        } else {
            const float pi = own[loc];
            const float st = OP == VPS_OP_ROW ? (float)acc : V * (float)acc;
            const float mp = on ? Dg * pi + st : 0.f;
            if constexpr (KIND == K_APPLY) {
                a.out[g] = mp;
            } else if constexpr (KIND == K_CG_INIT) {
                const float b = Bg;
                const float r = on ? b - mp : 0.f;
                a.out[g] = r;
                a.out2[g] = r;
                red0 += (double)b * (double)b;
                red1 += (double)r * (double)r;
            } else {
                // t = M p is read by the next kernel only: streaming store
                __builtin_nontemporal_store(mp, &a.out[g]);
                red0 += (double)pi * (double)mp;
                if constexpr (KIND == K_SPMV_FIRST) red3 += (double)pi * (double)pi;
                if constexpr (KIND == K_SPMV_FIRST)
                    red2 += (double)mp * (double)mp;
            }
        }
    }
    if constexpr (KIND == K_SPMV_FIRST) {
        // four block sums (p.t, r.t, t.t, r.r; r = p after SPMV_FIRST): the
        // lanes' values through LDS, then wave w reduces sum w -- one DPP
        // tree per wave instead of four
        static_assert(sizeof(psum) >= 4 * kS3Threads * sizeof(double), "partials fit psum");
        __syncthreads();   // every wave has read its plane sums
        double* sp = reinterpret_cast<double*>(&psum[0][0][0]);
        sp[t] = red0;
        sp[kS3Threads + t] = KIND == K_SPMV_FIRST ? red0 : red1;
        sp[2 * kS3Threads + t] = red2;
        sp[3 * kS3Threads + t] = red3;
        __syncthreads();
        const double* q = sp + kS3Threads * wv;
        double v = (q[l] + q[l + 64]) + (q[l + 128] + q[l + 192]);
        v = wave_sum_all(v);
        double* dst = wv == 0 ? a.part0 : (wv == 1 ? a.part1 : (wv == 2 ? a.part2 : a.part3));
        if (l == 0 && dst) dst[blockIdx.x] = v;
    } else if constexpr (KIND != K_DIAG && KIND != K_APPLY) {
        constexpr bool three = false;
        red0 = wave_sum_all(red0);
        if (KIND == K_CG_INIT) red1 = wave_sum_all(red1);
        if (three) {
            red2 = wave_sum_all(red2);
            red3 = wave_sum_all(red3);
        }
        if (l == 0) {
            sred[0][wv] = red0;
            sred[1][wv] = red1;
            sred[2][wv] = red2;
            sred[3][wv] = red3;
        }
        __syncthreads();
        if (t == 0) {
            const double s0 = (sred[0][0] + sred[0][1]) + (sred[0][2] + sred[0][3]);
            a.part0[blockIdx.x] = s0;
            if (KIND == K_CG_INIT)
                a.part1[blockIdx.x] = (sred[1][0] + sred[1][1]) + (sred[1][2] + sred[1][3]);
            if (three && a.part2) {
                if (KIND == K_SPMV_FIRST) a.part1[blockIdx.x] = s0;
                a.part2[blockIdx.x] = (sred[2][0] + sred[2][1]) + (sred[2][2] + sred[2][3]);
                a.part3[blockIdx.x] = (sred[3][0] + sred[3][1]) + (sred[3][2] + sred[3][3]);
            }
        }
    }
}

// 3D CG iteration with persistent blocks (G <= kIterBlocks, a multiple of 8
// when G >= 8, every block owning one G-th of the slots, consecutive in
// column order: it walks up z-runs of bricks).  Each block
// first forms the global sums of the previous launch's block partials itself
// -- the same loads, order and arithmetic in every block, so every block
// reaches the same alpha_(kk-1) = r.r / p.t, the r.r after the update
// (expanded as r.r - 2 alpha r.t + alpha^2 t.t, anchored every iteration on
// the direct r.r of the stored r), Eigen's test on it and beta_kk
// without a reduction launch in between (about 1K partials per set, one
// round of loads).  Then per brick: x += alpha p_old, r = r_old - alpha t_old,
// p = r + beta p_old, t = M p with k_stencil3's tile and plane sums, and one partial
// per block of p.t, r.t, t.t and r.r.  A brick that sits on the block's
// previous brick shares 6 of its 14 halo planes with it: their plane sums
// and the updated p, r of its z = 0..2 (the previous brick's halo planes
// 11..13, stored by it) stay in LDS, and only planes 6..13 are loaded and
// summed.  The launch that ends the solve
// applies only the last x update; block 0 records the state.
static_assert(kIterBlocks <= 4 * kS3Threads, "one round of partial loads per thread");
struct IterArgs {
    const double* prev[4];   // the previous launch's partials (p.t, r.t, t.t, r.r)
    double* out[4];          // this launch's partials, one per block
    int nprev;               // their count
    int resume;              // alpha, beta from the state (after a replacement)
};

// Block sums of four per-thread values v[0..3] in a fixed order: the 4 x 256
// values go through LDS scratch (8 KB), wave w adds the four values of lane
// l, l + 64, l + 128, l + 192 of sum w and reduces them with one DPP tree;
// out[w] (LDS) holds sum w for the whole block afterwards.
__device__ __forceinline__ void block_sums4(const double v[4], double* scratch, double* out) {
    const int t = threadIdx.x, l = t & 63, wv = t >> 6;
#pragma unroll
    for (int q = 0; q < 4; q++) scratch[q * kS3Threads + t] = v[q];
    __syncthreads();
    const double* a = scratch + kS3Threads * wv;
    double s = (a[l] + a[l + 64]) + (a[l + 128] + a[l + 192]);
    s = wave_sum_all(s);
    if (l == 0) out[wv] = s;
    __syncthreads();
}

template <int OP>
__global__ void __launch_bounds__(kS3Threads) k_cg_iter3(StencilArgs a, CoefF cc, IterArgs it,
                                                          const int32_t* __restrict__ order) {
    __shared__ float tile[kT * kPS];
    __shared__ __align__(16) float psum[kT][4][64];
    // p and r of the own nodes by tile plane 3..13: planes 3..10 are the
    // brick's own z = 0..7, planes 11..13 the next brick's z = 0..2 when it
    // is the dz = +1 neighbour
    __shared__ __align__(16) float own[kOwnX];
    __shared__ __align__(16) float own_r[kOwnX];
    __shared__ uint32_t sbase[2][27];
    __shared__ double sred[4][4];
    __shared__ float s_ab[2];
    __shared__ int s_mode;
    CgState* st = a.st;
    const int kk = a.kk;
    const int t = threadIdx.x;
    // Everything the prologue reads goes out at once: the stop tag and the
    // state (written by earlier launches only, so plain loads), the previous
    // launch's partials (at most 4 x 256 per sum) and this block's first
    // slots.  Every lane loads (indices past the partials read the last
    // one): loads under a condition make the compiler's waits for them
    // cover every later load.
    const int stop = st->stop;
    const double thr_rec = st->thr_rec;
    const float st_alpha = st->alpha, st_beta = st->beta;
    double v4[4][4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int i = min(t + kS3Threads * u, it.nprev - 1);
#pragma unroll
        for (int q = 0; q < 4; q++) v4[q][u] = it.prev[q][i];
    }
    const int l = t & 63, wv = t >> 6;
    const int x = l & 7, y = l >> 3;
    const int nsl = (int)a.nslots;
    const int G = (int)gridDim.x;
    // this block's bricks: positions [lo, hi) of the column order, so that a
    // block walks up z-runs (the blocks of one XCD hold consecutive ranges)
    const int rbk = (G & 7) == 0 ? xcd_remap(blockIdx.x, G) : (int)blockIdx.x;
    const int lo = __builtin_amdgcn_readfirstlane((int)((int64_t)rbk * nsl / G));
    const int hi = __builtin_amdgcn_readfirstlane((int)((int64_t)(rbk + 1) * nsl / G));
    int cur = order[min(lo, nsl - 1)];
    const int nx0 = order[min(lo + 1, nsl - 1)];
    int nxt = lo + 1 < hi ? nx0 : -1;
    // a solve ended by an earlier launch (or before any): nothing to do
    if (stop != 0 && stop != kk + 1) return;   // uniform over the grid
    {
        const int nb0 = a.nbr_slot[(int64_t)cur * 27 + min(t, 26)];
        if (t < 27) sbase[0][t] = (uint32_t)nb0 * 512u;
    }
    // A brick's loads: the epilogue's operands at its two own nodes, and the
    // halo (see k_stencil3) from unit u0 on -- a chained brick loads tile
    // planes 6..13 only: planes 0..5 are the previous brick's 8..13, whose
    // plane sums it copies.
    float4 ra[kUnitIters], rb[kUnitIters], rv[kUnitIters], rc[kUnitIters];
    float Vz[2], Dz[2], Pz[2], Xz[2];
    auto issue = [&](int64_t base, const uint32_t* sbc, int u0) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int64_t g = base + x + 8 * y + 64 * (2 * wv + q);
            Vz[q] = a.vvol[g];
            Dz[q] = a.vdiag[g];
            Pz[q] = a.in[g];
            Xz[q] = a.xio[g];
        }
#pragma unroll
        for (int i = 0; i < kUnitIters; i++) {
            const int u = u0 + t + kS3Threads * i;
            ra[i] = rb[i] = rv[i] = rc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (u < kUnits) {
                const int k = u & 3, line = u >> 2;
                const int yt = line % kT, zt = line / kT;
                const int yy = yt + 5, zz = zt + 5;
                const int bxo = k == 0 ? 0 : (k == 3 ? 2 : 1);
                const int nb = (zz >> 3) * 9 + (yy >> 3) * 3 + bxo;
                const int xl = (k == 0 || k == 2) ? 4 : 0;
                // byte offset in 32 bits (the node arrays stay below 4 GB):
                // the loads take the array base from scalar registers
                const uint32_t g =
                    4u * (sbc[nb] + (uint32_t)(xl + 8 * ((yy & 7) + 8 * (zz & 7))));
                ra[i] = *(const float4*)((const char*)a.in + g);
                rb[i] = *(const float4*)((const char*)a.in2 + g);
                rc[i] = *(const float4*)((const char*)a.in3 + g);
                if constexpr (OP == VPS_OP_ROW)
                    rv[i] = *(const float4*)((const char*)a.vvol + g);
            }
        }
    };
    // alpha_(kk-1), beta_kk
    if (it.resume) {
        if (t == 0) {
            s_mode = 0;
            s_ab[0] = st_alpha;
            s_ab[1] = st_beta;
        }
    } else {
        double sv[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            double w[4];
#pragma unroll
            for (int u = 0; u < 4; u++) w[u] = t + kS3Threads * u < it.nprev ? v4[q][u] : 0.0;
            sv[q] = (w[0] + w[1]) + (w[2] + w[3]);
        }
        // the four sums through LDS (psum is free here): wave w reduces sum
        // w with one DPP tree instead of every wave reducing all four
        block_sums4(sv, reinterpret_cast<double*>(&psum[0][0][0]), sred[0]);
        if (t == 0) {
            double sum[4];
#pragma unroll
            for (int q = 0; q < 4; q++) sum[q] = sred[0][q];
            const double pt = sum[0], rt = sum[1], tt = sum[2], rr = sum[3];
            const int k = a.iter_base + kk - 1;   // Eigen's index of the reduced SpMV
            const double ald = rr / pt;
            const float al = (float)ald;
            const double a64 = (double)al;   // the alpha the vectors are updated with
            const double rn = (rr - 2.0 * a64 * rt) + a64 * a64 * tt;
            int done = 0, iter = 0;
            float be = 0.f;
            if (!isfinite(ald) || !isfinite(rn)) {
                done = 3;
                iter = k;
            } else if (rn < thr_rec) {
                done = 1;
                iter = k;
            } else if (k + 1 >= a.max_iter) {
                done = 2;
                iter = k + 1;
            } else {
                be = (float)(rn / rr);
            }
            s_mode = done == 0 ? 0 : (done == 3 ? 2 : 1);
            s_ab[0] = al;
            s_ab[1] = be;
            if (blockIdx.x == 0) {
                st->alpha = al;
                st->pt = pt;
                st->rr_prev = rr;
                st->rr = rn;
                if (done) {
                    st->iter = iter;
                    st->done = done;
                    __hip_atomic_store(&st->stop, kk + 1, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                } else {
                    st->beta = be;
                }
            }
        }
    }
    __syncthreads();
    const int mode = s_mode;
    const float al = s_ab[0], be = s_ab[1];
    if (mode == 2) return;   // non-finite alpha: no update
    if (mode == 1) {
        // the solve ended: the last x update only (own nodes of every brick)
        for (int v = blockIdx.x; v < nsl; v += G) {
            const int64_t b0 = (int64_t)((G & 7) == 0 ? xcd_remap(v, nsl) : v) * 512;
            for (int e = t; e < 512; e += kS3Threads)
                a.xio[b0 + e] += al * a.in[b0 + e];
        }
        return;
    }
    double red0 = 0, red1 = 0, red2 = 0, red3 = 0;
    int sb = 0;             // which sbase row is the current brick's
    bool chained = false;   // the current brick sits on the previous one
    for (int p = lo; p < hi; p++) {
        const int64_t base = (int64_t)cur * 512;
        if (p != lo) __syncthreads();   // the previous brick's psum and own have been read
        const uint32_t* sbc = sbase[sb];
        const int u0 = chained ? kChainUnit0 : 0;
        issue(base, sbc, u0);
        // the next brick is the dz = +1 neighbour: it is chained to this one
        const uint32_t up = __builtin_amdgcn_readfirstlane(sbc[22]);
        const bool has_next = nxt >= 0 && up == (uint32_t)nxt * 512u;
        // the brick after next (its slot) and the next brick's neighbour slots
        const int nn2 = order[min(p + 2, nsl - 1)];
        const int nbv = a.nbr_slot[(int64_t)max(nxt, 0) * 27 + min(t, 26)];
        if (chained) {
            float4* ps4 = reinterpret_cast<float4*>(&psum[0][0][0]);
            for (int e = t; e < 6 * 64; e += kS3Threads) ps4[e] = ps4[e + 8 * 64];
        }
#pragma unroll
        for (int i = 0; i < kUnitIters; i++) {
            const int u = u0 + t + kS3Threads * i;
            if (u < kUnits) {
                const int k = u & 3, line = u >> 2;
                const int yt = line % kT, zt = line / kT;
                float rn[4], pn[4];
                rn[0] = rb[i].x - al * rc[i].x;
                rn[1] = rb[i].y - al * rc[i].y;
                rn[2] = rb[i].z - al * rc[i].z;
                rn[3] = rb[i].w - al * rc[i].w;
                pn[0] = rn[0] + be * ra[i].x;
                pn[1] = rn[1] + be * ra[i].y;
                pn[2] = rn[2] + be * ra[i].z;
                pn[3] = rn[3] + be * ra[i].w;
                float* dst = &tile[zt * kPS + yt * kRS + 4 + 4 * k];
                if constexpr (OP == VPS_OP_ROW) {
                    dst[0] = rv[i].x * pn[0];
                    dst[1] = rv[i].y * pn[1];
                    dst[2] = rv[i].z * pn[2];
                    dst[3] = rv[i].w * pn[3];
                } else {
#pragma unroll
                    for (int c = 0; c < 4; c++) dst[c] = pn[c];
                }
                if ((k == 1 || k == 2) && yt >= 3 && yt < 11 && zt >= 3) {
                    // p and r of the own nodes, and of the next brick's
                    // z = 0..2 in planes 11..13
                    const int loc = 4 * (k - 1) + 8 * (yt - 3) + 64 * (zt - 3);
                    if (chained && zt >= 11) {
                        // the previous brick left this brick's z = 0..2 here;
                        // the same thread moves them before overwriting them
#pragma unroll
                        for (int c = 0; c < 4; c++) {
                            own[loc - 512 + c] = own[loc + c];
                            own_r[loc - 512 + c] = own_r[loc + c];
                        }
                    }
#pragma unroll
                    for (int c = 0; c < 4; c++) {
                        own[loc + c] = pn[c];
                        own_r[loc + c] = rn[c];
                    }
                }
            }
        }
        if (t < 27 && nxt >= 0) sbase[sb ^ 1][t] = (uint32_t)nbv * 512u;
        __syncthreads();
        for (int zp = (chained ? 6 : 0) + wv; zp < kT; zp += 4) {
            float P[4];
            plane4<kRS, float>(&tile[zp * kPS + (y + 3) * kRS + (x + 8)], cc, P);
#pragma unroll
            for (int q = 0; q < 4; q++) psum[zp][q][l] = P[q];
        }
        __syncthreads();
        // t = M p, and x, p, r of the own nodes
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int z = 2 * wv + q;
            float acc = 0.f;
            acc += psum[z][3][l];
            acc += psum[z + 1][2][l];
            acc += psum[z + 2][1][l];
            acc += psum[z + 3][0][l];
            acc += psum[z + 4][1][l];
            acc += psum[z + 5][2][l];
            acc += psum[z + 6][3][l];
            const int loc = x + 8 * y + 64 * z;
            const float V = Vz[q], Dg = Dz[q];
            const float pi = own[loc], ri = own_r[loc];
            const float stv = OP == VPS_OP_ROW ? acc : V * acc;
            const float mp = V > 0.f ? Dg * pi + stv : 0.f;
            __builtin_nontemporal_store(mp, &a.out[base + loc]);
            __builtin_nontemporal_store(pi, &a.out2[base + loc]);
            __builtin_nontemporal_store(ri, &a.out3[base + loc]);
            __builtin_nontemporal_store(Xz[q] + al * Pz[q], &a.xio[base + loc]);
            const double rvv = (double)ri;
            red0 += (double)pi * (double)mp;
            red1 += rvv * (double)mp;
            red2 += (double)mp * (double)mp;
            red3 += rvv * rvv;
        }
        const int nn = p + 2 < hi ? nn2 : -1;
        cur = nxt;
        nxt = nn;
        sb ^= 1;
        chained = has_next;
    }
    // this block's partials, one DPP tree per wave (see block_sums4)
    double rsum[4] = {red0, red1, red2, red3};
    __syncthreads();   // the last brick's plane sums have been read
    block_sums4(rsum, reinterpret_cast<double*>(&psum[0][0][0]), sred[1]);
    if (t < 4) {
        double* dst = t == 0 ? it.out[0] : (t == 1 ? it.out[1] : (t == 2 ? it.out[2] : it.out[3]));
        dst[blockIdx.x] = sred[1][t];
    }
}

// 2D: one 64-lane wave per 8x8 brick, one output per lane.
template <int KIND, int OP, typename T, typename C>
__global__ void __launch_bounds__(64) k_stencil2(StencilArgs a, C cc) {
    __shared__ T tile[kHaloIters2 * 64];
    __shared__ float own[64];
    __shared__ double own64[KIND == K_TRUE_RES ? 64 : 1];
    __shared__ int64_t sbase[10];
    float beta = 0.f;
    if constexpr (KIND == K_SPMV_FIRST || KIND == K_SPMV_FUSED) {
        if (!spmv_prologue(a, &beta)) return;
    }
    if constexpr (KIND == K_TRUE_RES) {
        if (a.spec && a.st->done != 1) return;   // same value in every lane
    }
    const int slot = xcd_remap(blockIdx.x, (int)a.nslots);
    const int brick = a.brick_of_slot[slot];
    const int t = threadIdx.x;
    if (t < 10) {
        int64_t s = a.zero_slot;
        if (t < 9) {
            const int bx = brick % a.k.bdims[0];
            const int by = brick / a.k.bdims[0];
            const int nx = bx + (t % 3) - 1, ny = by + t / 3 - 1;
            if (nx >= 0 && ny >= 0 && nx < a.k.bdims[0] && ny < a.k.bdims[1]) {
                const int sl = a.slot_of_brick[(int64_t)ny * a.k.bdims[0] + nx];
                if (sl >= 0) s = sl;
            }
        }
        sbase[t] = s * 64;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kHaloIters2; i++) {
        const int e = t + 64 * i;
        const int v = kHalo2.v[e];
        const int64_t g = sbase[v >> 9] + (v & 511);
        float pn = 0.f;
        if constexpr (KIND == K_TRUE_RES) {
            // x = FP64 base + FP32 correction
            const double xs = (double)a.in[g] + a.in_base[g];
            tile[e] = OP == VPS_OP_ROW ? (double)a.vvol[g] * xs : xs;
            if ((v >> 9) == 4) own64[v & 511] = xs;
        } else {
            tile[e] = load_input<KIND, OP, T>(a, g, beta, &pn);
        }
        if constexpr (KIND != K_DIAG && KIND != K_TRUE_RES) {
            if ((v >> 9) == 4) {
                own[v & 511] = pn;
                if constexpr (KIND == K_SPMV_FUSED) a.out2[g] = pn;
            }
        }
    }
    __syncthreads();
    const int x = t & 7, y = t >> 3;
    const T acc = plane1<T>(&tile[(y + 3) * kT + (x + 3)], cc);
    const int loc = x + 8 * y;
    const int64_t g = (int64_t)slot * 64 + loc;
    const float V = a.vvol[g];
    const bool on = V > 0.f;
    double red0 = 0, red1 = 0;
    if constexpr (KIND == K_DIAG) {
        const double s = -(double)acc;
        const double d = on ? (s < a.alpha0 ? a.alpha0 : s) : 0.0;
        a.out_d[g] = d;
        a.out_f[g] = (float)d;
        a.out_s[g] = on ? (float)s : 0.f;
    } else if constexpr (KIND == K_TRUE_RES) {
        const double xi = own64[loc];
        const double st = OP == VPS_OP_ROW ? (double)acc : (double)V * (double)acc;
        const double mx = a.vdiag64[g] * xi + st;
        const double rt = on ? (double)a.vb[g] - mx : 0.0;
        a.out[g] = (float)rt;
        red0 = rt * rt;
    } else {
        const float pi = own[loc];
        const float st = OP == VPS_OP_ROW ? (float)acc : V * (float)acc;
        const float mp = on ? a.vdiag[g] * pi + st : 0.f;
        if constexpr (KIND == K_APPLY) {
            a.out[g] = mp;
        } else if constexpr (KIND == K_CG_INIT) {
            const float b = a.vb[g];
            const float r = on ? b - mp : 0.f;
            a.out[g] = r;
            a.out2[g] = r;
            red0 = (double)b * (double)b;
            red1 = (double)r * (double)r;
        } else {
            a.out[g] = mp;
            red0 = (double)pi * (double)mp;
        }
    }
    if constexpr (KIND != K_DIAG && KIND != K_APPLY) {
        red0 = wave_sum_all(red0);
        if (KIND == K_CG_INIT) red1 = wave_sum_all(red1);
        if (t == 0) {
            a.part0[blockIdx.x] = red0;
            if (KIND == K_CG_INIT) a.part1[blockIdx.x] = red1;
        }
    }
}

// x += alpha p; r -= alpha t; partial r.r.  alpha = rr_{k-1} / (p.t)_k,
// both re-summed from the producers' partials.
__global__ void __launch_bounds__(kUpdThreads) k_cg_update(
    int64_t n4, float4* x, float4* r, const float4* p, const float4* tt,
    CgState* st, const double* pt_part, int npt, const double* rr_prev,
    int kk, double* rr_out) {
    __shared__ double sbuf[kUpdThreads / 64];
    __shared__ int s_done;
    // every load that does not depend on alpha is issued first: the state
    // flag (read once per block, so the exit below is uniform), both
    // partial sums and this thread's vector elements
    if (threadIdx.x == 0) s_done = st->done;
    const double abs0 = st->abs0;
    double pt = 0, rp = 0;
    for (int i = threadIdx.x; i < npt; i += kUpdThreads) pt += pt_part[i];
    if (kk != 0)
        for (int i = threadIdx.x; i < kUpdBlocks; i += kUpdThreads) rp += rr_prev[i];
    const int64_t i0 = (int64_t)blockIdx.x * kUpdThreads + threadIdx.x;
    const int64_t stride = (int64_t)kUpdBlocks * kUpdThreads;
    float4 xv0{}, rv0{}, pv0{}, tv0{};
    if (i0 < n4) {
        xv0 = x[i0];
        rv0 = r[i0];
        pv0 = p[i0];
        tv0 = tt[i0];
    }
    __syncthreads();
    if (s_done) return;
    pt = block_sum<kUpdThreads>(pt, sbuf);
    rp = kk == 0 ? abs0 : block_sum<kUpdThreads>(rp, sbuf);
    const double ald = rp / pt;
    if (!isfinite(ald)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) st->done = 3;
        return;
    }
    const float al = (float)ald;
    double rr = 0;
    for (int64_t i = i0; i < n4; i += stride) {
        float4 xv, rv, pv, tv;
        if (i == i0) {
            xv = xv0; rv = rv0; pv = pv0; tv = tv0;
        } else {
            xv = x[i]; rv = r[i]; pv = p[i]; tv = tt[i];
        }
        xv.x += al * pv.x;
        xv.y += al * pv.y;
        xv.z += al * pv.z;
        xv.w += al * pv.w;
        rv.x -= al * tv.x;
        rv.y -= al * tv.y;
        rv.z -= al * tv.z;
        rv.w -= al * tv.w;
        store_nt4((float*)&x[i], xv.x, xv.y, xv.z, xv.w);
        store_nt4((float*)&r[i], rv.x, rv.y, rv.z, rv.w);
        rr += ((double)rv.x * rv.x + (double)rv.y * rv.y) +
              ((double)rv.z * rv.z + (double)rv.w * rv.w);
    }
    rr = block_sum<kUpdThreads>(rr, sbuf);
    if (threadIdx.x == 0) rr_out[blockIdx.x] = rr;
}

// CG start: b.b and r0.r0 from the init partials -> tolerance, flags.
__global__ void __launch_bounds__(1024) k_cg_start(const double* pbb,
                                                   const double* prr, int n,
                                                   double tol_in, double rec_scale,
                                                   int guard, CgState* st) {
    __shared__ double sbuf[16];
    double va[4], vb[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {   // loads issued together
        const int i = threadIdx.x + 1024 * u;
        va[u] = i < n ? pbb[i] : 0.0;
        vb[u] = i < n ? prr[i] : 0.0;
    }
    double a = (va[0] + va[1]) + (va[2] + va[3]);
    double b = (vb[0] + vb[1]) + (vb[2] + vb[3]);
    for (int i = threadIdx.x + 4096; i < n; i += 1024) {
        a += pbb[i];
        b += prr[i];
    }
    const double bb = block_sum<1024>(a, sbuf);
    const double rr = block_sum<1024>(b, sbuf);
    if (threadIdx.x != 0) return;
    *st = CgState{};   // the whole state starts here (no host upload)
    st->bb = bb;
    st->rr = rr;
    st->abs0 = rr;
    st->iter = 0;
    st->true_rr = -1.0;
    st->final_kk = -1;
    // IterativeSolver.h:25 setTolerance(max(tol, eps / ||b||)); Eigen's
    // threshold max(tol^2 ||b||^2, DBL_MIN) (ConjugateGradient.h)
    const double eps_tol = DBL_EPSILON / sqrt(bb);
    const double tol = tol_in > eps_tol ? tol_in : eps_tol;
    const double thr = tol * tol * bb;
    st->thr = thr < DBL_MIN ? DBL_MIN : thr;
    st->thr_rec = st->thr * rec_scale;
    if (!isfinite(bb)) st->done = 3;
    else if (bb == 0.0) st->done = 4;   // Eigen: x = 0, zero iterations
    else if (!isfinite(rr) || (guard && !(rr <= bb))) st->done = 5;   // start rejected
    else st->done = rr < st->thr ? 1 : 0;
    st->stop = st->done ? kStopBefore : 0;
}

// 3D residual replacement: r (the buffer the next iteration launch reads) takes the
// FP64-checked true residual rt; partial r.r per block.
__global__ void __launch_bounds__(kUpdThreads) k_copy_rr(int64_t n4, const float4* rt,
                                                         float4* r, double* part) {
    __shared__ double sbuf[kUpdThreads / 64];
    double rr = 0;
    for (int64_t i = (int64_t)blockIdx.x * kUpdThreads + threadIdx.x; i < n4;
         i += (int64_t)kUpdBlocks * kUpdThreads) {
        const float4 v = rt[i];
        r[i] = v;
        rr += ((double)v.x * v.x + (double)v.y * v.y) +
              ((double)v.z * v.z + (double)v.w * v.w);
    }
    rr = block_sum<kUpdThreads>(rr, sbuf);
    if (threadIdx.x == 0) part[blockIdx.x] = rr;
}

// x is kept in two parts during a solve: an FP64 base xb and the FP32
// correction x that the iterations update.  A correction is small next to
// the solution once the start is close, so its FP32 rounding stays far
// below the tolerance even where ||b|| is small next to ||M|| ||x||.
//   kFoldSplit  xb = x0, x = 0             (start of the solve)
//   kFoldMerge  xb += x, x = 0             (each residual replacement)
//   kFoldFinal  x = (float)(xb + x)        (end of the solve: the pressures)
//   kFoldZero   x = 0, xb = 0              (b = 0: Eigen's x = 0)
enum FoldMode { kFoldSplit = 0, kFoldMerge = 1, kFoldFinal = 2, kFoldZero = 3 };
__global__ void __launch_bounds__(kUpdThreads) k_fold_x(int64_t n4, float4* x, double4* xb,
                                                        int mode, SolveGuard guard) {
    if (!solve_accepted(guard)) return;
    for (int64_t i = (int64_t)blockIdx.x * kUpdThreads + threadIdx.x; i < n4;
         i += (int64_t)kUpdBlocks * kUpdThreads) {
        const float4 d = x[i];
        if (mode == kFoldSplit) {
            xb[i] = make_double4(d.x, d.y, d.z, d.w);
            x[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        } else if (mode == kFoldMerge) {
            const double4 b = xb[i];
            xb[i] = make_double4(b.x + d.x, b.y + d.y, b.z + d.z, b.w + d.w);
            x[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        } else if (mode == kFoldFinal) {
            const double4 b = xb[i];
            x[i] = make_float4((float)(b.x + d.x), (float)(b.y + d.y), (float)(b.z + d.z),
                               (float)(b.w + d.w));
        } else {
            xb[i] = make_double4(0.0, 0.0, 0.0, 0.0);
            x[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
}

// ... and the iteration resumes from the same x: alpha = 0 (x already has
// the last update), beta = r_true.r_true / r.r before that update.
__global__ void __launch_bounds__(1024) k_resume(const double* part, CgState* st) {
    __shared__ double sbuf[16];
    double a = 0;
    for (int i = threadIdx.x; i < kUpdBlocks; i += 1024) a += part[i];
    a = block_sum<1024>(a, sbuf);
    if (threadIdx.x == 0) {
        st->alpha = 0.f;
        st->beta = (float)(a / st->rr_prev);
        st->rr = a;
        st->true_rr = -1.0;
        st->final_kk = -1;
        st->done = 0;
        st->stop = 0;
    }
}

// Deterministic one-block sum of n partials into *out.
// st != nullptr: only once the recursive residual has converged (the
// speculative check launched after every batch)
__global__ void __launch_bounds__(1024) k_sum_partials(const double* p, int n,
                                                       double* out, const CgState* st) {
    __shared__ double sbuf[16];
    if (st && st->done != 1) return;   // same value in every thread
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {   // loads issued together
        const int i = threadIdx.x + 1024 * u;
        v[u] = i < n ? p[i] : 0.0;
    }
    double a = (v[0] + v[1]) + (v[2] + v[3]);
    for (int i = threadIdx.x + 4096; i < n; i += 1024) a += p[i];
    a = block_sum<1024>(a, sbuf);
    if (threadIdx.x == 0) *out = a;
}

// The first SpMV leaves one partial per brick; the first iteration launch
// reads them folded to G: partial g of sum q is the sum, in index order, of
// bricks [g n / G, (g + 1) n / G), so that no block of that launch reads all
// n (every block of an iteration launch sums every partial it is given).
__global__ void __launch_bounds__(256) k_fold_partials(const double* p0, const double* p1,
                                                       const double* p2, const double* p3,
                                                       int n, int G, double* out0,
                                                       double* out1, double* out2,
                                                       double* out3) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 4 * G) return;
    const int q = i / G, g = i - q * G;
    const double* p = q == 0 ? p0 : (q == 1 ? p1 : (q == 2 ? p2 : p3));
    double* out = q == 0 ? out0 : (q == 1 ? out1 : (q == 2 ? out2 : out3));
    const int a = (int)((int64_t)g * n / G), b = (int)((int64_t)(g + 1) * n / G);
    double s = 0.0;
    for (int j = a; j < b; j++) s += p[j];
    out[g] = s;
}

// Residual replacement: r = rt; p = rt (restart direction); partial r.r.
__global__ void __launch_bounds__(kUpdThreads) k_replace(int64_t n4,
                                                         const float4* rt,
                                                         float4* r, float4* p0,
                                                         double* part) {
    __shared__ double sbuf[kUpdThreads / 64];
    double rr = 0;
    for (int64_t i = (int64_t)blockIdx.x * kUpdThreads + threadIdx.x; i < n4;
         i += (int64_t)kUpdBlocks * kUpdThreads) {
        const float4 v = rt[i];
        r[i] = v;
        p0[i] = v;
        rr += ((double)v.x * v.x + (double)v.y * v.y) +
              ((double)v.z * v.z + (double)v.w * v.w);
    }
    rr = block_sum<kUpdThreads>(rr, sbuf);
    if (threadIdx.x == 0) part[blockIdx.x] = rr;
}

// Residual replacement keeping the direction: r = rt; partial r.r into the
// ring slot that the next SpMV reads as the newest r.r; the solve is open
// again (block 0 clears the flag: no kernel of this launch reads it).
__global__ void __launch_bounds__(kUpdThreads) k_replace_keep(int64_t n4,
                                                              const float4* rt, float4* r,
                                                              double* part, CgState* st) {
    __shared__ double sbuf[kUpdThreads / 64];
    double rr = 0;
    for (int64_t i = (int64_t)blockIdx.x * kUpdThreads + threadIdx.x; i < n4;
         i += (int64_t)kUpdBlocks * kUpdThreads) {
        const float4 v = rt[i];
        r[i] = v;
        rr += ((double)v.x * v.x + (double)v.y * v.y) +
              ((double)v.z * v.z + (double)v.w * v.w);
    }
    rr = block_sum<kUpdThreads>(rr, sbuf);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = rr;
        if (blockIdx.x == 0) {
            st->done = 0;
            st->true_rr = -1.0;
        }
    }
}

// After a replacement: the restart's r.r becomes abs0.
__global__ void __launch_bounds__(1024) k_restart(const double* part,
                                                  CgState* st) {
    __shared__ double sbuf[16];
    double a = 0;
    for (int i = threadIdx.x; i < kUpdBlocks; i += 1024) a += part[i];
    a = block_sum<1024>(a, sbuf);
    if (threadIdx.x == 0) {
        st->abs0 = a;
        st->rr = a;
        st->final_kk = -1;
        st->done = a < st->thr ? 1 : 0;
        st->stop = st->done ? kStopBefore : 0;
    }
}

__global__ void __launch_bounds__(256) k_alpha0(int64_t n, const float* vvol,
                                                const double* diag,
                                                double* part) {
    double m = -DBL_MAX;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256)
        if (vvol[i] > 0.f) m = diag[i] > m ? diag[i] : m;
    m = wave_max(m);
    __shared__ double sw[4];
    if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sw[0];
        for (int i = 1; i < 4; i++) t = sw[i] > t ? sw[i] : t;
        part[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(64) k_max_partials(const double* p, int n,
                                                     double* out) {
    double m = -DBL_MAX;
    for (int i = threadIdx.x; i < n; i += 64) m = p[i] > m ? p[i] : m;
    m = wave_max(m);
    if (threadIdx.x == 0) *out = m;
}

inline unsigned blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

// Coefficients cc[k2] = 2 W'_H(r) / (r + 1e-6), r = sqrt(k2) * spacing,
// the cubic-spline derivative of SmoothedParticles.h:67-72 with radius H.
static void make_coef(const vps_context* c, CoefD* d, CoefF* f) {
    const double s = c->p.lattice_spacing;
    const double H = s * 4;
    const double invH = 1 / H;
    const double c0 = 0.31830988618379067154 * (1 / (H * H)) *
                      (c->dim == 2 ? (double)40 / 7 : 8 * invH);
    const double c1 = 6 * invH * c0;
    for (int k2 = 0; k2 < 16; k2++) {
        const double r = std::sqrt((double)k2) * s;
        const double x = r * invH;
        double dw;
        if (x < .5) dw = c1 * x * (3 * x - 2);
        else if (x < 1) dw = c1 * (1 - x) * (x - 1);
        else dw = 0;
        const double v = k2 == 0 ? 0.0 : 2. * dw / (r + 1e-6);
        d->c[k2] = v;
        f->c[k2] = (float)v;
    }
}

// Partials layout (doubles): [3 x nbricks] SpMV ring, [3 x 256] update
// ring, [2 x nbricks] init, [nbricks] true residual, [256] replacement,
// [2048] misc.
static double* part_pt(vps_context* c, int k) {
    return c->partials.p + (int64_t)(k % 3) * c->nbricks;
}
static double* part_rr(vps_context* c, int k) {
    return c->partials.p + 3 * c->nbricks + (int64_t)(k % 3) * kUpdBlocks;
}
static double* part_init(vps_context* c, int which) {
    return c->partials.p + 3 * c->nbricks + 3 * kUpdBlocks +
           (int64_t)which * c->nbricks;
}
static double* part_true(vps_context* c) {
    return c->partials.p + 5 * c->nbricks + 3 * kUpdBlocks;
}
static double* part_repl(vps_context* c) {
    return c->partials.p + 6 * c->nbricks + 3 * kUpdBlocks;
}
static double* part_misc(vps_context* c) {
    return c->partials.p + 6 * c->nbricks + 4 * kUpdBlocks;
}
// the second set of the 3D iteration's partials (the first: the SpMV ring)
static double* part_iter(vps_context* c, int q) {
    return c->partials.p + 6 * c->nbricks + 4 * kUpdBlocks + 2048 + (int64_t)q * kIterBlocks;
}
// the first SpMV's per-brick partials folded to one per iteration block
static double* part_fold(vps_context* c, int q) {
    return c->partials.p + 6 * c->nbricks + 4 * kUpdBlocks + 2048 +
           (int64_t)(4 + q) * kIterBlocks;
}

int64_t partials_size(int64_t nbricks) {
    return 6 * nbricks + 4 * kUpdBlocks + 2048 + 8 * kIterBlocks;
}

static StencilArgs base_args(vps_context* c) {
    StencilArgs a{};
    a.k = c->k;
    a.slot_of_brick = c->brick_slot[c->which].p;
    a.brick_of_slot = c->slot_brick[c->which].p;
    a.nbr_slot = c->nbr_slot.p;
    a.nslots = c->nslots;
    a.zero_slot = c->nbricks;
    a.vvol = c->vvol.p;
    a.vdiag = c->vdiag.p;
    a.vdiag64 = c->vdiag64.p;
    a.vb = c->vb.p;
    a.st = c->cg.p;
    a.alpha0 = c->alpha0;
    a.max_iter = 1;
    return a;
}

// With a pair of events the launch goes through hipExtLaunchKernelGGL, which
// stamps them at the start and the end of the kernel's own execution (no
// dispatch gap in the measured time).
template <int KIND, int OP, typename T, typename C>
static void launch(vps_context* c, const StencilArgs& a, const C& cc,
                   hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    const dim3 grid((unsigned)c->nslots), block(64), block3(kS3Threads);
    if (c->dim == 3) {
        if (e0)
            hipExtLaunchKernelGGL((k_stencil3<KIND, OP, T, C>), grid, block3, 0,
                                  c->stream, e0, e1, 0, a, cc);
        else
            hipLaunchKernelGGL((k_stencil3<KIND, OP, T, C>), grid, block3, 0,
                               c->stream, a, cc);
    } else {
        if (e0)
            hipExtLaunchKernelGGL((k_stencil2<KIND, OP, T, C>), grid, block, 0,
                                  c->stream, e0, e1, 0, a, cc);
        else
            hipLaunchKernelGGL((k_stencil2<KIND, OP, T, C>), grid, block, 0,
                               c->stream, a, cc);
    }
}

template <int KIND, typename T, typename C>
static void launch_op(vps_context* c, const StencilArgs& a, const C& cc,
                      hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    if (c->p.op_mode == VPS_OP_ROW) launch<KIND, VPS_OP_ROW, T, C>(c, a, cc, e0, e1);
    else launch<KIND, VPS_OP_TRANSPOSE, T, C>(c, a, cc, e0, e1);
}

int virtual_diag(vps_context* c, bool set_alpha0) {
    if (c->nslots == 0) return VPS_OK;
    CoefD cd;
    CoefF cf;
    make_coef(c, &cd, &cf);
    StencilArgs a = base_args(c);
    a.alpha0 = set_alpha0 ? -1e6 : c->alpha0;
    a.out_d = c->vdiag64.p;
    a.out_f = c->vdiag.p;
    a.out_s = c->vrowsum.p;
    launch<K_DIAG, 0, double, CoefD>(c, a, cd);
    VPS_HIP(c, hipGetLastError());
    if (set_alpha0) {
        const int64_t nn = c->nslots * c->brick_nodes;
        int64_t nb = blocks(nn, 256);
        if (nb > 1024) nb = 1024;
        double* part = part_misc(c);
        double* dout = (double*)(c->counters.p + 4);
        hipLaunchKernelGGL(k_alpha0, dim3((unsigned)nb), dim3(256), 0,
                           c->stream, nn, c->vvol.p, c->vdiag64.p, part);
        hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(64), 0, c->stream,
                           part, (int)nb, dout);
        VPS_HIP(c, hipMemcpyAsync(c->host_counters + 4, dout, 8,
                                  hipMemcpyDeviceToHost, c->stream));
        VPS_HIP(c, hipStreamSynchronize(c->stream));
        c->alpha0 = *(double*)(c->host_counters + 4);
    }
    return VPS_OK;
}

int apply_operator_f32(vps_context* c, const float* in, float* out,
                       int op_mode) {
    (void)op_mode;
    if (c->nslots == 0) return VPS_OK;
    CoefD cd;
    CoefF cf;
    make_coef(c, &cd, &cf);
    StencilArgs a = base_args(c);
    a.in = in;
    a.out = out;
    launch_op<K_APPLY, float, CoefF>(c, a, cf);
    VPS_HIP(c, hipGetLastError());
    return VPS_OK;
}

static int read_state(vps_context* c, CgState* out) {
    VPS_HIP(c, hipMemcpyAsync(c->host_cg, c->cg.p, sizeof(CgState),
                              hipMemcpyDeviceToHost, c->stream));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    *out = *c->host_cg;
    return VPS_OK;
}

static double ms_between(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();   // not a failure of the solve
        return 0.0;
    }
    return ms;
}

// Persistent blocks of the 3D iteration: all resident at once (blocks per CU
// from the occupancy of k_cg_iter3, at most three, times the CUs, at most
// kIterBlocks), so that no block waits for another to finish.
static int iter_blocks(vps_context* c) {
    static int cached[64][2] = {};
    const int dev = c->device >= 0 && c->device < 64 ? c->device : 0;
    const int op = c->p.op_mode == VPS_OP_ROW ? 1 : 0;
    if (cached[dev][op] > 0) return cached[dev][op];
    int per_cu = 0, cus = 0;
    const auto kern = c->p.op_mode == VPS_OP_ROW ? (const void*)k_cg_iter3<VPS_OP_ROW>
                                                 : (const void*)k_cg_iter3<VPS_OP_TRANSPOSE>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kS3Threads, 0) !=
            hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) !=
            hipSuccess ||
        per_cu <= 0 || cus <= 0) {
        (void)hipGetLastError();
        return kIterBlocks;
    }
    // three blocks per CU: measured faster than four (1M: 21.2 against
    // 21.7 us per launch, 8.1M: 120 against 125 us), with longer z-runs per
    // block and less contention for LDS and the caches
    const int n = std::max(8, std::min(kIterBlocks, std::min(per_cu, 3) * cus) & ~7);
    cached[dev][op] = n;
    return n;
}

// residual replacements of one solve before it fails (the FP64 base of x
// makes the true residual reachable: one or two are the rule)
constexpr int kMaxReplacements = 8;

int cg_solve(vps_context* c, vps_stats* stats) {
    const int w = c->which;
    const int64_t nn = c->nslots * c->brick_nodes;
    const int64_t n4 = nn / 4;
    if (c->nslots == 0 || c->nvirt == 0) {
        if (stats) {
            stats->cg_iterations = 0;
            stats->cg_rel_residual = 0;
        }
        return VPS_OK;
    }
    CoefD cd;
    CoefF cf;
    make_coef(c, &cd, &cf);
    const double tol_in = c->p.tolerance > 0 ? c->p.tolerance : 1e-6;
    const int64_t max_iter =
        c->p.max_iterations >= 0 ? c->p.max_iterations : 2 * c->nvirt;
    float* x = c->vx[w].p;
    float* r = c->vr.p;
    float* t = c->vt.p;
    float* p0 = c->vp[0].p;
    float* p1 = c->vp[1].p;
    const int nslots = (int)c->nslots;
    // x = xb (FP64) + x (FP32 correction) during the solve; every return
    // after the split leaves the pressures in x
    double* xb = c->vxb.p;
    auto fold = [&](int mode, SolveGuard g = SolveGuard{}) {
        hipLaunchKernelGGL(k_fold_x, dim3(kUpdBlocks), dim3(kUpdThreads), 0, c->stream, n4,
                           (float4*)x, (double4*)xb, mode, g);
    };
    // the post-solve work behind a batch, run only if the batch ends the
    // solve accepted (the state is read back before it, so the host learns
    // the outcome while the gradient runs)
    const bool spec = c->spec_post;
    c->spec_done = false;
    auto queue_post = [&]() -> int {
        const SolveGuard g{c->cg.p, c->check_true ? 2 : 1};
        fold(kFoldFinal, g);
        VPS_TRY(after_solve(c, g));
        VPS_HIP(c, hipEventRecord(c->ev[6], c->stream));
        VPS_TRY(pressure_gradient(c, g));
        return VPS_OK;
    };
    auto finish = [&](int rc) {
        fold(kFoldFinal);
        return rc;
    };

    {
        StencilArgs a = base_args(c);
        a.in = x;
        a.out = r;
        a.out2 = p0;
        a.part0 = part_init(c, 0);
        a.part1 = part_init(c, 1);
        launch_op<K_CG_INIT, float, CoefF>(c, a, cf);
        fold(kFoldSplit);
        // 3D: the recursion's threshold absorbs the FP32 drift measured at
        // the last solve (with a 10% margin); the FP64 check keeps tol
        double rec_scale = 1.0;
        if (c->dim == 3 && c->cg_drift > 0.9) rec_scale = std::max(0.25, 0.9 / c->cg_drift);
        hipLaunchKernelGGL(k_cg_start, dim3(1), dim3(1024), 0, c->stream,
                           part_init(c, 0), part_init(c, 1), nslots, tol_in, rec_scale,
                           c->start_guard ? 1 : 0, c->cg.p);
        VPS_HIP(c, hipGetLastError());
    }
    // The start state is read back together with the first batch: its
    // kernels return at once when k_cg_start has already ended the solve
    // (converged, zero or non-finite right-hand side).  Only a zero
    // iteration limit needs the state before launching.
    CgState hs{};
    if (max_iter == 0) {
        VPS_TRY(read_state(c, &hs));
        if (hs.done == 5) {
            c->err = "initial guess rejected (non-finite or worse than zero)";
            return finish(kStartRejected);
        }
        if (hs.done == 0) {
            c->err = "CG reached the iteration limit without converging";
            return finish(VPS_ERR_NO_CONVERGENCE);
        }
    }
    bool start_checked = max_iter == 0;
    const bool prof = c->kernel_timing;
    const int64_t max_timed = (int64_t)c->kev.size() / 4;
    int64_t timed = 0, spmv_launches = 0;
    int kk = 0;          // iteration since the (re)start
    int iter_base = 0;   // Eigen iterations before the (re)start
    int replacements = 0;
    // the first batch: the last solve's count and three more (a launch past
    // convergence returns at its prologue in a few us; a batch that stops
    // short costs a host round trip of about 30 us)
    int predicted = c->prev_iters > 0 ? c->prev_iters + 3 : 8;
    // 3D: ping-pong vectors of the one-launch iteration; resume_pending: the
    // next launch is an update after a replacement (no reduction before it)
    const bool three = c->dim == 3;
    float* pb[2] = {p0, p1};
    float* rbuf[2] = {r, c->vr2.p};
    float* tbuf[2] = {t, c->vt2.p};
    bool resume_pending = false;
    // 3D: persistent blocks of the iteration and its two partial sets
    const int giter = nslots >= 8 ? (std::min(nslots, iter_blocks(c)) & ~7) : nslots;
    double* setA[4] = {part_pt(c, 0), part_pt(c, 1), part_pt(c, 2), part_init(c, 1)};
    double* setB[4] = {part_iter(c, 0), part_iter(c, 1), part_iter(c, 2), part_iter(c, 3)};
    std::vector<int> timed_iter;
    double host_launch_us = 0;
    bool post_queued = false;   // the guarded post-solve work follows this state
    while (true) {
        if (hs.done == 0) {
            int batch = iter_base + kk < predicted ? predicted - (iter_base + kk) : 4;
            if (batch > 64) batch = 64;
            const auto h0 = std::chrono::steady_clock::now();
            for (int j = 0; j < batch && three; j++) {
                // 3D: FIRST, then per iteration one launch of persistent
                // blocks: the sums of the previous SpMV, the update, the SpMV
                StencilArgs s = base_args(c);
                s.kk = kk;
                s.iter_base = iter_base;
                s.max_iter = max_iter;
                s.part0 = part_pt(c, 0);
                s.part1 = part_pt(c, 1);
                s.part2 = part_pt(c, 2);
                s.part3 = part_init(c, 1);   // free once k_cg_start has run
                if (kk == 0) {
                    s.in = p0;
                    s.out = t;
                    launch_op<K_SPMV_FIRST, float, CoefF>(c, s, cf);
                    hipLaunchKernelGGL(k_fold_partials, dim3((unsigned)((4 * giter + 255) / 256)),
                                       dim3(256), 0, c->stream, setA[0], setA[1], setA[2],
                                       setA[3], nslots, giter, part_fold(c, 0), part_fold(c, 1),
                                       part_fold(c, 2), part_fold(c, 3));
                    spmv_launches++;
                    kk++;
                    continue;
                }
                // (an update resumed after a replacement: not timed)
                const bool do_time = prof && timed < max_timed && !resume_pending;
                hipEvent_t* ev = do_time ? &c->kev[4 * timed] : nullptr;
                const int o = (kk - 1) & 1, nw = kk & 1;
                s.in = pb[o];
                s.in2 = rbuf[o];
                s.in3 = tbuf[o];
                s.out = tbuf[nw];
                s.out2 = pb[nw];
                s.out3 = rbuf[nw];
                s.xio = x;
                // partials: FIRST writes set A (one per slot), folded to one
                // per block for launch 1; launch kk writes set B when kk is
                // odd, A when even, and reads the other
                IterArgs ia;
                const int32_t* corder = c->col_order.p;
                for (int q = 0; q < 4; q++) {
                    ia.out[q] = nw ? setB[q] : setA[q];
                    ia.prev[q] = kk == 1 ? part_fold(c, q) : (nw ? setA[q] : setB[q]);
                }
                ia.nprev = giter;
                ia.resume = resume_pending ? 1 : 0;
                resume_pending = false;
                const dim3 gi((unsigned)giter), bi(kS3Threads);
                if (c->p.op_mode == VPS_OP_ROW) {
                    if (ev)
                        hipExtLaunchKernelGGL(k_cg_iter3<VPS_OP_ROW>, gi, bi, 0, c->stream, ev[0],
                                              ev[1], 0, s, cf, ia, corder);
                    else
                        hipLaunchKernelGGL(k_cg_iter3<VPS_OP_ROW>, gi, bi, 0, c->stream, s, cf, ia,
                                           corder);
                } else {
                    if (ev)
                        hipExtLaunchKernelGGL(k_cg_iter3<VPS_OP_TRANSPOSE>, gi, bi, 0, c->stream,
                                              ev[0], ev[1], 0, s, cf, ia, corder);
                    else
                        hipLaunchKernelGGL(k_cg_iter3<VPS_OP_TRANSPOSE>, gi, bi, 0, c->stream, s,
                                           cf, ia, corder);
                }
                if (do_time) {
                    // an iteration of index kk - 1 (its reduction and update)
                    timed_iter.push_back(iter_base + kk - 1);
                    timed++;
                }
                spmv_launches++;
                kk++;
            }
            for (int j = 0; j < batch && !three; j++) {
                StencilArgs s = base_args(c);
                s.kk = kk;
                s.iter_base = iter_base;
                s.max_iter = max_iter;
                s.rr1 = part_rr(c, kk + 2);   // (kk - 1) mod 3
                s.rr2 = part_rr(c, kk + 1);   // (kk - 2) mod 3
                s.part0 = part_pt(c, kk);
                s.out = t;
                float* dir = (kk == 0) ? p0 : ((kk & 1) ? p1 : p0);
                // the first SpMV of a (re)start is not fused: only fused
                // launches are timed
                const bool do_time = prof && timed < max_timed && kk > 0;
                hipEvent_t* ev = do_time ? &c->kev[4 * timed] : nullptr;
                if (kk == 0) {
                    s.in = p0;
                    launch_op<K_SPMV_FIRST, float, CoefF>(c, s, cf);
                } else {
                    s.in = (kk & 1) ? p0 : p1;   // previous direction
                    s.in2 = r;
                    s.out2 = dir;
                    launch_op<K_SPMV_FUSED, float, CoefF>(c, s, cf,
                                                          ev ? ev[0] : nullptr,
                                                          ev ? ev[1] : nullptr);
                }
                if (ev)
                    hipExtLaunchKernelGGL(k_cg_update, dim3(kUpdBlocks), dim3(kUpdThreads),
                                          0, c->stream, ev[2], ev[3], 0, n4, (float4*)x,
                                          (float4*)r, (const float4*)dir,
                                          (const float4*)t, c->cg.p, part_pt(c, kk),
                                          nslots, part_rr(c, kk + 2), kk, part_rr(c, kk));
                else
                    hipLaunchKernelGGL(k_cg_update, dim3(kUpdBlocks), dim3(kUpdThreads),
                                       0, c->stream, n4, (float4*)x, (float4*)r,
                                       (const float4*)dir, (const float4*)t, c->cg.p,
                                       part_pt(c, kk), nslots, part_rr(c, kk + 2), kk,
                                       part_rr(c, kk));
                if (do_time) {
                    timed_iter.push_back(iter_base + kk);
                    timed++;
                }
                spmv_launches++;
                kk++;
            }
            if (c->check_true) {
                // the true-residual check rides on the same read-back: it
                // runs on the device only if this batch converged
                StencilArgs tr = base_args(c);
                tr.in = x;
                tr.in_base = xb;
                tr.out = three ? c->vrt.p : t;   // r_true
                tr.part0 = part_true(c);
                tr.spec = 1;
                launch_op<K_TRUE_RES, double, CoefD>(c, tr, cd);
                hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, c->stream,
                                   part_true(c), nslots, &c->cg.p->true_rr, c->cg.p);
            }
            VPS_HIP(c, hipGetLastError());
            host_launch_us += std::chrono::duration<double, std::micro>(
                                  std::chrono::steady_clock::now() - h0).count();
            // the state read-back, then (spec) the guarded post-solve work
            // behind it; the host waits for the read-back only
            VPS_HIP(c, hipMemcpyAsync(c->host_cg, c->cg.p, sizeof(CgState),
                                      hipMemcpyDeviceToHost, c->stream));
            VPS_HIP(c, hipEventRecord(c->ev[8], c->stream));
            if (spec) VPS_TRY(queue_post());
            VPS_HIP(c, hipEventSynchronize(c->ev[8]));
            hs = *c->host_cg;
            post_queued = spec;
            if (!start_checked) {
                start_checked = true;
                if (hs.done == 3) {
                    c->err = "non-finite right-hand side";
                    return finish(VPS_ERR_NONFINITE);
                }
                if (hs.done == 5) {
                    c->err = "initial guess rejected (non-finite or worse than zero)";
                    return finish(kStartRejected);
                }
                if (hs.done == 4) {   // b = 0: x = 0, zero iterations
                    fold(kFoldZero);
                    hs.done = 1;
                    hs.iter = 0;
                    hs.rr = 0;
                    hs.true_rr = 0;
                    break;
                }
            }
            if (hs.done == 0) {
                if (iter_base + kk > max_iter + 8) {
                    c->err = "CG loop did not terminate";
                    return finish(VPS_ERR_NO_CONVERGENCE);
                }
                // size the next batch from the mean residual decay so far
                const int done_it = hs.iter - iter_base;
                if (done_it > 0 && hs.abs0 > 0 && hs.rr > 0 && hs.rr < hs.abs0 &&
                    hs.thr > 0) {
                    const double rate = std::log(hs.rr / hs.abs0) / done_it;
                    const double left = std::log(hs.thr / hs.rr) / rate;
                    int more = left > 0 ? (int)std::ceil(1.1 * left) : 1;
                    if (more < 1) more = 1;
                    if (more > 64) more = 64;
                    predicted = iter_base + kk + more;
                }
                continue;
            }
        }
        if (hs.done == 3) {
            c->err = "non-finite value in CG";
            // from a guarded start: solve again from the plain one
            return finish(c->start_guard ? kStartRejected : VPS_ERR_NONFINITE);
        }
        if (hs.done == 2) break;
        // converged on the recursive residual; the true residual of this
        // state was computed right after the batch
        if (!c->check_true) {
            c->spec_done = post_queued;
            break;
        }
        if (three && replacements == 0 && hs.rr > 0 && hs.true_rr >= 0 &&
            std::isfinite(hs.true_rr / hs.rr))
            c->cg_drift = hs.true_rr / hs.rr;
        if (hs.true_rr < hs.thr) {
            c->spec_done = post_queued;
            break;
        }
        if (replacements >= kMaxReplacements) {
            c->err = "CG: the true residual stayed above the tolerance after " +
                     std::to_string(kMaxReplacements) + " residual replacements";
            return finish(VPS_ERR_NO_CONVERGENCE);
        }
        replacements++;
        // the correction so far goes into the FP64 base; the iteration goes
        // on with a fresh FP32 correction from the true residual
        fold(kFoldMerge);
        if (three && hs.iter > iter_base) {
            // residual replacement (van der Vorst and Ye): r takes the true
            // residual, the direction is kept and the iteration continues
            // from the same x.  The launch that applied the last update
            // (kk) runs again with alpha = 0 and beta from the new r.r.
            kk = hs.iter - iter_base + 1;
            hipLaunchKernelGGL(k_copy_rr, dim3(kUpdBlocks), dim3(kUpdThreads), 0, c->stream,
                               n4, (const float4*)c->vrt.p, (float4*)rbuf[(kk - 1) & 1],
                               part_repl(c));
            hipLaunchKernelGGL(k_resume, dim3(1), dim3(1024), 0, c->stream, part_repl(c),
                               c->cg.p);
            VPS_HIP(c, hipGetLastError());
            resume_pending = true;
            hs.done = 0;
            predicted = iter_base + kk + 2;
            continue;
        }
        if (!three && hs.iter > iter_base) {
            // residual replacement (van der Vorst and Ye): r takes the true
            // residual, the direction is kept and the iteration continues
            // where convergence was declared.  The SpMV that declared it is
            // launched again: its prologue forms beta from the new r.r,
            // written into the ring slot of the previous update.
            kk = hs.iter - iter_base + 1;
            hipLaunchKernelGGL(k_replace_keep, dim3(kUpdBlocks), dim3(kUpdThreads), 0,
                               c->stream, n4, (const float4*)t, (float4*)r,
                               part_rr(c, kk + 2), c->cg.p);
            VPS_HIP(c, hipGetLastError());
            hs.done = 0;   // k_replace_keep reopened the solve on the device
            predicted = iter_base + kk + 2;
            continue;
        }
        // converged before any iteration of this (re)start: restart from the
        // true residual
        iter_base = hs.iter;
        kk = 0;
        hipLaunchKernelGGL(k_replace, dim3(kUpdBlocks), dim3(kUpdThreads), 0,
                           c->stream, n4, (const float4*)(three ? c->vrt.p : t), (float4*)r,
                           (float4*)p0, part_repl(c));
        hipLaunchKernelGGL(k_restart, dim3(1), dim3(1024), 0, c->stream,
                           part_repl(c), c->cg.p);
        VPS_HIP(c, hipGetLastError());
        VPS_TRY(read_state(c, &hs));
        post_queued = false;
        predicted = iter_base + 4;
    }
    if (!c->spec_done) fold(kFoldFinal);
    VPS_HIP(c, hipGetLastError());
    c->prev_iters = hs.iter;
    double spmv_ms = 0, upd_ms = 0;
    int64_t counted = 0;
    for (int64_t j = 0; j < timed; j++) {
        if (timed_iter[(size_t)j] > hs.iter) break;   // no-op launches
        spmv_ms += ms_between(c->kev[4 * j], c->kev[4 * j + 1]);
        // 3D: the update is part of the iteration launch
        if (!three) upd_ms += ms_between(c->kev[4 * j + 2], c->kev[4 * j + 3]);
        counted++;
    }
    if (stats) {
        stats->cg_iterations = hs.iter;
        // b = 0 solves exactly with x = 0 (relative residual 0)
        stats->cg_rel_residual = hs.bb > 0 ? std::sqrt(hs.rr / hs.bb) : 0.0;
        stats->cg_true_rel_residual =
            hs.true_rr >= 0 ? (hs.bb > 0 ? std::sqrt(hs.true_rr / hs.bb) : 0.0) : -1.0;
        stats->rhs_norm = std::sqrt(hs.bb);
        stats->spmv_launches = spmv_launches;
        stats->ms_spmv_avg = counted ? spmv_ms / counted : 0;
        stats->ms_update_avg = counted ? upd_ms / counted : 0;
        stats->reserved[0] = replacements;
        stats->reserved[1] = (int64_t)host_launch_us;   // host time spent launching
    }
    if (hs.done == 2) {
        c->err = "CG reached the iteration limit without converging";
        return VPS_ERR_NO_CONVERGENCE;
    }
    return VPS_OK;
}

}  // namespace vps
