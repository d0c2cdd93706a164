// vps_internal.h -- shared definitions of the MI355X projection library.
//
// Data layout in HBM (one context, one device):
//   real particles   sorted by cell every step; SoA-of-vectors:
//                    rpos  double4 (x, y, z, 0)      exact FP64 membership
//                    rvel  float4  (vx, vy, vz, V)   V frozen at initialize
//                    rrho  float                      current density
//                    rid   int32                      caller index
//   boundary         same layout, sorted once (bpos, bvel = (v, V),
//                    bnrm = (n, rho))
//   cell grid        uniform, cell size h (1 + 1e-12), x-fastest linear id,
//                    cell_start[C + 1]; the 3 cells of an x-row are
//                    contiguous in the sorted arrays, so a query walks
//                    3^(D-1) contiguous ranges
//   virtual lattice  bricks of 8^D nodes; brick_slot[brick] -> slot or -1;
//                    every per-node field is stored slot-major
//                    (slot * 8^D + local, local x fastest) so a brick is one
//                    contiguous 2 KB (3D) run per field
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/vps.h"

namespace vps {

constexpr int kBrick = 8;                  // brick side in lattice nodes
constexpr int kBrickNodes3 = 512;
constexpr int kBrickNodes2 = 64;
constexpr int kStencilR = 3;               // |offset| < 4 lattice spacings
constexpr int kMaxReduceBlocks = 65536;

// Constants every kernel needs, passed by value.
struct KConst {
    int dim;
    // kernel radii: h for real/boundary (4 * particle radius), H for virtual
    // (4 * lattice spacing), squared in FP64 exactly as the reference does
    double h, h2, H, H2;
    double inv_h64, c0h64;                 // FP64 weight for the 1e-6 test
    float hf, inv_h, c0h, c1h;             // FP32 kernel constants, radius h
    float Hf, inv_H, c0H, c1H;             // FP32 kernel constants, radius H
    // lattice
    double origin[3];
    double spacing;
    int ndims[3];
    int bdims[3];
    // cell grid
    double cell_lo[3];
    double cell_inv;
    double cell_size;                      // 1 / cell_inv (cell corners)
    int cdims[3];
    // physics
    float mass;
    float rho0;
    float kappa;
    // FP32 pre-test of |x_j - q|^2 < h^2 on brick-relative coordinates:
    // below h2f_lo inside, at or above h2f_hi outside, in between the exact
    // FP64 test decides (the band covers the FP32 rounding of the relative
    // coordinates and of the squared distance)
    float h2f_lo;
    float h2f_hi;
    float pad[3];
};

// ---- FP64 helpers with the reference's rounding ----------------------
// Node position origin + c * spacing (Grid::dataPosition, Grid.h:51): one
// product and one sum, each rounded.
__device__ __forceinline__ double node_coord(double o, int c, double s) {
#pragma clang fp contract(off)
    return o + (double)c * s;
}

// (a - b).squaredNorm() as ((dx*dx + dy*dy) + dz*dz), no contraction.
__device__ __forceinline__ double sq_dist3(double ax, double ay, double az,
                                           double bx, double by, double bz) {
#pragma clang fp contract(off)
    const double dx = ax - bx, dy = ay - by, dz = az - bz;
    return (dx * dx + dy * dy) + dz * dz;
}
__device__ __forceinline__ double sq_dist2(double ax, double ay, double bx,
                                           double by) {
#pragma clang fp contract(off)
    const double dx = ax - bx, dy = ay - by;
    return dx * dx + dy * dy;
}
template <int D>
__device__ __forceinline__ double sq_dist(const double* a, const double* b) {
    if constexpr (D == 2) return sq_dist2(a[0], a[1], b[0], b[1]);
    else return sq_dist3(a[0], a[1], a[2], b[0], b[1], b[2]);
}
__device__ __forceinline__ double dsub(double a, double b) {
#pragma clang fp contract(off)
    return a - b;
}
__device__ __forceinline__ double dsq(double a) {
#pragma clang fp contract(off)
    return a * a;
}
__device__ __forceinline__ double dadd(double a, double b) {
#pragma clang fp contract(off)
    return a + b;
}

// FP64 cubic spline W (SmoothedParticles.h:60-65), for the virtual-set
// threshold test only.
__device__ __forceinline__ double kernel_w64(double d2, double inv_kr,
                                             double c0) {
#pragma clang fp contract(off)
    const double x = sqrt(d2) * inv_kr;
    if (x < .5) return c0 * (6 * x * x * (x - 1) + 1);
    else if (x < 1) return c0 * 2 * (1 - x) * (1 - x) * (1 - x);
    return 0;
}

// FP32 cubic spline value and first derivative at squared distance d2 < kr2.
// For x >= 1/2 the factor (1 - x) is formed as (kr^2 - d^2) / (kr (kr + r))
// from the exact FP64 difference kr2 - d2, so values near the support edge
// keep full FP32 relative accuracy.
struct KW {
    float w, dw, r;
};
__device__ __forceinline__ KW kernel_f32(double d2, double kr2, float kr,
                                         float inv_kr, float c0, float c1) {
    KW o;
    const float r = sqrtf((float)d2);
    const float x = r * inv_kr;
    o.r = r;
    if (!(d2 < kr2)) {
        o.w = 0.f;
        o.dw = 0.f;
    } else if (x < 0.5f) {
        o.w = c0 * (6.f * x * x * (x - 1.f) + 1.f);
        o.dw = c1 * x * (3.f * x - 2.f);
    } else {
        const float omx = (float)(kr2 - d2) / (kr * (kr + r));
        o.w = 2.f * c0 * omx * omx * omx;
        o.dw = -c1 * omx * omx;
    }
    return o;
}

// FP32 cubic spline from x = r / kr (valid for 0 <= x): W and
// S = -W'(r) / r, the factor of gradK(d) = S d, without divisions
// (SmoothedParticles.h:60-72: W' = c1 x (3x - 2) below x = 1/2, so
// -W'/r = c1 (2 - 3x) / kr there; -c1 (1 - x)^2 above, so -W'/r =
// c1 (1 - x)^2 / r).  rinv = 1 / r is only used for x >= 1/2.
// Both pieces are evaluated and selected (no branch: the lanes of a wave
// mix both sides); max(1 - x, 0) makes the outer piece vanish for x >= 1.
__device__ __forceinline__ float spline_w(float x, float c0) {
    const float in = c0 * (6.f * x * x * (x - 1.f) + 1.f);
    const float omx = fmaxf(1.f - x, 0.f);
    const float out = 2.f * c0 * omx * omx * omx;
    return x < 0.5f ? in : out;
}
__device__ __forceinline__ float spline_s(float x, float rinv, float c1,
                                          float inv_kr) {
    const float in = c1 * inv_kr * (2.f - 3.f * x);
    const float omx = fmaxf(1.f - x, 0.f);
    const float out = c1 * omx * omx * rinv;
    return x < 0.5f ? in : out;
}

// The same spline with no select: from a = max(1 - x, 0) and
// b = max(1 - 2x, 0), W = c0 (2 a^3 - b^3) and S = c1 (a^2 - b^2) / r on
// every piece (below x = 1/2, 2 a^3 - b^3 = 6 x^2 (x - 1) + 1 and
// a^2 - b^2 = x (2 - 3x); above, b = 0; from x = 1 on, a = b = 0).  Both
// values are <= 1, so each max is the clamp to [0, 1], which folds into the
// fma that forms it.  The differences lose at most one ulp of 1 to
// cancellation, below the rounding of the sums they enter.
struct SplineAB {
    float a, b;
};
__device__ __forceinline__ SplineAB spline_ab(float r, float inv_kr) {
    return {__builtin_amdgcn_fmed3f(fmaf(-inv_kr, r, 1.f), 0.f, 1.f),
            __builtin_amdgcn_fmed3f(fmaf(-2.f * inv_kr, r, 1.f), 0.f, 1.f)};
}
__device__ __forceinline__ float spline_w_ab(SplineAB s, float c0) {
    return c0 * fmaf(s.a + s.a, s.a * s.a, -(s.b * s.b * s.b));
}
__device__ __forceinline__ float spline_s_ab(SplineAB s, float rinv, float c1) {
    return c1 * rinv * fmaf(s.a, s.a, -(s.b * s.b));
}

// Candidate masks of an FP32 pre-test on cell-relative offsets (w < 0 marks
// a particle clamped into a border cell): `fast` inside h by the FP32 band,
// `exact` where the FP64 test decides (the band around h^2, r = 0, clamped
// particles); neither: outside h.
struct Pretest {
    bool fast, exact;
};
__device__ __forceinline__ Pretest pretest(const KConst& k, float d2f, float w) {
    const bool ok = w >= 0.f;
    const bool fast = ok & (d2f > 0.f) & (d2f < k.h2f_lo);
    const bool exact = !fast & (!ok | (d2f < k.h2f_hi));
    return Pretest{fast, exact};
}

// ---- launch geometry ----------------------------------------------------
// Workgroups are dealt round-robin to the 8 XCDs of an MI355X.  Remapping
// the hardware block id so that each XCD gets a contiguous range of logical
// blocks keeps spatial neighbours (which share halo data) in one L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ int cell_coord(double p, double lo, double inv,
                                          int n) {
    int c = (int)floor((p - lo) * inv);
    return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

// Corner C_c = lo + c * size of cell c along one axis, in FP64.  Any fixed
// value per cell serves: staged gathers form q - x_j as
// (q - C_q) - (x_j - C_j) + (C_q - C_j).
__device__ __forceinline__ double cell_corner(const KConst& k, int a, int c) {
    return k.cell_lo[a] + (double)c * k.cell_size;
}

// Walks the candidates of a radius-h query: the 3^(D-1) x-rows of 3 cells
// around the query cell, each one contiguous range of the cell-sorted set.
// Positions are loaded 4 at a time so their latencies overlap; f(j, x_j)
// is called for every candidate in order.
template <int D, typename F>
__device__ __forceinline__ void for_each_candidate(const KConst& k,
                                                   const int32_t* cstart,
                                                   const double4* pos,
                                                   double qx, double qy,
                                                   double qz, F&& f) {
    constexpr int U = 4;
    const int cx = cell_coord(qx, k.cell_lo[0], k.cell_inv, k.cdims[0]);
    const int cy = cell_coord(qy, k.cell_lo[1], k.cell_inv, k.cdims[1]);
    const int cz = D == 3 ? cell_coord(qz, k.cell_lo[2], k.cell_inv, k.cdims[2]) : 0;
    const int x0 = cx > 0 ? cx - 1 : 0;
    const int x1 = cx + 1 < k.cdims[0] ? cx + 1 : k.cdims[0] - 1;
    const int zlo = D == 3 ? (cz > 0 ? cz - 1 : 0) : 0;
    const int zhi = D == 3 ? (cz + 1 < k.cdims[2] ? cz + 1 : k.cdims[2] - 1) : 0;
    const int ylo = cy > 0 ? cy - 1 : 0;
    const int yhi = cy + 1 < k.cdims[1] ? cy + 1 : k.cdims[1] - 1;
    for (int z = zlo; z <= zhi; z++)
        for (int y = ylo; y <= yhi; y++) {
            const int row = (z * k.cdims[1] + y) * k.cdims[0];
            const int b = cstart[row + x0], e = cstart[row + x1 + 1];
            for (int j = b; j < e; j += U) {
                double4 pj[U];
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (j + u < e) pj[u] = pos[j + u];
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (j + u < e) f(j + u, pj[u]);
            }
        }
}

// ---- reductions ----------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// FP64 wave sum without the LDS crossbar, for call sites where all 64 lanes
// are active: DPP butterflies within each row of 16 lanes (xor 1, xor 2,
// half-row mirror, row mirror leave the row sum in every lane of the row),
// then the four row sums read as scalars and added in a fixed order.  Every
// lane returns the same value.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)u, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}
__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const int hi = __builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}
__device__ __forceinline__ double wave_sum_all(double v) {
    v += dpp_f64<0xb1>(v);    // quad_perm [1, 0, 3, 2]
    v += dpp_f64<0x4e>(v);    // quad_perm [2, 3, 0, 1]
    v += dpp_f64<0x141>(v);   // row_half_mirror
    v += dpp_f64<0x140>(v);   // row_mirror
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) +
           (readlane_f64(v, 32) + readlane_f64(v, 48));
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T u = __shfl_xor(v, o, 64);
        v = u > v ? u : v;
    }
    return v;
}

// Deterministic grid-wide sum: every block writes its partial, the last block
// to arrive sums the partials in index order.  Returns true in the block that
// produced the total (written to *out by thread 0).
template <int BS>
__device__ bool grid_sum_last(double v, double* partials, unsigned* counter,
                              double* out) {
    __shared__ double sred[BS / 64];
    __shared__ bool is_last;
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sred[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int i = 0; i < BS / 64; i++) s += sred[i];
        partials[blockIdx.x] = s;
        __threadfence();
        const unsigned t = atomicAdd(counter, 1u);
        is_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    double s = 0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += BS)
        s += __hip_atomic_load(&partials[i], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    s = wave_sum(s);
    __syncthreads();
    if (lane == 0) sred[wv] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int i = 0; i < BS / 64; i++) t += sred[i];
        *out = t;
        *counter = 0;
    }
    return true;
}

// ---- device buffers -------------------------------------------------------
template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
};

// CG scalars living on the device (no host round trip per iteration).
struct CgState {
    double bb;        // b.b
    double thr;       // tol^2 b.b (or DBL_MIN)
    double abs_new;   // r.r of the current residual
    double rr;        // latest r.r
    double pt;        // p.t
    double true_rr;   // ||b - M x||^2 from the FP64 check
    double thr_scale; // tol^2
    double abs0;      // r.r at the (re)start of the iteration
    float alpha, beta;
    int iter;         // Eigen iteration counter
    int done;         // 0 running, 1 converged, 2 max iterations, 3 nonfinite,
                      // 4 b = 0, 5 guarded start rejected (r0.r0 not <= b.b)
    int first;        // 1 before the first SpMV of a (re)start
    int pad;
    unsigned counter[4];
    double rr_prev;   // 3D iteration: r.r before the last update
    double thr_rec;   // 3D iteration: threshold of the recursive r.r
    int final_kk;     // 3D iteration: the launch that applies the last x update
    // 3D iteration launches: 0 while running; kk + 1 of the launch that
    // ended the solve; kStopBefore when it ended before any launch (read by
    // the iteration launches only: one word, so a block never mixes a stop
    // declared in its own launch with one from an earlier launch)
    int stop;
};
constexpr int kStopBefore = 0x7fffffff;

// Work after a solve that is queued before the host has read the CG state:
// it runs only if the solve ended accepted.  mode 0: always (the host has
// decided); 1: converged (done == 1); 2: converged with the FP64 true
// residual below the threshold (the host's test when the check is on).
struct SolveGuard {
    const CgState* st = nullptr;
    int mode = 0;
};
__device__ __forceinline__ bool solve_accepted(const SolveGuard& g) {
    if (g.mode == 0) return true;
    const CgState* s = g.st;
    return s->done == 1 && (g.mode == 1 || s->true_rr < s->thr);
}
// cg_solve's status when a guarded warm start (learned or extrapolated) is
// rejected: its initial residual is not finite or larger than ||b|| (worse
// than a zero start), or the iteration from it overflowed; the caller
// retries from the plain start
constexpr int kStartRejected = 100;

}  // namespace vps

// ---- the context ---------------------------------------------------------
struct vps_context {
    vps_params p{};
    vps::KConst k{};
    int dim = 3;
    int brick_nodes = 512;
    hipStream_t stream = nullptr;
    int device = 0;
    std::string err;

    // real particles (double buffered for the reorder)
    int64_t n = 0;
    vps::DBuf<double4> rpos[2];
    vps::DBuf<float4> rvel[2];
    vps::DBuf<int32_t> rid[2];
    int cur = 0;
    vps::DBuf<float> rrho;
    vps::DBuf<float> rrinv;   // 1 / rrho (3D, for the staged virtual sums)
    // sorted order: FP32 offset of each particle from its cell corner, w =
    // +-1/rho (negative: clamped into a border cell, exact test only)
    vps::DBuf<float4> rofs;
    // cell-sorted boundary, 3D: FP32 offset from the cell corner, w = +-1/rho_b
    // (negative: clamped into a border cell, exact test only)
    vps::DBuf<float4> bofs;
    vps::DBuf<int4> grad_chunks;        // 3D gradient: particle and cell ranges of spans
    vps::DBuf<int32_t> grad_over;       // spans left to the global-memory path
    int grad_span_cells = 0;            // cells per span (0: global-memory gradient)
    int grad_nspan = 0;                 // spans per cell row
    int64_t grad_nchunks = 0;           // non-empty spans of the last generation
    vps::DBuf<float4> vel_next;         // velocities after a Jacobi pass (DEM, viscosity)
    vps::DBuf<double> couple_stage;     // grain coupling: boundary x, n, v
    int64_t couple_walls = -1;          // walls already staged there (-1: none)
    vps::DBuf<int32_t> cell_of, slot_of, perm;
    int64_t ncells = 0;
    vps::DBuf<int32_t> cell_count, cell_start;

    // boundary
    int64_t nb = 0;
    vps::DBuf<double4> bpos;
    vps::DBuf<float4> bvel;   // (v, V_b)
    vps::DBuf<float4> bnrm;   // (n, rho_b)
    vps::DBuf<int32_t> bid;
    vps::DBuf<int32_t> bcell_start;
    vps::DBuf<double4> bpos_spare;   // sort targets of the boundary arrays
    vps::DBuf<float4> bvel_spare, bnrm_spare;
    vps::DBuf<int32_t> bid_spare;
    std::vector<double> host_bx, host_bn, host_bv;

    // lattice / bricks
    int64_t nnodes = 0, nbricks = 0;
    int64_t nslots = 0;                 // active bricks this step
    int which = 0;                      // current slot map (double buffered)
    int last_solve_buf = -1;            // slot map holding the last pressures
    int64_t last_solve_nvirt = 0;       // virtual count at the last solve
    vps::DBuf<uint32_t> brick_flag;
    vps::DBuf<int32_t> brick_slot[2];
    vps::DBuf<int32_t> slot_brick[2];
    // slots of the 27 neighbour bricks of every slot of the current layout
    // (the zero brick nbricks where absent), built after each generation
    vps::DBuf<int32_t> nbr_slot;
    // 3D slots in column order (every z-run of bricks bottom to top, runs in
    // the order of their first slots) and the run lengths / offsets it is
    // built from, after each generation
    vps::DBuf<int32_t> col_run, col_order;
    // per-node fields, slot-major
    vps::DBuf<float> vrho, vvol, vb, vdiag, vx[2], vr, vp[2], vt, vq, vrowsum;
    // 3D CG iteration: second r and t (ping-pong with vr, vt) and the
    // FP64-checked true residual
    vps::DBuf<float> vr2, vt2, vrt;
    vps::DBuf<double> vdiag64;
    vps::DBuf<double> vxb;             // CG: FP64 base of the solution x
    vps::DBuf<int32_t> vrank;
    // lattice bitmask for reference (x-fastest) ranks
    int64_t nwords = 0;
    vps::DBuf<uint32_t> bits, word_prefix;
    int64_t nvirt = 0;
    vps::DBuf<float> p_by_rank[2];

    // scalars
    double rho0 = 0, kappa = 0, alpha0 = -1e6;
    bool initialized = false;
    bool generated = false;
    bool check_true = true;
    int prev_iters = 0;
    // 3D CG: ||b - M x||^2 / (recursive r.r) at the last solve's convergence;
    // the FP32 recursion runs to tol^2 b.b divided by it (at most 4x) so the
    // FP64 check usually passes without a replacement
    double cg_drift = 1.0;
    int64_t start_fallbacks = 0;   // solves restarted from the plain start
    bool start_guard = false;      // cg_solve: the start may be rejected
    // cg_solve queues the post-solve work (fold, after_solve, gradient)
    // behind each batch, guarded, before reading the state back; spec_done:
    // the accepted batch's queued work stands (the caller skips its own)
    bool spec_post = false, spec_done = false;

    // scratch
    vps::DBuf<double> partials;
    vps::DBuf<vps::CgState> cg;
    vps::DBuf<uint32_t> scan_tmp;
    vps::DBuf<int64_t> counters;        // small device counters
    vps::CgState* host_cg = nullptr;    // pinned
    int64_t* host_counters = nullptr;   // pinned

    // timing
    hipEvent_t ev[16] = {};
    std::vector<hipEvent_t> kev;        // per-launch events (SpMV/update)
    bool kernel_timing = false;
    bool async_steps = false;   // steps return without waiting for their end

    // learned warm start (vps_mlp.hip)
    void* mlp = nullptr;
};

// ---- error handling -------------------------------------------------------
#define VPS_HIP(ctx, call)                                                    \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);   \
            return e_ == hipErrorOutOfMemory ? VPS_ERR_OOM : VPS_ERR_HIP;     \
        }                                                                     \
    } while (0)

#define VPS_TRY(expr)           \
    do {                        \
        int rc_ = (expr);       \
        if (rc_) return rc_;    \
    } while (0)

namespace vps {

template <typename T>
int dalloc(vps_context* c, DBuf<T>& b, size_t n) {
    if (b.n >= n && b.p) return VPS_OK;
    if (b.p) {
        (void)hipFree(b.p);
        b.p = nullptr;
        b.n = 0;
    }
    if (n == 0) n = 1;
    VPS_HIP(c, hipMalloc((void**)&b.p, n * sizeof(T)));
    b.n = n;
    return VPS_OK;
}
template <typename T>
void dfree(DBuf<T>& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
}

// launchers (defined in the .hip files)
// 3D gradient set-up at creation: span length and the span list buffer
int gradient_setup(vps_context* c);
// appends the non-empty cell spans of the current sort (count in counters[2])
int gradient_spans(vps_context* c);
// exclusive scan of in into out (out may alias in); clear_in (== in, not
// aliased with out) zeroes the input once read
int scan_exclusive_i32(vps_context* c, const int32_t* in, int32_t* out,
                       int64_t n, int32_t* total_dev, int32_t* clear_in = nullptr);
int cells_sort_real(vps_context* c);
int cells_sort_boundary(vps_context* c);
int upload_boundary_device(vps_context* c, int64_t n, const double* dx,
                           const double* dn, const double* dv);
int real_density(vps_context* c, bool set_volume);
int boundary_density(vps_context* c);
int virtual_generate(vps_context* c);
int virtual_diag(vps_context* c, bool set_alpha0);
int cg_solve(vps_context* c, vps_stats* st);
int64_t partials_size(int64_t nbricks);
int real_offsets(vps_context* c);
int apply_operator_f32(vps_context* c, const float* in, float* out,
                       int op_mode);
int pressure_gradient(vps_context* c, SolveGuard g = SolveGuard{});
// test support: terms per particle (sorted order) of the FP32 walks
int density_member_counts(vps_context* c, int which, int32_t* members);
int gradient_member_counts(vps_context* c, int32_t* members);
int viscosity_member_counts(vps_context* c, int32_t* members);
int advect(vps_context* c, double dt, double g);
int query_neighbors(vps_context* c, int which, int64_t nq, const double* q,
                    int64_t* offsets, int64_t* idx, int64_t cap,
                    int64_t* needed);
int reduce_max_f32(vps_context* c, const float* in, int64_t n, double* out);
int gather_virtual(vps_context* c, int64_t* node, double* rho, double* vol,
                   double* div, double* diag, double* pressure);
int warm_start(vps_context* c, int mode, float* x0);
// after_solve: the kernel (guarded) and, with an unguarded call, the host
// bookkeeping; after_solve_commit: the bookkeeping alone
int after_solve(vps_context* c, SolveGuard g = SolveGuard{});
void after_solve_commit(vps_context* c);

// learned warm start
int mlp_create(vps_context* c);
void mlp_destroy(vps_context* c);
int mlp_warm_start(vps_context* c, int mode, float* x0);
int mlp_after_solve(vps_context* c, const vps_stats* st);
bool mlp_active(const vps_context* c);
// the rest of the SPH step (vps_sph.hip)
int move_collide(vps_context* c, double dt, const vps_collider* col, int32_t ncol);
int apply_gravity(vps_context* c, double dt, double g);
int apply_viscosity(vps_context* c, double dt, double nu);
// checkpoint of the learned warm start: parameters, Adam moments and, for
// slot map lb (nslots slots; lb < 0: none), the node history of the last solve
int mlp_state_save(vps_context* c, int lb, int64_t nslots, std::vector<char>& out);
int mlp_state_load(vps_context* c, int lb, int64_t nslots, const char* p, size_t n);


}  // namespace vps
