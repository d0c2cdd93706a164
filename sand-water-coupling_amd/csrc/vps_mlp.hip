// vps_mlp.hip -- learned CG warm start (no counterpart in the reference).
//
// Every lattice node keeps the history of its pressure over the last three
// solves: p1 (last), dp1 = p1 - p2, dp2 = p2 - p3, and the depth h (how many
// of them are known; zero for a node absent from the last solve).  The
// polynomial extrapolation through the known values is the prior
//     e = 0 (h <= 1),  dp1 (h = 2),  2 dp1 - dp2 (h = 3)   (p_next ~ p1 + e)
// and a per-virtual-particle MLP 8 -> 64 -> 64 -> 1 (tanh-form GELU, BF16 weights and
// activations, FP32 accumulation) learns the correction c = tau * net(f)
// to it.  The start is the Galerkin projection of the solution onto the span
// of the three fields (zero on nodes that are new):
//     x0 = y0 p1 + y1 e + y2 c,   (V^T M V) y = V^T b,   V = [p1 e c]
// (the residual of x0 is orthogonal to the three fields; for a symmetric M
// it is the start of least M-norm error in that span, which holds the
// node-keyed start p1, the extrapolation p1 + e and the network's own
// p1 + e + c).  It costs three applications of M and twelve dot products per
// solve.  The network's features are
//     f0 = p1 / ps, f1 = dp1 / dps, f2 = dp2 / dps    the node history
//     f3 = D_I / ds                                the CG right-hand side
//     f4 = diag_I / alpha0 - 1, f5 = s_I / alpha0 - 1   clamped / unclamped
//          Laplacian diagonal (s < alpha0 marks the free surface)
//     f6 = V_I / (m / rho0) - 1                    virtual volume
//     f7 = history depth (0, 0.5, 1 for h = 1, 2, 3)
// ps = RMS of the last solution, dps = RMS of its increment dp1 (over nodes
// with h >= 1), ds = RMS of the last right-hand side, tau = RMS of the
// extrapolation's own error at the last solve (over nodes with h >= 1): the
// network's output is in units of the error it corrects, so its BF16 and
// Adam-step noise stays far below that error as the flow settles.  The
// scales live on the device and are refreshed after every solve, so no host
// round trip is added.  The output layer starts at zero: an untrained network
// (c = 0) gives the Galerkin start over p1 and e.  CG always runs to the same
// tolerance, so the warm start changes the iteration count, not the answer.
//
// Kernels (all matrix work on MFMA v_mfma_f32_16x16x16_bf16, 64-lane
// waves, tiles of 16 nodes):
//   forward: activations are kept transposed (hidden x node), so the C
//     layout of one layer's MFMA output is exactly the B operand layout of
//     the next layer -- no data movement between layers;
//   backward: dH1 = W2^T dZ2 on MFMA from the same layouts; the weight
//     gradients dW2 = dZ2 H1^T and dW1 = dZ1 X^T contract over nodes, which
//     needs the transposed layout: each wave stages its bf16 tiles through
//     LDS (8 KB) and reads them back with the lane<->row swap;
//   per-wave gradient partials are reduced in a fixed order, then Adam.
// MFMA lane layouts (16x16x16): A lane l holds A[l%16][4(l/16)+j], B lane l
// holds B[4(l/16)+j][l%16], C lane l holds C[4(l/16)+i][l%16].
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "vps_internal.h"

namespace vps {
namespace {

constexpr int kIn = 8;
constexpr int kHid = 64;
constexpr int OFF_W1 = 0;                       // W1[64][8]
constexpr int OFF_B1 = OFF_W1 + kHid * kIn;     // 512
constexpr int OFF_W2 = OFF_B1 + kHid;           // 576, W2[64][64] (out, in)
constexpr int OFF_B2 = OFF_W2 + kHid * kHid;    // 4672
constexpr int OFF_W3 = OFF_B2 + kHid;           // 4736, W3[64]
constexpr int OFF_B3 = OFF_W3 + kHid;           // 4800
constexpr int kNParams = OFF_B3 + 1;            // 4801
// device scales: inverses of ps, dps, ds, tau; tau; dps; ps
enum {
    SC_INV_PS = 0, SC_INV_DPS = 1, SC_INV_DS = 2, SC_INV_TAU = 3, SC_TAU = 4,
    SC_DPS = 5, SC_PS = 6, SC_N = 8
};
constexpr int kTrainBlocks = 256;               // one block of 4 waves per CU
constexpr int kTrainWaves = 4 * kTrainBlocks;
// largest batch of vps_mlp_loss_and_grad(_device): 4,096 tiles of 16 samples
// per training wave
constexpr int64_t kMaxLossBatch = (int64_t)16 * kTrainWaves * 4096;
constexpr int kInferBlocks = 1024;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> bf16, round to nearest even (v_cvt_pk_bf16_f32 on gfx950).
__device__ __forceinline__ short f2bf(float f) {
    return __builtin_bit_cast(short, (__bf16)f);
}
typedef __bf16 bf16x4v __attribute__((ext_vector_type(4)));
// four values rounded to BF16 with two v_cvt_pk_bf16_f32 (same rounding as
// f2bf)
__device__ __forceinline__ bf16x4 f2bf4(float a, float b, float c, float d) {
    const f32x4 v = {a, b, c, d};
    return __builtin_bit_cast(bf16x4, __builtin_convertvector(v, bf16x4v));
}
__device__ __forceinline__ float bf2f(short b) {
    return __uint_as_float(((uint32_t)(uint16_t)b) << 16);
}
__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// GELU in its tanh form: gelu(z) = 0.5 z (1 + tanh(u)) = z sigmoid(2u),
// u = sqrt(2/pi) (z + 0.044715 z^3).  sigmoid(2u) = 1 / (1 + 2^(-2u log2 e)):
// one v_exp_f32 and one v_rcp_f32 per value (the network's values are
// rounded to BF16, far above their 1-ulp errors).  Large |u| saturates to
// 0 or 1: 2^+big = inf and 1 / inf = 0.
__device__ __forceinline__ float gelu_sig(float z) {
    // 2 sqrt(2/pi) log2(e)
    const float u2l = 2.3022081985f * z * (1.f + 0.044715f * z * z);
    return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-u2l));
}
__device__ __forceinline__ float gelu(float z) { return z * gelu_sig(z); }
// d/dz [z s(2u)] = s + z s (1 - s) 2u'(z)
__device__ __forceinline__ float gelu_d(float z) {
    const float s = gelu_sig(z);
    const float du2 = 1.5957691216057308f * (1.f + 0.134145f * z * z);
    return s + z * s * (1.f - s) * du2;
}

struct Wregs {
    bf16x4 w1[4];        // A tiles of W1 (hidden x feature), per m-tile
    bf16x4 w2[4][4];     // A tiles of W2 [m-tile out][k-tile in]
    bf16x4 w3[4];        // A tiles of W3 (row 0 only), per k-tile
    f32x4 b1[4], b2[4];  // biases in C layout
    float b3;
};

template <bool W2REG>
__device__ __forceinline__ void load_weights(const float* P, Wregs& W) {
    const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = 4 * q + j;
            W.w1[mt][j] = k < kIn ? f2bf(P[OFF_W1 + (mt * 16 + r) * kIn + k]) : (short)0;
            W.b1[mt][j] = P[OFF_B1 + mt * 16 + 4 * q + j];
            W.b2[mt][j] = P[OFF_B2 + mt * 16 + 4 * q + j];
            W.w3[mt][j] = r == 0 ? f2bf(P[OFF_W3 + mt * 16 + 4 * q + j]) : (short)0;
        }
        if constexpr (W2REG) {
#pragma unroll
            for (int kt = 0; kt < 4; kt++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    W.w2[mt][kt][j] =
                        f2bf(P[OFF_W2 + (mt * 16 + r) * kHid + kt * 16 + 4 * q + j]);
        }
    }
    W.b3 = P[OFF_B3];
}

struct Fwd {
    f32x4 z1[4], z2[4];
    bf16x4 h1[4], h2[4];
    float out;   // lanes 0..15: output of node l
};

// W2 A tile (m-tile mo, k-tile kt) read from an LDS copy W2[out][in] (bf16)
__device__ __forceinline__ bf16x4 w2_tile(const uint16_t* sW, int mo, int kt) {
    const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
    const uint16_t* p = sW + (mo * 16 + r) * kHid + kt * 16 + 4 * q;
    bf16x4 v;
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = (short)p[j];
    return v;
}

template <bool LDSW2>
__device__ __forceinline__ void forward(const Wregs& W, const uint16_t* sW2,
                                        bf16x4 xb, Fwd& F) {
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        F.z1[mt] = mfma(W.w1[mt], xb, W.b1[mt]);
        F.h1[mt] = f2bf4(gelu(F.z1[mt][0]), gelu(F.z1[mt][1]), gelu(F.z1[mt][2]),
                         gelu(F.z1[mt][3]));
    }
#pragma unroll
    for (int mo = 0; mo < 4; mo++) {
        f32x4 acc = W.b2[mo];
#pragma unroll
        for (int kt = 0; kt < 4; kt++) {
            if constexpr (LDSW2) acc = mfma(w2_tile(sW2, mo, kt), F.h1[kt], acc);
            else acc = mfma(W.w2[mo][kt], F.h1[kt], acc);
        }
        F.z2[mo] = acc;
        F.h2[mo] = f2bf4(gelu(acc[0]), gelu(acc[1]), gelu(acc[2]), gelu(acc[3]));
    }
    f32x4 z3 = {W.b3, W.b3, W.b3, W.b3};
#pragma unroll
    for (int kt = 0; kt < 4; kt++) z3 = mfma(W.w3[kt], F.h2[kt], z3);
    F.out = z3[0];
}

// Feature sources: the current step's slot-major fields.
struct FeatSrc {
    const float* p1;
    const float* dp1;
    const float* dp2;
    const uint8_t* hh;
    const float* vb;
    const float* vdiag;
    const float* vrow;
    const float* vvol;
    const float* scale;   // device scales (SC_*)
    float inv_a0, inv_vs;
    float* corr;          // the network's correction c per node, or null
};

// Polynomial extrapolation of the node history (the prior of the network).
__device__ __forceinline__ float prior(int h, float dp1, float dp2) {
    return h >= 3 ? 2.f * dp1 - dp2 : (h == 2 ? dp1 : 0.f);
}

__device__ __forceinline__ float feature(const FeatSrc& s, int64_t g, int f) {
    switch (f) {
        case 0: return s.p1[g] * s.scale[SC_INV_PS];
        case 1: return s.dp1[g] * s.scale[SC_INV_DPS];
        case 2: return s.dp2[g] * s.scale[SC_INV_DPS];
        case 3: return s.vb[g] * s.scale[SC_INV_DS];
        case 4: return s.vdiag[g] * s.inv_a0 - 1.f;
        case 5: return s.vrow[g] * s.inv_a0 - 1.f;
        case 6: return s.vvol[g] * s.inv_vs - 1.f;
        default: {
            const int h = s.hh[g];
            return h >= 3 ? 1.f : (h == 2 ? 0.5f : 0.f);
        }
    }
}

// B operand (feature x node) of the 16-node tile: lane l carries features
// 4(l/16)..4(l/16)+3 of node l%16 (zero beyond the 8 features).
template <bool DIRECT>
__device__ __forceinline__ bf16x4 tile_features(const FeatSrc& s,
                                                const float* direct, int64_t n,
                                                int64_t node, bool ok,
                                                float fo[4]) {
    const int q = (threadIdx.x & 63) >> 4;
    bf16x4 xb = {0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 4; j++) {
        float v = 0.f;
        if (ok && q < 2) {
            const int f = 4 * q + j;
            if constexpr (DIRECT) v = direct[node * kIn + f];
            else v = feature(s, node, f);
        }
        fo[j] = v;
        xb[j] = f2bf(v);
    }
    (void)n;
    return xb;
}

// ---- inference: x0 = p1 + e + tau * net(f) ------------------------------
template <bool DIRECT>
__global__ void __launch_bounds__(256) k_mlp_infer(const float* P, FeatSrc fs,
                                                   const float* direct,
                                                   int64_t nn, float* out) {
    Wregs W;
    load_weights<true>(P, W);
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t nwaves = (int64_t)gridDim.x * 4;
    const int64_t ntiles = (nn + 15) / 16;
    for (int64_t tile = wave; tile < ntiles; tile += nwaves) {
        const int64_t g = tile * 16 + (lane & 15);
        bool ok = g < nn;
        if constexpr (!DIRECT) {
            // only nodes with a history get the network; whole tiles without
            // one (empty brick slots, new nodes) skip it (wave-uniform)
            ok = ok && fs.vvol[g] > 0.f && fs.hh[g] >= 1;
            if (!__any(ok)) {
                if (lane < 16 && g < nn) out[g] = 0.f;
                continue;
            }
        }
        float fo[4];
        const bf16x4 xb = tile_features<DIRECT>(fs, direct, nn, g, ok, fo);
        Fwd F;
        forward<false>(W, nullptr, xb, F);
        if (lane < 16 && g < nn) {
            if constexpr (DIRECT) {
                out[g] = F.out;
            } else {
                const int h = fs.hh[g];
                out[g] = ok ? fs.p1[g] + prior(h, fs.dp1[g], fs.dp2[g]) +
                                  fs.scale[SC_TAU] * F.out
                            : 0.f;
            }
        }
    }
}

// Raw per-lane inputs of one 16-node tile (loaded one tile ahead): lane
// (r, q) holds the 4 feature sources of node r for q < 2, and every lane the
// node's V and history depth (which decide whether it takes the network).
struct RawTile {
    float v[4];
    float vol;
    int h;
    int64_t g;
};

__device__ __forceinline__ void load_raw(const FeatSrc& s, int64_t tile, int64_t nn,
                                         RawTile& R) {
    const int lane = threadIdx.x & 63, q = lane >> 4;
    R.g = tile * 16 + (lane & 15);
    R.v[0] = R.v[1] = R.v[2] = R.v[3] = 0.f;
    R.vol = 0.f;
    R.h = 0;
    if (R.g >= nn) return;
    R.vol = s.vvol[R.g];
    R.h = s.hh[R.g];
    if (q == 0) {
        R.v[0] = s.p1[R.g];
        R.v[1] = s.dp1[R.g];
        R.v[2] = s.dp2[R.g];
        R.v[3] = s.vb[R.g];
    } else if (q == 1) {
        R.v[0] = s.vdiag[R.g];
        R.v[1] = s.vrow[R.g];
    }
}

// Inference over the slot nodes with the next tile's loads in flight while
// the current one runs through the network.
__global__ void __launch_bounds__(256) k_mlp_infer_pipe(const float* P, FeatSrc fs,
                                                        int64_t nn, float* out) {
    // the block reads the parameters once (coalesced) into LDS; each wave
    // builds its register tiles from there
    __shared__ float sP[kNParams];
    for (int i = threadIdx.x; i < kNParams; i += 256) sP[i] = P[i];
    __syncthreads();
    Wregs W;
    load_weights<true>(sP, W);
    const int lane = threadIdx.x & 63, q = lane >> 4;
    const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t nwaves = (int64_t)gridDim.x * 4;
    const int64_t ntiles = (nn + 15) / 16;
    const float inv_ps = fs.scale[SC_INV_PS], inv_dps = fs.scale[SC_INV_DPS];
    const float inv_ds = fs.scale[SC_INV_DS], tau = fs.scale[SC_TAU];
    RawTile cur, nxt;
    if (wave < ntiles) load_raw(fs, wave, nn, cur);
    for (int64_t tile = wave; tile < ntiles; tile += nwaves) {
        if (tile + nwaves < ntiles) load_raw(fs, tile + nwaves, nn, nxt);
        const bool ok = cur.g < nn && cur.vol > 0.f && cur.h >= 1;
        if (__any(ok)) {
            float f[4];
            if (q == 0) {
                f[0] = cur.v[0] * inv_ps;
                f[1] = cur.v[1] * inv_dps;
                f[2] = cur.v[2] * inv_dps;
                f[3] = cur.v[3] * inv_ds;
            } else if (q == 1) {
                f[0] = cur.v[0] * fs.inv_a0 - 1.f;
                f[1] = cur.v[1] * fs.inv_a0 - 1.f;
                f[2] = cur.vol * fs.inv_vs - 1.f;
                f[3] = cur.h >= 3 ? 1.f : (cur.h == 2 ? 0.5f : 0.f);
            } else {
                f[0] = f[1] = f[2] = f[3] = 0.f;
            }
            const bf16x4 xb = f2bf4(ok ? f[0] : 0.f, ok ? f[1] : 0.f, ok ? f[2] : 0.f,
                                    ok ? f[3] : 0.f);
            Fwd F;
            forward<false>(W, nullptr, xb, F);
            // lanes 0..15 hold node r's output and (q = 0) its raw history
            if (lane < 16 && cur.g < nn) {
                out[cur.g] = ok ? cur.v[0] + prior(cur.h, cur.v[1], cur.v[2]) + tau * F.out
                                : 0.f;
                if (fs.corr) fs.corr[cur.g] = ok ? tau * F.out : 0.f;
            }
        } else if (lane < 16 && cur.g < nn) {
            out[cur.g] = 0.f;
            if (fs.corr) fs.corr[cur.g] = 0.f;
        }
        cur = nxt;
    }
}

// The fields of the Galerkin start: P = p1 and E = e on the nodes with a
// history, zero elsewhere (C, the network's correction, is written by the
// inference).
__global__ void __launch_bounds__(256) k_galerkin_fields(FeatSrc fs, int64_t nn, float* P,
                                                         float* E) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= nn) return;
    const int h = fs.hh[g];
    const bool ok = fs.vvol[g] > 0.f && h >= 1;
    P[g] = ok ? fs.p1[g] : 0.f;
    E[g] = ok ? prior(h, fs.dp1[g], fs.dp2[g]) : 0.f;
}

// Block partials of v_i . M v_j (i, j = 0..2) and v_i . b for V = [P E C]
// (grid-stride over the nodes: kGramBlocks partials for the solve to sum).
constexpr int kNGram = 12;
constexpr int kGramBlocks = 512;
__global__ void __launch_bounds__(256) k_galerkin_gram(int64_t nn, const float* P,
                                                       const float* E, const float* Cc,
                                                       const float* MP, const float* ME,
                                                       const float* MC, const float* b,
                                                       double* part) {
    double a[kNGram];
#pragma unroll
    for (int j = 0; j < kNGram; j++) a[j] = 0;
    for (int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x; g < nn;
         g += (int64_t)gridDim.x * 256) {
        const double v[3] = {P[g], E[g], Cc[g]};
        const double mv[3] = {MP[g], ME[g], MC[g]};
        const double bg = b[g];
#pragma unroll
        for (int i = 0; i < 3; i++) {
#pragma unroll
            for (int j = 0; j < 3; j++) a[3 * i + j] += v[i] * mv[j];
            a[9 + i] += v[i] * bg;
        }
    }
    __shared__ double sw[kNGram][4];
#pragma unroll
    for (int j = 0; j < kNGram; j++) {
        a[j] = wave_sum(a[j]);
        if ((threadIdx.x & 63) == 0) sw[j][threadIdx.x >> 6] = a[j];
    }
    __syncthreads();
    if (threadIdx.x < kNGram) {
        const double* w = sw[threadIdx.x];
        part[(int64_t)blockIdx.x * kNGram + threadIdx.x] = (w[0] + w[1]) + (w[2] + w[3]);
    }
}

// Sums the partials (fixed order) and solves the 3 x 3 Galerkin system
// (V^T M V) y = V^T b by elimination with partial pivoting over the fields
// that are present (a field with v . M v = 0 takes weight 0); y = (1, 1, 1),
// the network's own start, if the system is singular or the weights are not
// finite.
__global__ void __launch_bounds__(1024) k_galerkin_solve(const double* part, int nblk,
                                                         double* y) {
    __shared__ double sw[kNGram][16];
    double a[kNGram];
#pragma unroll
    for (int j = 0; j < kNGram; j++) a[j] = 0;
    for (int i = threadIdx.x; i < nblk; i += 1024)
#pragma unroll
        for (int j = 0; j < kNGram; j++) a[j] += part[(int64_t)i * kNGram + j];
#pragma unroll
    for (int j = 0; j < kNGram; j++) {
        a[j] = wave_sum(a[j]);
        if ((threadIdx.x & 63) == 0) sw[j][threadIdx.x >> 6] = a[j];
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double t[kNGram];
#pragma unroll
    for (int j = 0; j < kNGram; j++) {
        t[j] = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) t[j] += sw[j][i];
    }
    // a field that is absent (v . M v not positive) gets the row and column
    // of the identity and a zero right-hand side, so its weight is 0; all
    // indices below are compile-time constants (registers, no scratch)
    double G[3][3], r[3], dg[3];
    bool act[3];
#pragma unroll
    for (int i = 0; i < 3; i++) act[i] = t[4 * i] > 0 && isfinite(t[4 * i]);
    const bool any = act[0] || act[1] || act[2];
#pragma unroll
    for (int p = 0; p < 3; p++) {
        r[p] = act[p] ? t[9 + p] : 0.0;
#pragma unroll
        for (int q = 0; q < 3; q++)
            G[p][q] = act[p] && act[q] ? t[3 * p + q] : (p == q ? 1.0 : 0.0);
        dg[p] = fabs(G[p][p]);
    }
    // Gaussian elimination with partial pivoting; a pivot not above 1e-10 of
    // its row's diagonal (a field nearly in the span of the others) rejects
    bool ok = any;
#pragma unroll
    for (int k = 0; k < 3; k++) {
#pragma unroll
        for (int p = k + 1; p < 3; p++) {
            if (fabs(G[p][k]) > fabs(G[k][k])) {
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    const double tmp = G[k][q];
                    G[k][q] = G[p][q];
                    G[p][q] = tmp;
                }
                const double tr = r[k];
                r[k] = r[p];
                r[p] = tr;
            }
        }
        ok = ok && fabs(G[k][k]) > 1e-10 * dg[k];
        const double piv = ok ? G[k][k] : 1.0;
#pragma unroll
        for (int p = k + 1; p < 3; p++) {
            const double f = G[p][k] / piv;
#pragma unroll
            for (int q = 0; q < 3; q++) G[p][q] -= f * G[k][q];
            r[p] -= f * r[k];
        }
    }
    double x[3];
#pragma unroll
    for (int p = 2; p >= 0; p--) {
        double sum = r[p];
#pragma unroll
        for (int q = p + 1; q < 3; q++) sum -= G[p][q] * x[q];
        x[p] = sum / (ok ? G[p][p] : 1.0);
    }
    const bool fin = isfinite(x[0]) && isfinite(x[1]) && isfinite(x[2]);
    double out[3];
#pragma unroll
    for (int i = 0; i < 3; i++) out[i] = ok && fin ? (act[i] ? x[i] : 0.0) : 1.0;
    for (int i = 0; i < 3; i++) y[i] = out[i];
}

__global__ void __launch_bounds__(256) k_galerkin_combine(int64_t nn, const float* P,
                                                          const float* E, const float* Cc,
                                                          const double* y, float* x0) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= nn) return;
    const float y0 = (float)y[0], y1 = (float)y[1], y2 = (float)y[2];
    x0[g] = fmaf(y2, Cc[g], fmaf(y1, E[g], y0 * P[g]));
}

// x0 = p1 + e (the untrained network's answer, without the MLP).
__global__ void __launch_bounds__(256) k_extrapolate(FeatSrc fs, int64_t nn,
                                                     float* out) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= nn) return;
    const int h = fs.hh[g];
    out[g] = fs.vvol[g] > 0.f && h >= 1 ? fs.p1[g] + prior(h, fs.dp1[g], fs.dp2[g])
                                        : 0.f;
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// ---- training: forward + backward over sampled nodes ---------------------
// Per wave: tiles_per_wave tiles of 16 samples.  The 4 waves of a block sum
// their gradients through LDS (fixed order); the block writes one row of
// kNParams floats and its loss partial.  Loss = sum (out - t)^2 / nsamples.
template <bool DIRECT>
__global__ void __launch_bounds__(256) k_mlp_train(
    const float* P, FeatSrc fs, const float* vx, const int32_t* slot_of_rank,
    int64_t nv, const float* fdirect, const float* tdirect, int64_t nsamples,
    int tiles_per_wave, uint64_t seed, float* grad_part, double* loss_part) {
    // per-wave bf16 staging tiles; reused as the FP32 gradient exchange of
    // the block reduction (2 x kNParams floats) after the sample loop
    constexpr int kTileBytes = 4 * (3 * kHid * 16 + 16 * 16) * 2;
    constexpr int kRedBytes = 2 * kNParams * 4;
    constexpr int kStageBytes = kTileBytes > kRedBytes ? kTileBytes : kRedBytes;
    __shared__ __align__(16) uint8_t s_stage[kStageBytes];
    __shared__ double s_loss[4];
    auto s_dz2 = reinterpret_cast<uint16_t(*)[kHid * 16]>(s_stage);
    auto s_h1 = reinterpret_cast<uint16_t(*)[kHid * 16]>(s_stage + 4 * kHid * 16 * 2);
    auto s_dz1 = reinterpret_cast<uint16_t(*)[kHid * 16]>(s_stage + 8 * kHid * 16 * 2);
    auto s_x = reinterpret_cast<uint16_t(*)[16 * 16]>(s_stage + 12 * kHid * 16 * 2);
    float* s_red = reinterpret_cast<float*>(s_stage);
    __shared__ uint16_t sW2[kHid * kHid];    // W2[out][in]
    __shared__ uint16_t sW2T[kHid * kHid];   // W2^T[in][out]
    const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
    const int wl = threadIdx.x >> 6;
    const int64_t wave = (int64_t)blockIdx.x * 4 + wl;
    for (int e = threadIdx.x; e < kHid * kHid; e += 256) {
        const uint16_t b = (uint16_t)f2bf(P[OFF_W2 + e]);
        sW2[e] = b;
        sW2T[(e % kHid) * kHid + e / kHid] = b;
    }
    Wregs W;
    load_weights<false>(P, W);
    // W3 in C layout
    f32x4 w3c[4];
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
        for (int i = 0; i < 4; i++) w3c[mi][i] = bf2f(f2bf(P[OFF_W3 + mi * 16 + 4 * q + i]));
    __syncthreads();
    f32x4 gW2[4][4], gW1[4], gb1[4], gb2[4], gW3[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
        gW1[a] = {0, 0, 0, 0};
        gb1[a] = {0, 0, 0, 0};
        gb2[a] = {0, 0, 0, 0};
        gW3[a] = {0, 0, 0, 0};
#pragma unroll
        for (int b = 0; b < 4; b++) gW2[a][b] = {0, 0, 0, 0};
    }
    float gb3 = 0.f;
    double loss = 0;
    const float inv_n = 1.f / (float)nsamples;
    for (int it = 0; it < tiles_per_wave; it++) {
        const int64_t k = (wave * tiles_per_wave + it) * 16 + r;   // sample id
        bool ok = k < nsamples;
        int64_t node = k;
        float target = 0.f;
        if constexpr (DIRECT) {
            if (ok) target = tdirect[k];
        } else {
            // uniform over the virtual particles of this solve
            node = ok ? slot_of_rank[mix64(seed ^ (uint64_t)k) % (uint64_t)nv] : 0;
            const int h = fs.hh[node];
            ok = ok && h >= 1;
            if (ok)
                target = (vx[node] - fs.p1[node] -
                          prior(h, fs.dp1[node], fs.dp2[node])) *
                         fs.scale[SC_INV_TAU];
        }
        float fo[4];
        const bf16x4 xb = tile_features<DIRECT>(fs, fdirect, nsamples, node, ok, fo);
        Fwd F;
        forward<true>(W, sW2, xb, F);
        // dL/dout for the 16 nodes (lanes 0..15), broadcast to the columns
        float e = (lane < 16 && ok) ? F.out - target : 0.f;
        if (lane < 16) loss += (double)e * (double)e;
        const float dout = 2.f * e * inv_n;
        const float d = __shfl(dout, r, 64);
        if (lane < 16) gb3 += d;
        // layer 3
        f32x4 dz2[4];
#pragma unroll
        for (int kt = 0; kt < 4; kt++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                gW3[kt][i] += bf2f(F.h2[kt][i]) * d;
                dz2[kt][i] = w3c[kt][i] * d * gelu_d(F.z2[kt][i]);
                gb2[kt][i] += dz2[kt][i];
            }
        bf16x4 dz2b[4];
#pragma unroll
        for (int kt = 0; kt < 4; kt++)
#pragma unroll
            for (int i = 0; i < 4; i++) dz2b[kt][i] = f2bf(dz2[kt][i]);
        // dH1 = W2^T dZ2; dZ1 = dH1 * gelu'(Z1)
        f32x4 dz1[4];
#pragma unroll
        for (int mi = 0; mi < 4; mi++) {
            f32x4 acc = {0, 0, 0, 0};
#pragma unroll
            for (int kt = 0; kt < 4; kt++) acc = mfma(w2_tile(sW2T, mi, kt), dz2b[kt], acc);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                dz1[mi][i] = acc[i] * gelu_d(F.z1[mi][i]);
                gb1[mi][i] += dz1[mi][i];
            }
        }
        // stage dZ2, H1, dZ1 (hidden x node) and X (feature x node) in LDS
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int row = t * 16 + 4 * q + i;
                s_dz2[wl][row * 16 + r] = (uint16_t)dz2b[t][i];
                s_h1[wl][row * 16 + r] = (uint16_t)F.h1[t][i];
                s_dz1[wl][row * 16 + r] = (uint16_t)f2bf(dz1[t][i]);
            }
#pragma unroll
        for (int j = 0; j < 4; j++) s_x[wl][(4 * q + j) * 16 + r] = (uint16_t)xb[j];
        __syncthreads();
        // node-contracted products: A(hidden x node), B(node x hidden/feature)
        bf16x4 a_dz2[4], b_h1[4], a_dz1[4];
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                a_dz2[t][j] = (short)s_dz2[wl][(t * 16 + r) * 16 + 4 * q + j];
                b_h1[t][j] = (short)s_h1[wl][(t * 16 + r) * 16 + 4 * q + j];
                a_dz1[t][j] = (short)s_dz1[wl][(t * 16 + r) * 16 + 4 * q + j];
            }
        bf16x4 b_x;
#pragma unroll
        for (int j = 0; j < 4; j++) b_x[j] = (short)s_x[wl][r * 16 + 4 * q + j];
        __syncthreads();
#pragma unroll
        for (int mo = 0; mo < 4; mo++)
#pragma unroll
            for (int mi = 0; mi < 4; mi++)
                gW2[mo][mi] = mfma(a_dz2[mo], b_h1[mi], gW2[mo][mi]);
#pragma unroll
        for (int mt = 0; mt < 4; mt++) gW1[mt] = mfma(a_dz1[mt], b_x, gW1[mt]);
    }
    // reduce the per-lane bias / W3 partials over the 16 node lanes
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                gb1[t][i] += __shfl_xor(gb1[t][i], o, 64);
                gb2[t][i] += __shfl_xor(gb2[t][i], o, 64);
                gW3[t][i] += __shfl_xor(gW3[t][i], o, 64);
            }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        gb3 += __shfl_xor(gb3, o, 64);
        loss += __shfl_xor(loss, o, 64);
    }
    // Every gradient entry this lane holds, with its parameter index (each
    // index is held by exactly one lane of the wave); f returns the new value.
    auto each = [&](auto&& f) {
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int row = t * 16 + 4 * q + i;
                if (r < kIn) gW1[t][i] = f(OFF_W1 + row * kIn + r, gW1[t][i]);
                if (r == 0) {
                    gb1[t][i] = f(OFF_B1 + row, gb1[t][i]);
                    gb2[t][i] = f(OFF_B2 + row, gb2[t][i]);
                    gW3[t][i] = f(OFF_W3 + row, gW3[t][i]);
                }
#pragma unroll
                for (int mi = 0; mi < 4; mi++)
                    gW2[t][mi][i] = f(OFF_W2 + row * kHid + mi * 16 + r, gW2[t][mi][i]);
            }
        if (lane == 0) gb3 = f(OFF_B3, gb3);
    };
    // block sum in a fixed order: (w0 + w2) + (w1 + w3)
    if (lane == 0) s_loss[wl] = loss;
    __syncthreads();                      // staging tiles no longer read
    if (wl >= 2)
        each([&](int idx, float v) {
            s_red[(wl - 2) * kNParams + idx] = v;
            return v;
        });
    __syncthreads();
    if (wl < 2) each([&](int idx, float v) { return v + s_red[wl * kNParams + idx]; });
    __syncthreads();
    if (wl == 1)
        each([&](int idx, float v) {
            s_red[idx] = v;
            return v;
        });
    __syncthreads();
    if (wl == 0) {
        float* G = grad_part + (int64_t)blockIdx.x * kNParams;
        each([&](int idx, float v) {
            G[idx] = v + s_red[idx];
            return v;
        });
        if (lane == 0)
            loss_part[blockIdx.x] = ((s_loss[0] + s_loss[2]) + (s_loss[1] + s_loss[3])) *
                                    (double)inv_n;
    }
}

// Sum of the per-block partials (fixed order).  Block = 64 parameters x 4
// row groups; each thread sums rows g, g + 4, ... of its parameter, the 4
// groups are added in order.
__global__ void __launch_bounds__(256) k_mlp_reduce(
    const float* grad_part, int nrows, const double* loss_part, float* grad,
    double* loss) {
    __shared__ float s_g[4][64];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + tx;
    if (blockIdx.x == 0 && threadIdx.x < 64) {
        double s = 0;
        for (int w = threadIdx.x; w < nrows; w += 64) s += loss_part[w];
        s = wave_sum(s);
        if (threadIdx.x == 0) *loss = s;
    }
    float g = 0.f;
    if (p < kNParams) {
#pragma unroll 8
        for (int w = ty; w < nrows; w += 4) g += grad_part[(int64_t)w * kNParams + p];
    }
    s_g[ty][tx] = g;
    __syncthreads();
    if (ty != 0 || p >= kNParams) return;
    grad[p] = (s_g[0][tx] + s_g[1][tx]) + (s_g[2][tx] + s_g[3][tx]);
}

// One Adam step (beta 0.9 / 0.999, eps 1e-8) on the reduced gradient, in one
// block so that the whole step is taken or skipped: a non-finite loss or
// gradient entry (a batch the network overflows on) leaves P, m and v as
// they are and counts the skip.
__global__ void __launch_bounds__(1024) k_mlp_adam(const float* grad, const double* loss,
                                                   float* P, float* m, float* v, float lr,
                                                   float b1t, float b2t, int32_t* skipped) {
    bool bad = !isfinite(*loss);
    for (int p = threadIdx.x; p < kNParams; p += 1024) bad |= !isfinite(grad[p]);
    if (__syncthreads_or(bad)) {
        if (threadIdx.x == 0) *skipped += 1;
        return;
    }
    for (int p = threadIdx.x; p < kNParams; p += 1024) {
        const float g = grad[p];
        const float mm = 0.9f * m[p] + 0.1f * g;
        const float vv = 0.999f * v[p] + 0.001f * g * g;
        m[p] = mm;
        v[p] = vv;
        const float mh = mm / (1.f - b1t), vh = vv / (1.f - b2t);
        P[p] -= lr * mh / (sqrtf(vh) + 1e-8f);
    }
}

// Node history in the current slot layout, gathered from the last solve.
__global__ void __launch_bounds__(256) k_history_gather(
    int64_t nn, int NB, const int32_t* brick_of_slot, const float* vvol,
    const int32_t* prev_slot_of_brick, const float* P_prev, const float* DP_prev,
    const float* DP2_prev, const uint8_t* H_prev, float* hp1, float* hdp1,
    float* hdp2, uint8_t* hh) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= nn) return;
    float a = 0.f, b = 0.f, c = 0.f;
    int h = 0;
    if (prev_slot_of_brick && vvol[g] > 0.f) {
        const int slot = (int)(g / NB), loc = (int)(g % NB);
        const int ps = prev_slot_of_brick[brick_of_slot[slot]];
        if (ps >= 0) {
            const int64_t pg = (int64_t)ps * NB + loc;
            h = H_prev[pg];
            if (h >= 1) a = P_prev[pg];
            if (h >= 2) b = DP_prev[pg];
            if (h >= 3) c = DP2_prev[pg];
        }
    }
    hp1[g] = a;
    hdp1[g] = b;
    hdp2[g] = c;
    hh[g] = (uint8_t)h;
}

// After the solve: history of this layout, the rank -> slot map of the
// training sampler, and per-block partials (x.x, dp1.dp1, r.r, count) for
// the scales of the next step, with r = x - p1 - e the extrapolation error
// and count the nodes with h >= 1.
constexpr int kNPart = 4;
__global__ void __launch_bounds__(256) k_history_store(
    int64_t nn, const float* vvol, const int32_t* vrank, const float* x,
    const float* hp1, const float* hdp1, const float* hdp2, const uint8_t* hh,
    float* DP, float* DP2, uint8_t* H, int32_t* slot_of_rank, double* part) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double a[kNPart] = {0, 0, 0, 0};
    if (g < nn) {
        const bool on = vvol[g] > 0.f;
        const int h = hh[g];
        const float xv = x[g];
        const float dp = on && h >= 1 ? xv - hp1[g] : 0.f;
        H[g] = on ? (uint8_t)(h + 1 > 3 ? 3 : h + 1) : (uint8_t)0;
        DP[g] = dp;
        DP2[g] = on && h >= 2 ? hdp1[g] : 0.f;
        if (on) {
            slot_of_rank[vrank[g]] = (int32_t)g;
            a[0] = (double)xv * (double)xv;
            if (h >= 1) {
                const double r = (double)(dp - prior(h, hdp1[g], hdp2[g]));
                a[1] = (double)dp * (double)dp;
                a[2] = r * r;
                a[3] = 1;
            }
        }
    }
    __shared__ double sw[kNPart][4];
#pragma unroll
    for (int j = 0; j < kNPart; j++) {
        a[j] = wave_sum(a[j]);
        if ((threadIdx.x & 63) == 0) sw[j][threadIdx.x >> 6] = a[j];
    }
    __syncthreads();
    if (threadIdx.x < kNPart) {
        const double* w = sw[threadIdx.x];
        part[(int64_t)blockIdx.x * kNPart + threadIdx.x] = (w[0] + w[1]) + (w[2] + w[3]);
    }
}

// Sums the block partials (fixed order) and writes the scales of the next
// step: ps = sqrt(x.x / nv), dps = sqrt(dp.dp / count), tau = sqrt(r.r /
// count), ds = sqrt(b.b / nv).
__global__ void __launch_bounds__(1024) k_scales(const double* part, int nblk,
                                                 const CgState* st, int64_t nv,
                                                 float* scale) {
    __shared__ double sw[kNPart][16];
    double a[kNPart] = {0, 0, 0, 0};
    for (int i = threadIdx.x; i < nblk; i += 1024)
#pragma unroll
        for (int j = 0; j < kNPart; j++) a[j] += part[(int64_t)i * kNPart + j];
#pragma unroll
    for (int j = 0; j < kNPart; j++) {
        a[j] = wave_sum(a[j]);
        if ((threadIdx.x & 63) == 0) sw[j][threadIdx.x >> 6] = a[j];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[kNPart] = {0, 0, 0, 0};
        for (int j = 0; j < kNPart; j++)
            for (int i = 0; i < 16; i++) t[j] += sw[j][i];
        const double inv_nv = nv > 0 ? 1.0 / (double)nv : 0.0;
        const double ps = sqrt(t[0] * inv_nv);
        const double dps = t[3] > 0 ? sqrt(t[1] / t[3]) : 0.0;
        const double tau = t[3] > 0 ? sqrt(t[2] / t[3]) : 0.0;
        const double ds = sqrt(st->bb * inv_nv);
        scale[SC_INV_PS] = ps > 0 ? (float)(1.0 / ps) : 0.f;
        scale[SC_INV_DPS] = dps > 0 ? (float)(1.0 / dps) : 0.f;
        scale[SC_INV_DS] = ds > 0 ? (float)(1.0 / ds) : 0.f;
        scale[SC_INV_TAU] = tau > 0 ? (float)(1.0 / tau) : 0.f;
        scale[SC_TAU] = (float)tau;
        scale[SC_DPS] = (float)dps;
        scale[SC_PS] = (float)ps;
    }
}

// Features of every slot node (read-back for tests).
__global__ void __launch_bounds__(256) k_features(FeatSrc fs, int64_t nn,
                                                  float* out) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= nn) return;
    for (int f = 0; f < kIn; f++) out[g * kIn + f] = feature(fs, g, f);
}

inline unsigned blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

struct MlpState {
    DBuf<float> P, m, v, grad, grad_part;
    DBuf<double> loss_part, loss;
    DBuf<int32_t> skipped;       // Adam steps skipped (non-finite loss or gradient)
    DBuf<float> scale[2];        // SC_* of the current / next step
    int cur_scale = 0;           // scales of the next warm start
    int step_scale = 0;          // scales used by the last warm start
    DBuf<float> DP[2], DP2[2];
    DBuf<uint8_t> H[2];
    DBuf<float> hp1, hdp1, hdp2;
    DBuf<uint8_t> hh;
    DBuf<double> gpart, gy;      // Galerkin start: block partials, weights
    DBuf<int32_t> slot_of_rank;
    DBuf<double> part;
    bool training = false;
    int batch = 65536;
    int steps_per_solve = 1;
    double lr = 1e-3;
    uint64_t seed = 1;
    int64_t adam_t = 0;
    bool have_history = false;   // history arrays valid for last_solve_buf
    int hist_solves = 0;         // consecutive solves with history kept
};

static MlpState* S(vps_context* c) { return (MlpState*)c->mlp; }

static void init_params(std::vector<float>& p, uint64_t seed) {
    p.assign(kNParams, 0.f);
    uint64_t s = seed;
    auto uni = [&](float a) {
        const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
        return (float)((2 * u - 1) * a);
    };
    const float a1 = std::sqrt(6.f / (kIn + kHid)), a2 = std::sqrt(6.f / (2 * kHid));
    for (int i = 0; i < kHid * kIn; i++) p[OFF_W1 + i] = uni(a1);
    for (int i = 0; i < kHid * kHid; i++) p[OFF_W2 + i] = uni(a2);
    // output layer zero: the untrained network predicts "no change"
}

// ---- scratch of vps_mlp_loss_and_grad_device (allocates nothing) ---------
// One call's per-block partial losses (kTrainBlocks doubles) and partial
// gradients (kTrainBlocks x kNParams floats): 4.92 MB on every device the
// library runs on, held until the process ends.  Every call overwrites all
// of it (each training block writes its whole row and its loss), so nothing
// needs clearing; calls that overlap on one device race on it.
constexpr size_t kScratchLossBytes = (size_t)kTrainBlocks * sizeof(double);
constexpr size_t kScratchBytes =
    kScratchLossBytes + (size_t)kTrainBlocks * kNParams * sizeof(float);
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];
static_assert(kScratchLossBytes % 256 == 0, "gradient partials stay 256-byte aligned");

// Device address of `scratch` per device: the host-side name is only a
// stand-in, so kernels get the address hipGetSymbolAddress returns, fetched
// once per device at context creation with that device current.
constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_dev[kScratchDevices] = {};

static int scratch_setup(vps_context* c) {
    if (c->device < 0 || c->device >= kScratchDevices) {
        c->err = "device index beyond the scratch table";
        return VPS_ERR_INVALID_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[c->device]) return VPS_OK;
    void* ptr = nullptr;
    VPS_HIP(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) {
        c->err = "hipGetSymbolAddress(scratch) returned no address";
        return VPS_ERR_HIP;
    }
    g_scratch_dev[c->device] = ptr;
    return VPS_OK;
}

static void* scratch_ptr(int device) {
    if (device < 0 || device >= kScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_dev[device];
}

int mlp_create(vps_context* c) {
    VPS_TRY(scratch_setup(c));
    MlpState* m = new MlpState();
    c->mlp = m;
    VPS_TRY(dalloc(c, m->P, kNParams));
    VPS_TRY(dalloc(c, m->m, kNParams));
    VPS_TRY(dalloc(c, m->v, kNParams));
    VPS_TRY(dalloc(c, m->grad, kNParams));
    VPS_TRY(dalloc(c, m->grad_part, (size_t)kTrainBlocks * kNParams));
    VPS_TRY(dalloc(c, m->loss_part, kTrainBlocks));
    VPS_TRY(dalloc(c, m->loss, 1));
    VPS_TRY(dalloc(c, m->skipped, 1));
    VPS_HIP(c, hipMemset(m->skipped.p, 0, sizeof(int32_t)));
    for (int b = 0; b < 2; b++) {
        VPS_TRY(dalloc(c, m->scale[b], SC_N));
        VPS_HIP(c, hipMemset(m->scale[b].p, 0, SC_N * sizeof(float)));
    }
    VPS_HIP(c, hipMemset(m->loss.p, 0, sizeof(double)));
    std::vector<float> p;
    init_params(p, 0x5EEDull);
    VPS_HIP(c, hipMemcpy(m->P.p, p.data(), kNParams * 4, hipMemcpyHostToDevice));
    VPS_HIP(c, hipMemset(m->m.p, 0, kNParams * 4));
    VPS_HIP(c, hipMemset(m->v.p, 0, kNParams * 4));
    return VPS_OK;
}

// History arrays are allocated on first use (the plain warm starts never
// touch them).
static int ensure_history(vps_context* c) {
    MlpState* m = S(c);
    if (m->hp1.p) return VPS_OK;
    const size_t cap = (size_t)(c->nbricks + 1) * c->brick_nodes;
    for (int b = 0; b < 2; b++) {
        VPS_TRY(dalloc(c, m->DP[b], cap));
        VPS_TRY(dalloc(c, m->DP2[b], cap));
        VPS_TRY(dalloc(c, m->H[b], cap));
        VPS_HIP(c, hipMemset(m->H[b].p, 0, cap));
    }
    VPS_TRY(dalloc(c, m->hp1, cap));
    VPS_TRY(dalloc(c, m->hdp1, cap));
    VPS_TRY(dalloc(c, m->hdp2, cap));
    VPS_TRY(dalloc(c, m->hh, cap));
    VPS_TRY(dalloc(c, m->slot_of_rank, cap));
    VPS_TRY(dalloc(c, m->part, kNPart * (cap / 256 + 16)));
    VPS_TRY(dalloc(c, m->gpart, kNGram * kGramBlocks));
    VPS_TRY(dalloc(c, m->gy, 3));
    return VPS_OK;
}

void mlp_destroy(vps_context* c) {
    MlpState* m = S(c);
    if (!m) return;
    dfree(m->P);
    dfree(m->m);
    dfree(m->v);
    dfree(m->grad);
    dfree(m->grad_part);
    dfree(m->loss_part);
    dfree(m->loss);
    dfree(m->skipped);
    for (int b = 0; b < 2; b++) {
        dfree(m->scale[b]);
        dfree(m->DP[b]);
        dfree(m->DP2[b]);
        dfree(m->H[b]);
    }
    dfree(m->hp1);
    dfree(m->hdp1);
    dfree(m->hdp2);
    dfree(m->hh);
    dfree(m->gpart);
    dfree(m->gy);
    dfree(m->slot_of_rank);
    dfree(m->part);
    delete m;
    c->mlp = nullptr;
}

bool mlp_active(const vps_context* c) {
    const MlpState* m = (const MlpState*)c->mlp;
    return m && (m->training || c->p.warm_start == VPS_WARM_LEARNED ||
                 c->p.warm_start == VPS_WARM_EXTRAPOLATE);
}

static FeatSrc feat_src(vps_context* c, int scale_buf) {
    MlpState* m = S(c);
    FeatSrc f{};
    f.p1 = m->hp1.p;
    f.dp1 = m->hdp1.p;
    f.dp2 = m->hdp2.p;
    f.hh = m->hh.p;
    f.vb = c->vb.p;
    f.vdiag = c->vdiag.p;
    f.vrow = c->vrowsum.p;
    f.vvol = c->vvol.p;
    f.scale = m->scale[scale_buf].p;
    f.inv_a0 = c->alpha0 > 0 ? (float)(1.0 / c->alpha0) : 0.f;
    const double vs = c->p.mass / (c->rho0 > 0 ? c->rho0 : 1.0);
    f.inv_vs = (float)(1.0 / vs);
    return f;
}

// Gathers the node history into the current slot layout, then the warm
// start of the chosen mode.
int mlp_warm_start(vps_context* c, int mode, float* x0) {
    MlpState* m = S(c);
    VPS_TRY(ensure_history(c));
    const int w = c->which;
    const int lb = c->last_solve_buf;
    const int64_t nn = c->nslots * c->brick_nodes;
    const bool hist = m->have_history && lb >= 0;
    if (!hist) m->hist_solves = 0;
    m->step_scale = m->cur_scale;
    if (nn == 0) return warm_start(c, mode, x0);
    hipLaunchKernelGGL(k_history_gather, dim3(blocks(nn, 256)), dim3(256), 0,
                       c->stream, nn, c->brick_nodes, c->slot_brick[w].p,
                       c->vvol.p, hist ? c->brick_slot[lb].p : nullptr,
                       hist ? c->vx[lb].p : nullptr, hist ? m->DP[lb].p : nullptr,
                       hist ? m->DP2[lb].p : nullptr, hist ? m->H[lb].p : nullptr,
                       m->hp1.p, m->hdp1.p, m->hdp2.p, m->hh.p);
    VPS_HIP(c, hipGetLastError());
    if (mode == VPS_WARM_LEARNED) {
        // The Galerkin start over [p1 e c], with the CG's scratch vectors:
        // inputs of M in vp[0], vp[1] and vr2 (their zero brick stays zero),
        // products in vt, vt2 and vr.  The CG overwrites all of them.
        float* P = c->vp[0].p;
        float* E = c->vp[1].p;
        float* Cc = c->vr2.p;
        FeatSrc fs = feat_src(c, m->step_scale);
        fs.corr = Cc;
        // about 8 tiles of 16 nodes per wave
        const int64_t nblk = std::min<int64_t>(std::max<int64_t>((nn + 511) / 512, 1), 8192);
        hipLaunchKernelGGL(k_mlp_infer_pipe, dim3((unsigned)nblk), dim3(256), 0,
                           c->stream, m->P.p, fs, nn, x0);
        hipLaunchKernelGGL(k_galerkin_fields, dim3(blocks(nn, 256)), dim3(256), 0, c->stream,
                           fs, nn, P, E);
        VPS_TRY(apply_operator_f32(c, P, c->vt.p, c->p.op_mode));
        VPS_TRY(apply_operator_f32(c, E, c->vt2.p, c->p.op_mode));
        VPS_TRY(apply_operator_f32(c, Cc, c->vr.p, c->p.op_mode));
        const int nb = (int)blocks(nn, 256);
        const int ng = std::min(nb, kGramBlocks);
        hipLaunchKernelGGL(k_galerkin_gram, dim3(ng), dim3(256), 0, c->stream, nn, P, E, Cc,
                           c->vt.p, c->vt2.p, c->vr.p, c->vb.p, m->gpart.p);
        hipLaunchKernelGGL(k_galerkin_solve, dim3(1), dim3(1024), 0, c->stream, m->gpart.p,
                           ng, m->gy.p);
        hipLaunchKernelGGL(k_galerkin_combine, dim3(nb), dim3(256), 0, c->stream, nn, P, E,
                           Cc, m->gy.p, x0);
    } else if (mode == VPS_WARM_EXTRAPOLATE) {
        hipLaunchKernelGGL(k_extrapolate, dim3(blocks(nn, 256)), dim3(256), 0,
                           c->stream, feat_src(c, m->step_scale), nn, x0);
    } else {
        return warm_start(c, mode, x0);
    }
    VPS_HIP(c, hipGetLastError());
    return VPS_OK;
}

// After a converged solve: node history of this layout, the scales of the
// next step, and (when training) Adam steps on sampled batches.  Everything
// stays on the stream: no host synchronisation.
int mlp_after_solve(vps_context* c, const vps_stats* st) {
    (void)st;
    MlpState* m = S(c);
    const int w = c->which;
    const int64_t nn = c->nslots * c->brick_nodes;
    if (nn == 0) return VPS_OK;
    const int nb = (int)blocks(nn, 256);
    hipLaunchKernelGGL(k_history_store, dim3(nb), dim3(256), 0, c->stream, nn,
                       c->vvol.p, c->vrank.p, c->vx[w].p, m->hp1.p, m->hdp1.p,
                       m->hdp2.p, m->hh.p, m->DP[w].p, m->DP2[w].p, m->H[w].p,
                       m->slot_of_rank.p, m->part.p);
    // targets need dp1 known at the sampled nodes and a dps scale: two
    // earlier solves with history
    if (m->training && m->hist_solves >= 2 && c->nvirt > 0) {
        const int64_t ns = m->batch;
        const int tpw = (int)((ns + 16 * kTrainWaves - 1) / (16 * kTrainWaves));
        for (int it = 0; it < m->steps_per_solve; it++) {
            m->adam_t++;
            const float b1t = (float)std::pow(0.9, (double)m->adam_t);
            const float b2t = (float)std::pow(0.999, (double)m->adam_t);
            hipLaunchKernelGGL(k_mlp_train<false>, dim3(kTrainBlocks), dim3(256),
                               0, c->stream, m->P.p, feat_src(c, m->step_scale), c->vx[w].p,
                               m->slot_of_rank.p, c->nvirt, nullptr, nullptr, ns,
                               tpw, m->seed ^ (uint64_t)m->adam_t * 0x100000001ull,
                               m->grad_part.p, m->loss_part.p);
            hipLaunchKernelGGL(k_mlp_reduce, dim3(blocks(kNParams, 64)), dim3(256),
                               0, c->stream, m->grad_part.p, kTrainBlocks,
                               m->loss_part.p, m->grad.p, m->loss.p);
            hipLaunchKernelGGL(k_mlp_adam, dim3(1), dim3(1024), 0, c->stream, m->grad.p,
                               m->loss.p, m->P.p, m->m.p, m->v.p, (float)m->lr, b1t, b2t,
                               m->skipped.p);
        }
    }
    const int nxt = 1 - m->cur_scale;
    hipLaunchKernelGGL(k_scales, dim3(1), dim3(1024), 0, c->stream, m->part.p, nb,
                       c->cg.p, c->nvirt, m->scale[nxt].p);
    VPS_HIP(c, hipGetLastError());
    m->cur_scale = nxt;
    m->have_history = true;
    m->hist_solves++;
    return VPS_OK;
}

// Checkpoint: parameters, Adam moments and step, and (lb >= 0, history kept)
// the slot-major history of the last solve with the scales of the next step.
int mlp_state_save(vps_context* c, int lb, int64_t nslots, std::vector<char>& out) {
    MlpState* m = S(c);
    out.clear();
    auto put = [&](const void* p, size_t n) {
        out.insert(out.end(), (const char*)p, (const char*)p + n);
    };
    std::vector<float> buf(kNParams);
    for (const DBuf<float>* b : {&m->P, &m->m, &m->v}) {
        VPS_HIP(c, hipMemcpy(buf.data(), b->p, kNParams * 4, hipMemcpyDeviceToHost));
        put(buf.data(), kNParams * 4);
    }
    put(&m->adam_t, 8);
    const int32_t hist = (lb >= 0 && m->have_history && m->hp1.p) ? 1 : 0;
    put(&hist, 4);
    if (hist) {
        const size_t nn = (size_t)nslots * c->brick_nodes;
        std::vector<float> f(nn);
        std::vector<uint8_t> h(nn);
        std::vector<float> sc(SC_N);
        VPS_HIP(c, hipMemcpy(f.data(), m->DP[lb].p, nn * 4, hipMemcpyDeviceToHost));
        put(f.data(), nn * 4);
        VPS_HIP(c, hipMemcpy(f.data(), m->DP2[lb].p, nn * 4, hipMemcpyDeviceToHost));
        put(f.data(), nn * 4);
        VPS_HIP(c, hipMemcpy(h.data(), m->H[lb].p, nn, hipMemcpyDeviceToHost));
        put(h.data(), nn);
        VPS_HIP(c, hipMemcpy(sc.data(), m->scale[m->cur_scale].p, SC_N * 4,
                             hipMemcpyDeviceToHost));
        put(sc.data(), SC_N * 4);
        const int32_t hs = m->hist_solves;
        put(&hs, 4);
    }
    return VPS_OK;
}

int mlp_state_load(vps_context* c, int lb, int64_t nslots, const char* p, size_t n) {
    MlpState* m = S(c);
    const char* end = p + n;
    auto get = [&](void* dst, size_t k) {
        if ((size_t)(end - p) < k) return false;
        std::memcpy(dst, p, k);
        p += k;
        return true;
    };
    std::vector<float> buf(kNParams);
    for (DBuf<float>* b : {&m->P, &m->m, &m->v}) {
        if (!get(buf.data(), kNParams * 4)) {
            c->err = "checkpoint: truncated learned parameters";
            return VPS_ERR_IO;
        }
        VPS_HIP(c, hipMemcpy(b->p, buf.data(), kNParams * 4, hipMemcpyHostToDevice));
    }
    int32_t hist = 0;
    if (!get(&m->adam_t, 8) || !get(&hist, 4)) {
        c->err = "checkpoint: truncated learned state";
        return VPS_ERR_IO;
    }
    m->have_history = false;
    m->hist_solves = 0;
    if (hist && lb >= 0) {
        VPS_TRY(ensure_history(c));
        const size_t nn = (size_t)nslots * c->brick_nodes;
        std::vector<float> f(nn);
        std::vector<uint8_t> h(nn);
        std::vector<float> sc(SC_N);
        int32_t hs = 0;
        bool ok = get(f.data(), nn * 4);
        if (ok) VPS_HIP(c, hipMemcpy(m->DP[lb].p, f.data(), nn * 4, hipMemcpyHostToDevice));
        ok = ok && get(f.data(), nn * 4);
        if (ok) VPS_HIP(c, hipMemcpy(m->DP2[lb].p, f.data(), nn * 4, hipMemcpyHostToDevice));
        ok = ok && get(h.data(), nn) && get(sc.data(), SC_N * 4) && get(&hs, 4);
        if (!ok) {
            c->err = "checkpoint: truncated learned history";
            return VPS_ERR_IO;
        }
        VPS_HIP(c, hipMemcpy(m->H[lb].p, h.data(), nn, hipMemcpyHostToDevice));
        m->cur_scale = 0;
        VPS_HIP(c, hipMemcpy(m->scale[0].p, sc.data(), SC_N * 4, hipMemcpyHostToDevice));
        m->have_history = true;
        m->hist_solves = hs;
    }
    return VPS_OK;
}

}  // namespace vps

using namespace vps;

extern "C" {

int vps_mlp_set_training(vps_context* c, int32_t enable, int32_t batch,
                         double learning_rate, int32_t steps_per_solve,
                         uint64_t seed) {
    if (!c) return VPS_ERR_INVALID_ARG;
    if (batch < 16 || batch > (1 << 22) || !(learning_rate > 0) ||
        steps_per_solve < 1 || steps_per_solve > 64) {
        c->err = "mlp training: batch in [16, 4M], learning_rate > 0, "
                 "steps_per_solve in [1, 64]";
        return VPS_ERR_INVALID_ARG;
    }
    MlpState* m = S(c);
    VPS_TRY(ensure_history(c));
    m->training = enable != 0;
    m->batch = batch;
    m->lr = learning_rate;
    m->steps_per_solve = steps_per_solve;
    m->seed = seed;
    return VPS_OK;
}

int64_t vps_mlp_parameter_count(void) { return kNParams; }

int vps_mlp_get_parameters(vps_context* c, float* out) {
    if (!c || !out) return VPS_ERR_INVALID_ARG;
    VPS_HIP(c, hipMemcpy(out, S(c)->P.p, kNParams * 4, hipMemcpyDeviceToHost));
    return VPS_OK;
}

int vps_mlp_set_parameters(vps_context* c, const float* in) {
    if (!c || !in) return VPS_ERR_INVALID_ARG;
    for (int i = 0; i < kNParams; i++)
        if (!std::isfinite(in[i])) {
            c->err = "mlp parameters: non-finite value";
            return VPS_ERR_NONFINITE;
        }
    MlpState* m = S(c);
    VPS_HIP(c, hipMemcpy(m->P.p, in, kNParams * 4, hipMemcpyHostToDevice));
    return VPS_OK;
}

int vps_mlp_loss_and_grad(vps_context* c, int64_t n, const float* features,
                          const float* targets, double* loss, float* grad) {
    if (!c || n <= 0 || !features || !targets) return VPS_ERR_INVALID_ARG;
    if (n > kMaxLossBatch) {
        c->err = "mlp loss_and_grad: too many samples";
        return VPS_ERR_INVALID_ARG;
    }
    MlpState* m = S(c);
    DBuf<float> df, dt;
    VPS_TRY(dalloc(c, df, (size_t)n * kIn));
    VPS_TRY(dalloc(c, dt, (size_t)n));
    VPS_HIP(c, hipMemcpy(df.p, features, (size_t)n * kIn * 4, hipMemcpyHostToDevice));
    VPS_HIP(c, hipMemcpy(dt.p, targets, (size_t)n * 4, hipMemcpyHostToDevice));
    const int tpw = (int)((n + 16 * kTrainWaves - 1) / (16 * kTrainWaves));
    FeatSrc none{};
    DBuf<double> dl;
    VPS_TRY(dalloc(c, dl, 1));
    hipLaunchKernelGGL(k_mlp_train<true>, dim3(kTrainBlocks), dim3(256), 0,
                       c->stream, m->P.p, none, nullptr, nullptr, n, df.p, dt.p, n,
                       tpw, 0ull, m->grad_part.p, m->loss_part.p);
    hipLaunchKernelGGL(k_mlp_reduce, dim3(blocks(kNParams, 64)), dim3(256), 0,
                       c->stream, m->grad_part.p, kTrainBlocks, m->loss_part.p,
                       m->grad.p, dl.p);
    VPS_HIP(c, hipGetLastError());
    if (grad)
        VPS_HIP(c, hipMemcpyAsync(grad, m->grad.p, kNParams * 4,
                                  hipMemcpyDeviceToHost, c->stream));
    double l = 0;
    VPS_HIP(c, hipMemcpyAsync(c->host_counters + 20, dl.p, 8,
                              hipMemcpyDeviceToHost, c->stream));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    std::memcpy(&l, c->host_counters + 20, 8);
    if (loss) *loss = l;
    dfree(df);
    dfree(dt);
    dfree(dl);
    return VPS_OK;
}

int vps_mlp_loss_and_grad_device(vps_context* c, int64_t n, const float* d_features,
                                 const float* d_targets, double* d_loss, float* d_grad) {
    if (!c) return VPS_ERR_INVALID_ARG;
    if (n <= 0 || !d_features || !d_targets || !d_loss || !d_grad) {
        c->err = "mlp loss_and_grad_device: empty batch or NULL pointer";
        return VPS_ERR_INVALID_ARG;
    }
    if (n > kMaxLossBatch) {
        c->err = "mlp loss_and_grad_device: batch above the maximum of " +
                 std::to_string(kMaxLossBatch) + " samples";
        return VPS_ERR_INVALID_ARG;
    }
    char* base = (char*)scratch_ptr(c->device);
    if (!base) {
        c->err = "mlp loss_and_grad_device: scratch not set up on this device";
        return VPS_ERR_STATE;
    }
    MlpState* m = S(c);
    double* loss_part = (double*)base;
    float* grad_part = (float*)(base + kScratchLossBytes);
    const int tpw = (int)((n + 16 * kTrainWaves - 1) / (16 * kTrainWaves));
    FeatSrc none{};
    hipLaunchKernelGGL(k_mlp_train<true>, dim3(kTrainBlocks), dim3(256), 0, c->stream,
                       m->P.p, none, nullptr, nullptr, n, d_features, d_targets, n, tpw,
                       0ull, grad_part, loss_part);
    hipLaunchKernelGGL(k_mlp_reduce, dim3(blocks(kNParams, 64)), dim3(256), 0, c->stream,
                       grad_part, kTrainBlocks, loss_part, d_grad, d_loss);
    VPS_HIP(c, hipGetLastError());
    return VPS_OK;
}

int vps_mlp_forward(vps_context* c, int64_t n, const float* features,
                    float* out) {
    if (!c || n <= 0 || !features || !out) return VPS_ERR_INVALID_ARG;
    MlpState* m = S(c);
    DBuf<float> df, dout;
    VPS_TRY(dalloc(c, df, (size_t)n * kIn));
    VPS_TRY(dalloc(c, dout, (size_t)n));
    VPS_HIP(c, hipMemcpy(df.p, features, (size_t)n * kIn * 4, hipMemcpyHostToDevice));
    FeatSrc none{};
    hipLaunchKernelGGL(k_mlp_infer<true>, dim3(kInferBlocks), dim3(256), 0,
                       c->stream, m->P.p, none, df.p, n, dout.p);
    VPS_HIP(c, hipGetLastError());
    VPS_HIP(c, hipMemcpyAsync(out, dout.p, (size_t)n * 4, hipMemcpyDeviceToHost,
                              c->stream));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    dfree(df);
    dfree(dout);
    return VPS_OK;
}

int vps_mlp_training_state(vps_context* c, double* last_loss, int64_t* steps) {
    if (!c) return VPS_ERR_INVALID_ARG;
    const MlpState* m = S(c);
    double l = 0;
    int32_t skipped = 0;
    VPS_HIP(c, hipMemcpyAsync(c->host_counters + 20, m->loss.p, 8,
                              hipMemcpyDeviceToHost, c->stream));
    VPS_HIP(c, hipMemcpyAsync(c->host_counters + 21, m->skipped.p, 4,
                              hipMemcpyDeviceToHost, c->stream));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    std::memcpy(&l, c->host_counters + 20, 8);
    std::memcpy(&skipped, c->host_counters + 21, 4);
    if (last_loss) *last_loss = l;
    if (steps) *steps = m->adam_t - skipped;
    return VPS_OK;
}

int vps_mlp_get_features(vps_context* c, float* out) {
    if (!c || !out) return VPS_ERR_INVALID_ARG;
    MlpState* m = S(c);
    if (!m->hp1.p || c->nslots == 0) {
        c->err = "mlp features: no step with history tracking yet";
        return VPS_ERR_STATE;
    }
    const int64_t nn = c->nslots * c->brick_nodes;
    DBuf<float> d;
    VPS_TRY(dalloc(c, d, (size_t)nn * kIn));
    hipLaunchKernelGGL(k_features, dim3(blocks(nn, 256)), dim3(256), 0, c->stream,
                       feat_src(c, m->step_scale), nn, d.p);
    std::vector<float> f((size_t)nn * kIn);
    std::vector<int32_t> rank((size_t)nn);
    VPS_HIP(c, hipMemcpyAsync(f.data(), d.p, (size_t)nn * kIn * 4,
                              hipMemcpyDeviceToHost, c->stream));
    VPS_HIP(c, hipMemcpyAsync(rank.data(), c->vrank.p, (size_t)nn * 4,
                              hipMemcpyDeviceToHost, c->stream));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    for (int64_t g = 0; g < nn; g++)
        if (rank[(size_t)g] >= 0)
            std::memcpy(out + (int64_t)rank[(size_t)g] * kIn, f.data() + g * kIn,
                        kIn * 4);
    dfree(d);
    return VPS_OK;
}

}  // extern "C"
