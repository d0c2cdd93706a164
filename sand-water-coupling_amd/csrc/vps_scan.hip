// vps_scan.hip -- device-wide exclusive scan and max reduction.
//
// Used for the cell counting sort (cell_start), brick compaction (slots) and
// the lattice-bitmask ranks.  Three passes: per-block totals, one-block scan
// of the totals, per-block scan plus offset.  All integer, so the result is
// exact and independent of scheduling.
#include "vps_internal.h"

namespace vps {
namespace {

constexpr int kScanBS = 256;
constexpr int kScanIPT = 4;
constexpr int kScanTile = kScanBS * kScanIPT;

__device__ __forceinline__ int block_exclusive_scan(int v, int* total) {
    __shared__ int swave[kScanBS / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // inclusive wave scan
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) swave[wv] = x;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kScanBS / 64; i++) {
        if (i < wv) off += swave[i];
        tot += swave[i];
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
}

__global__ void __launch_bounds__(kScanBS) k_tile_sums(const int32_t* in,
                                                       int64_t n,
                                                       int32_t* sums) {
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanIPT; k++) {
        const int64_t i = base + (int64_t)k * kScanBS + threadIdx.x;
        if (i < n) s += in[i];
    }
    s = wave_sum(s);
    __shared__ int sw[kScanBS / 64];
    if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int i = 0; i < kScanBS / 64; i++) t += sw[i];
        sums[blockIdx.x] = t;
    }
}

// One block scans the tile totals in place; writes the grand total.
__global__ void __launch_bounds__(1024) k_scan_sums(int32_t* sums, int nt,
                                                     int32_t* total) {
    __shared__ int sw[16];
    const int per = (nt + 1023) / 1024;
    const int b0 = threadIdx.x * per;
    int local = 0;
    for (int i = 0; i < per; i++)
        if (b0 + i < nt) local += sums[b0 + i];
    // block exclusive scan of 'local' over 1024 threads
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int x = local;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sw[wv] = x;
    __syncthreads();
    int off = 0, tot = 0;
    for (int i = 0; i < 16; i++) {
        if (i < wv) off += sw[i];
        tot += sw[i];
    }
    int run = off + x - local;
    for (int i = 0; i < per; i++)
        if (b0 + i < nt) {
            const int v = sums[b0 + i];
            sums[b0 + i] = run;
            run += v;
        }
    if (threadIdx.x == 0 && total) *total = tot;
}

// clear != nullptr (the input itself, not aliased with out): the input is
// zeroed once read, ready for the next counting pass.
__global__ void __launch_bounds__(kScanBS) k_tile_scan(const int32_t* in,
                                                       int64_t n,
                                                       const int32_t* sums,
                                                       int32_t* out, int32_t* clear) {
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    // each thread owns kScanIPT consecutive elements
    int v[kScanIPT];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanIPT; k++) {
        const int64_t i = base + (int64_t)threadIdx.x * kScanIPT + k;
        v[k] = i < n ? in[i] : 0;
        s += v[k];
    }
    if (clear) {
#pragma unroll
        for (int k = 0; k < kScanIPT; k++) {
            const int64_t i = base + (int64_t)threadIdx.x * kScanIPT + k;
            if (i < n) clear[i] = 0;
        }
    }
    int tot;
    int run = block_exclusive_scan(s, &tot) + sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanIPT; k++) {
        const int64_t i = base + (int64_t)threadIdx.x * kScanIPT + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
}

__global__ void __launch_bounds__(256) k_max_f32(const float* in, int64_t n,
                                                 double* partials,
                                                 unsigned* counter,
                                                 double* out) {
    float m = -3.4e38f;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256)
        m = fmaxf(m, in[i]);
    m = wave_max(m);
    __shared__ float sw[4];
    __shared__ bool last;
    if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = sw[0];
        for (int i = 1; i < 4; i++) t = fmaxf(t, sw[i]);
        partials[blockIdx.x] = t;
        __threadfence();
        last = atomicAdd(counter, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (last && threadIdx.x == 0) {
        __threadfence();
        double t = -3.4e38;
        for (unsigned i = 0; i < gridDim.x; i++) {
            const double v = __hip_atomic_load(&partials[i], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
            t = v > t ? v : t;
        }
        *out = t;
        *counter = 0;
    }
}

}  // namespace

int scan_exclusive_i32(vps_context* c, const int32_t* in, int32_t* out,
                       int64_t n, int32_t* total_dev, int32_t* clear_in) {
    if (n <= 0) {
        if (total_dev) VPS_HIP(c, hipMemsetAsync(total_dev, 0, 4, c->stream));
        return VPS_OK;
    }
    const int64_t nt = (n + kScanTile - 1) / kScanTile;
    VPS_TRY(dalloc(c, c->scan_tmp, (size_t)nt));
    int32_t* sums = (int32_t*)c->scan_tmp.p;
    hipLaunchKernelGGL(k_tile_sums, dim3((unsigned)nt), dim3(kScanBS), 0,
                       c->stream, in, n, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, c->stream, sums,
                       (int)nt, total_dev);
    hipLaunchKernelGGL(k_tile_scan, dim3((unsigned)nt), dim3(kScanBS), 0,
                       c->stream, in, n, sums, out, clear_in);
    VPS_HIP(c, hipGetLastError());
    return VPS_OK;
}

int reduce_max_f32(vps_context* c, const float* in, int64_t n, double* out) {
    int64_t nb = (n + 255) / 256;
    if (nb > 2048) nb = 2048;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_max_f32, dim3((unsigned)nb), dim3(256), 0, c->stream,
                       in, n, c->partials.p, (unsigned*)&c->cg.p->counter[3],
                       out);
    VPS_HIP(c, hipGetLastError());
    return VPS_OK;
}

}  // namespace vps
