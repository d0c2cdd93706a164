// vps_sampling.hip -- wall and block sampling on the GPU.
//
// Replaces BoundaryParticles::addSurface (Cores/Structures/SmoothedParticles.
// cpp:111-155) and Shapes::generateBox without Poisson sampling
// (Cores/Utilities/Shapes.h:32-58), with the implicit surfaces they test
// against (Cores/Geometries/ImplicitSurface.h:27-113; ComplementarySurface,
// Surface.h:50-69; Surface::isInside(phi) = phi <= 0, Surface.h:25).
//
// Candidates, in the reference's order:
//   2D  rows j < ys of a hexagonal lattice of spacing 2r, odd rows shifted
//       by half a spacing and one point shorter;
//   3D  the three face grids, then the node grid, of StaggeredGrid(0,
//       2 sqrt(2) r, round((half - r) 2 / spacing), center)
//       (StaggeredGrid.cpp:20-34), each x-fastest (Grid.cpp:231-243).
// One thread per candidate evaluates the position and the surface in FP64
// with every product and sum rounded separately (the reference's order), a
// scan keeps the inside points in candidate order.  The lattice constants
// (row counts, resolutions, grid origins) are formed on the host in the same
// order of operations.
#include <cmath>
#include <vector>

#include "vps_internal.h"
#include "vps_surface.h"

namespace vps {
namespace {

inline unsigned blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

constexpr double kSqrt2 = 1.4142135623730951;     // std::numbers::sqrt2
constexpr double kSqrt3 = 1.7320508075688772;     // std::numbers::sqrt3
constexpr double kInvSqrt3 = 0.5773502691896257;  // std::numbers::inv_sqrt3

struct Lattice {
    int dim;
    double spacing;
    // 2D
    int xs, ys;
    double ox, oy;
    // 3D: grid g = 0..2 the face grids, 3 the node grid
    int64_t size[4][3];
    double org[4][3];
    int64_t start[5];   // first candidate of each grid; start[4] = total
};

double std_round(double x) { return std::round(x); }   // halfway away from zero

Lattice make_lattice(int dim, double r, const double* center, const double* half) {
#pragma clang fp contract(off)
    Lattice L{};
    L.dim = dim;
    if (dim == 2) {
        const double spacing = r * 2;
        L.spacing = spacing;
        L.xs = (int)std_round(half[0] * 2 / spacing);
        L.ys = (int)std_round((half[1] * 2 / spacing - 1) * 2 * kInvSqrt3 + 1);
        L.ox = center[0] - (double)(L.xs - 1) * spacing / 2;
        L.oy = center[1] - ((double)(L.ys - 1) * kSqrt3 / 2) * spacing / 2;
        const int64_t pairs = L.ys / 2;
        L.start[0] = 0;
        L.start[4] = pairs * (2 * (int64_t)L.xs - 1) + (L.ys & 1 ? L.xs : 0);
        if (L.xs <= 0 || L.ys <= 0) L.start[4] = 0;
        return L;
    }
    const double spacing = 2 * kSqrt2 * r;
    L.spacing = spacing;
    int res[3];
    double origin[3];
    for (int a = 0; a < 3; a++) {
        res[a] = (int)std_round((half[a] - r) * 2 / spacing);
        origin[a] = center[a] - (double)res[a] * spacing / 2;
    }
    int64_t acc = 0;
    for (int g = 0; g < 4; g++) {
        L.start[g] = acc;
        int64_t cnt = 1;
        for (int c = 0; c < 3; c++) {
            const int64_t sz = g < 3 ? res[c] + (c == g ? 1 : 0) : res[c] + 1;
            L.size[g][c] = sz < 0 ? 0 : sz;
            L.org[g][c] = g < 3 ? origin[c] + (c == g ? 0.0 : 1.0) * spacing / 2 : origin[c];
            cnt *= L.size[g][c];
        }
        acc += cnt;
    }
    L.start[4] = acc;
    return L;
}

__device__ void candidate(const Lattice& L, int64_t idx, double p[3]) {
#pragma clang fp contract(off)
    if (L.dim == 2) {
        const int64_t pair = 2 * (int64_t)L.xs - 1;
        const int64_t q = idx / pair, rem = idx - q * pair;
        const int j = (int)(2 * q + (rem >= L.xs ? 1 : 0));
        const int i = (int)(rem >= L.xs ? rem - L.xs : rem);
        const double yox = L.ox + (j & 1 ? 0.5 : 0.0) * L.spacing;
        const double yoy = L.oy + ((double)j * kSqrt3 / 2) * L.spacing;
        p[0] = yox + (double)i * L.spacing;
        p[1] = yoy + 0.0;
        p[2] = 0.0;
        return;
    }
    int g = 0;
    while (g < 3 && idx >= L.start[g + 1]) g++;
    const int64_t k = idx - L.start[g];
    const int64_t nx = L.size[g][0], ny = L.size[g][1];
    const int64_t c[3] = {k % nx, (k / nx) % ny, k / nx / ny};
    for (int a = 0; a < 3; a++) p[a] = L.org[g][a] + (double)c[a] * L.spacing;
}

__global__ void __launch_bounds__(256) k_sample(Lattice L, SurfK s, int use_surface,
                                                int32_t* flag, double* pos, double* nrm) {
#pragma clang fp contract(off)
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= L.start[4]) return;
    const int dim = L.dim;
    double p[3], n[3] = {0, 0, 0};
    candidate(L, idx, p);
    bool in = true;
    if (use_surface) {
        double cp[3];
        const double sd = surf_eval(s, dim, p, n, cp);
        if (!s.complement && s.kind == VPS_SURFACE_ELLIPSOID) {
            // ImplicitEllipsoid::isInside: sum (p / semi)^2 - 1 <= 0
            double q[3] = {0, 0, 0};
            for (int a = 0; a < dim; a++) q[a] = p[a] / s.b[a];
            in = (surf_sqnorm(dim, q) - 1) <= 0;
        } else {
            in = sd <= 0;
        }
    }
    flag[idx] = in ? 1 : 0;
    for (int a = 0; a < dim; a++) {
        pos[idx * dim + a] = p[a];
        nrm[idx * dim + a] = n[a];
    }
}

__global__ void __launch_bounds__(256) k_keep(int64_t total, int dim, const int32_t* flag,
                                              const int32_t* slot, const double* pos,
                                              const double* nrm, double* opos, double* onrm) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= total || !flag[idx]) return;
    const int64_t o = slot[idx];
    for (int a = 0; a < dim; a++) {
        opos[o * dim + a] = pos[idx * dim + a];
        onrm[o * dim + a] = nrm[idx * dim + a];
    }
}

// Samples on the GPU; the kept points come back to the host.
int sample(vps_context* c, const Lattice& L, const SurfK* s, std::vector<double>& x,
           std::vector<double>& n) {
    const int D = L.dim;
    const int64_t total = L.start[4];
    x.clear();
    n.clear();
    if (total <= 0) return VPS_OK;
    if (total > ((int64_t)1 << 31) - 1) {
        c->err = "sampling: more than 2^31 candidate points";
        return VPS_ERR_INVALID_ARG;
    }
    DBuf<int32_t> flag, slot;
    DBuf<double> pos, nrm, opos, onrm;
    VPS_TRY(dalloc(c, flag, (size_t)total));
    VPS_TRY(dalloc(c, slot, (size_t)total));
    VPS_TRY(dalloc(c, pos, (size_t)(total * D)));
    VPS_TRY(dalloc(c, nrm, (size_t)(total * D)));
    SurfK sk{};
    if (s) sk = *s;
    hipLaunchKernelGGL(k_sample, dim3(blocks(total, 256)), dim3(256), 0, c->stream, L, sk,
                       s ? 1 : 0, flag.p, pos.p, nrm.p);
    int32_t* dcount = (int32_t*)(c->counters.p + 4);
    VPS_TRY(scan_exclusive_i32(c, flag.p, slot.p, total, dcount));
    VPS_HIP(c, hipMemcpyAsync(c->host_counters + 4, dcount, 4, hipMemcpyDeviceToHost,
                              c->stream));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    const int64_t kept = *(int32_t*)(c->host_counters + 4);
    if (kept > 0) {
        VPS_TRY(dalloc(c, opos, (size_t)(kept * D)));
        VPS_TRY(dalloc(c, onrm, (size_t)(kept * D)));
        hipLaunchKernelGGL(k_keep, dim3(blocks(total, 256)), dim3(256), 0, c->stream, total,
                           D, flag.p, slot.p, pos.p, nrm.p, opos.p, onrm.p);
        VPS_HIP(c, hipGetLastError());
        x.resize((size_t)(kept * D));
        n.resize((size_t)(kept * D));
        VPS_HIP(c, hipMemcpyAsync(x.data(), opos.p, x.size() * 8, hipMemcpyDeviceToHost,
                                  c->stream));
        VPS_HIP(c, hipMemcpyAsync(n.data(), onrm.p, n.size() * 8, hipMemcpyDeviceToHost,
                                  c->stream));
        VPS_HIP(c, hipStreamSynchronize(c->stream));
    }
    dfree(flag);
    dfree(slot);
    dfree(pos);
    dfree(nrm);
    dfree(opos);
    dfree(onrm);
    return VPS_OK;
}

bool finite_vec(const double* v, int n) {
    for (int i = 0; i < n; i++)
        if (!std::isfinite(v[i])) return false;
    return true;
}

int check_region(vps_context* c, double radius, const double* center, const double* half) {
    const int D = c->dim;
    if (!(radius > 0) || !center || !half || !finite_vec(center, D) || !finite_vec(half, D)) {
        c->err = "sampling: radius > 0 and finite center / half lengths";
        return VPS_ERR_INVALID_ARG;
    }
    for (int a = 0; a < D; a++)
        if (!(half[a] > 0)) {
            c->err = "sampling: half lengths must be > 0";
            return VPS_ERR_INVALID_ARG;
        }
    return VPS_OK;
}

}  // namespace

bool make_surfk(const vps_surface& in, int D, SurfK* out, std::string* why) {
#pragma clang fp contract(off)
    if (in.kind < VPS_SURFACE_BOX || in.kind > VPS_SURFACE_ELLIPSOID ||
        !finite_vec(in.a, 3) || !finite_vec(in.b, 3)) {
        *why = "unknown surface kind or non-finite parameters";
        return false;
    }
    SurfK s{};
    s.kind = in.kind;
    s.complement = in.complement ? 1 : 0;
    for (int a = 0; a < 3; a++) {
        s.a[a] = a < D ? in.a[a] : 0.0;
        s.b[a] = a < D ? in.b[a] : 0.0;
    }
    if (s.kind == VPS_SURFACE_PLANE) {   // ImplicitPlane normalises its direction
        double z = s.b[0] * s.b[0] + s.b[1] * s.b[1];
        if (D == 3) z = z + s.b[2] * s.b[2];
        if (!(z > 0)) {
            *why = "zero plane direction";
            return false;
        }
        const double nz = std::sqrt(z);
        for (int a = 0; a < D; a++) s.pn[a] = s.b[a] / nz;
    }
    if (s.kind == VPS_SURFACE_ELLIPSOID)
        for (int a = 0; a < D; a++)
            if (!(s.b[a] != 0)) {
                *why = "zero ellipsoid semi-axis";
                return false;
            }
    *out = s;
    return true;
}

}  // namespace vps

using namespace vps;

extern "C" {

int vps_add_surface(vps_context* c, const vps_surface* surface, const double* center,
                    const double* half_lengths) {
    if (!c || !surface) return VPS_ERR_INVALID_ARG;
    const int D = c->dim;
    const double r = c->p.particle_radius;
    VPS_TRY(check_region(c, r, center, half_lengths));
    SurfK s{};
    std::string why;
    if (!make_surfk(*surface, D, &s, &why)) {
        c->err = "add_surface: " + why;
        return VPS_ERR_INVALID_ARG;
    }
    const Lattice L = make_lattice(D, r, center, half_lengths);
    std::vector<double> x, n;
    VPS_TRY(sample(c, L, &s, x, n));
    // append to the walls (zero velocity) and refresh the resident boundary
    std::vector<double> bx = c->host_bx, bn = c->host_bn, bv = c->host_bv;
    bx.insert(bx.end(), x.begin(), x.end());
    bn.insert(bn.end(), n.begin(), n.end());
    bv.resize(bx.size(), 0.0);
    const int64_t nb = (int64_t)(bx.size() / D);
    return vps_set_boundary(c, nb, bx.data(), bn.data(), bv.data());
}

int vps_generate_box(vps_context* c, const double* center, const double* half_lengths,
                     int64_t capacity, double* x, int64_t* count) {
    if (!c || !count) return VPS_ERR_INVALID_ARG;
    const int D = c->dim;
    VPS_TRY(check_region(c, c->p.particle_radius, center, half_lengths));
    const Lattice L = make_lattice(D, c->p.particle_radius, center, half_lengths);
    *count = L.start[4];
    if (!x) return VPS_OK;
    if (capacity < L.start[4]) {
        c->err = "generate_box: capacity below the point count";
        return VPS_ERR_INVALID_ARG;
    }
    std::vector<double> px, pn;
    VPS_TRY(sample(c, L, nullptr, px, pn));
    std::copy(px.begin(), px.end(), x);
    return VPS_OK;
}

int vps_get_boundary(vps_context* c, int64_t* count, double* x, double* normals, double* v) {
    if (!c || !count) return VPS_ERR_INVALID_ARG;
    const size_t n = c->host_bx.size();
    *count = (int64_t)(n / c->dim);
    if (x) std::copy(c->host_bx.begin(), c->host_bx.end(), x);
    if (normals) std::copy(c->host_bn.begin(), c->host_bn.end(), normals);
    if (v) std::copy(c->host_bv.begin(), c->host_bv.end(), v);
    return VPS_OK;
}

}  // extern "C"
