// vps_surface.h -- implicit surfaces on the device
// (Cores/Geometries/ImplicitSurface.h:27-113; ComplementarySurface,
// Surface.h:50-69): signed distance, closest normal and closest position of
// a box, sphere, plane or ellipsoid in FP64 with every product and sum
// rounded separately, as the reference evaluates them.
#ifndef VPS_SURFACE_H
#define VPS_SURFACE_H

#include "vps_internal.h"

namespace vps {

struct SurfK {
    int kind, complement;
    double a[3], b[3];
    double pn[3];   // plane: normalised direction
};

__device__ inline double surf_sqnorm(int dim, const double* v) {
#pragma clang fp contract(off)
    double s = v[0] * v[0] + v[1] * v[1];
    if (dim == 3) s = s + v[2] * v[2];
    return s;
}

__device__ inline void surf_normalize(int dim, double* v) {
#pragma clang fp contract(off)
    const double z = surf_sqnorm(dim, v);
    if (z > 0) {
        const double n = sqrt(z);
        for (int a = 0; a < dim; a++) v[a] = v[a] / n;
    }
}

// signed distance and closest normal of the surface itself
__device__ inline double surf_inner(const SurfK& s, int dim, const double* p, double* n,
                                  double* cp) {
#pragma clang fp contract(off)
    n[0] = n[1] = n[2] = 0.0;
    if (s.kind == VPS_SURFACE_BOX) {
        double d[3], phi[3];
        bool all_in = true;
        int ax = 0;
        for (int a = 0; a < dim; a++) {
            const double c = s.a[a] + s.b[a] / 2, h = s.b[a] / 2;
            d[a] = p[a] - c;
            phi[a] = fabs(d[a]) - h;
            all_in = all_in && phi[a] <= 0;
            if (a > 0 && phi[a] > phi[ax]) ax = a;   // first maximum
        }
        double sd;
        if (all_in) {
            sd = phi[ax];
            n[ax] = 1.0;
        } else {
            double m[3] = {0, 0, 0};
            for (int a = 0; a < dim; a++) m[a] = phi[a] > 0 ? phi[a] : 0.0;
            sd = sqrt(surf_sqnorm(dim, m));
            for (int a = 0; a < dim; a++) n[a] = m[a];
        }
        for (int a = 0; a < dim; a++) {
            const double sg = d[a] > 0 ? 1.0 : (d[a] < 0 ? -1.0 : 0.0);
            n[a] = n[a] * sg;
        }
        surf_normalize(dim, n);
        // ImplicitSurface::closestPosition: pos - signedDistance * normal
        for (int a = 0; a < dim; a++) cp[a] = p[a] - sd * n[a];
        return sd;
    }
    if (s.kind == VPS_SURFACE_SPHERE) {
        for (int a = 0; a < dim; a++) n[a] = p[a] - s.a[a];
        const double sd = sqrt(surf_sqnorm(dim, n)) - s.b[0];
        surf_normalize(dim, n);
        // ImplicitSphere::closestPosition: center + normal * radius
        for (int a = 0; a < dim; a++) cp[a] = s.a[a] + n[a] * s.b[0];
        return sd;
    }
    if (s.kind == VPS_SURFACE_PLANE) {
        double d[3];
        for (int a = 0; a < dim; a++) {
            d[a] = p[a] - s.a[a];
            n[a] = s.pn[a];
        }
        double sd = d[0] * s.pn[0] + d[1] * s.pn[1];
        if (dim == 3) sd = sd + d[2] * s.pn[2];
        for (int a = 0; a < dim; a++) cp[a] = p[a] - sd * n[a];
        return sd;
    }
    // ellipsoid (the reference measures from the origin, not its center)
    double q[3] = {0, 0, 0};
    for (int a = 0; a < dim; a++) q[a] = p[a] / s.b[a];
    const double s2 = surf_sqnorm(dim, q);
    const bool in = (s2 - 1) <= 0;
    const double sq = sqrt(s2);
    for (int a = 0; a < dim; a++) {
        cp[a] = p[a] / sq;   // ImplicitEllipsoid::closestPosition
        n[a] = p[a] - cp[a];
    }
    const double dist = sqrt(surf_sqnorm(dim, n));
    surf_normalize(dim, n);
    const double sg = in ? -1.0 : 1.0;
    for (int a = 0; a < dim; a++) n[a] = n[a] * sg;
    return dist * sg;
}

// Surface as sampled or collided against: the complement flips the signed
// distance and the normal and keeps the closest position.
__device__ inline double surf_eval(const SurfK& s, int dim, const double* p, double* n,
                                   double* cp) {
#pragma clang fp contract(off)
    double sd = surf_inner(s, dim, p, n, cp);
    if (s.complement) {
        sd = -sd;
        for (int a = 0; a < dim; a++) n[a] = -n[a];
    }
    return sd;
}
// Host side: a vps_surface with the plane direction normalised
// (ImplicitPlane's constructor); false for an unusable surface.
bool make_surfk(const vps_surface& in, int dim, SurfK* out, std::string* why);

}  // namespace vps

#endif  // VPS_SURFACE_H
