// vps_virtual_particle.hpp -- C++ adapter over the C ABI (vps.h) with the
// method names of the reference's projection interface, so that
// DualParticleSphLiquid<Dim> can hold a VirtualParticleGpu<Dim> where it
// holds a VirtualParticle<Dim>:
//
//   VirtualParticle<Dim>::generateParticles / setAlpha / setKappa / project
//       (Cores/Structures/VirtualParticle.h:76-87)
//   DualParticleSphLiquid<Dim>::initialize / advance / applyPressureForce
//       (Cores/Physics/DualParticleSphLiquid.cpp:78-126)
//
// Vectors are taken as `const double*` with stride Dim: the memory layout of
// std::vector<Eigen::Matrix<double, Dim, 1>> (Dim contiguous doubles per
// particle), so the reference passes `positions.data()->data()` without a
// copy.  No Eigen or torch types appear here.  Errors throw vps::Error with
// the library's status code and message (the reference calls std::exit).
#ifndef VPS_VIRTUAL_PARTICLE_HPP
#define VPS_VIRTUAL_PARTICLE_HPP

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "vps.h"

namespace vps {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what), code(code) {}
    int code;
};

inline void check(int rc, const vps_context* ctx, const char* where) {
    if (rc != VPS_OK)
        throw Error(rc, std::string(where) + ": " + vps_last_error(ctx));
}

template <int Dim>
class VirtualParticleGpu {
    static_assert(Dim == 2 || Dim == 3, "Dim is 2 or 3");

public:
    // The lattice of StaggeredGrid<Dim>::nodeGrid() and the particle radius
    // of DualParticleSphLiquid(grid, particleRadius) (DualParticleSphLiquid.h:27-29).
    explicit VirtualParticleGpu(const vps_params& params) {
        vps_params p = params;
        p.dim = Dim;
        check(vps_create(&p, &ctx_), nullptr, "vps_create");
    }
    ~VirtualParticleGpu() { vps_destroy(ctx_); }
    VirtualParticleGpu(const VirtualParticleGpu&) = delete;
    VirtualParticleGpu& operator=(const VirtualParticleGpu&) = delete;

    // The wall particles of BoundaryParticles::addSurface and the boundary
    // velocity field (DualParticleSphLiquid.cpp:79-84), sampled by the caller.
    void setBoundary(std::size_t n, const double* x, const double* normals,
                     const double* v) {
        check(vps_set_boundary(ctx_, (int64_t)n, x, normals, v), ctx_, "vps_set_boundary");
    }

    // DualParticleSphLiquid::initialize after the wall sampling
    // (DualParticleSphLiquid.cpp:86-98): real densities and frozen volumes,
    // rest density, boundary densities, setAlpha, setKappa, first
    // generateParticles.
    vps_stats initialize(std::size_t n, const double* x, const double* v) {
        check(vps_set_particles(ctx_, (int64_t)n, x, v), ctx_, "vps_set_particles");
        vps_stats st{};
        check(vps_initialize(ctx_, &st), ctx_, "vps_initialize");
        n_ = n;
        return st;
    }

    // The neighbour rebuild and density refresh at the end of
    // SmthParticleHydrodLiquid::moveParticles (SmthParticleHydrodLiquid.cpp:80-81),
    // with the caller's moved positions and current velocities.
    void moveParticles(const double* x, const double* v) {
        check(vps_set_particles(ctx_, (int64_t)n_, x, v), ctx_, "vps_set_particles");
        check(vps_refresh_real(ctx_), ctx_, "vps_refresh_real");
    }

    // VirtualParticle::generateParticles (VirtualParticle.cpp:45-62).
    void generateParticles() {
        check(vps_generate_particles(ctx_), ctx_, "vps_generate_particles");
    }

    // VirtualParticle::project (VirtualParticle.cpp:75-86): right-hand side,
    // Laplacian, CG and the pressure-gradient update, written back into v
    // (n * Dim doubles, the caller's particle order).
    vps_stats project(double* v) {
        vps_stats st{};
        const int rc = vps_project(ctx_, &st);
        check(rc, ctx_, "vps_project");
        scratch_.resize(n_ * Dim);
        check(vps_get_particles(ctx_, scratch_.data(), v), ctx_, "vps_get_particles");
        return st;
    }

    // Number of virtual particles of the last generation
    // (ParticlesBasedData::size of the virtual set).
    std::size_t size() const {
        int64_t nv = 0;
        check(vps_get_virtual_count(ctx_, &nv), ctx_, "vps_get_virtual_count");
        return (std::size_t)nv;
    }

    vps_context* handle() const { return ctx_; }

private:
    vps_context* ctx_ = nullptr;
    std::size_t n_ = 0;
    std::vector<double> scratch_;
};

// DualParticleSphLiquid<Dim> with its step on the GPU
// (Cores/Physics/DualParticleSphLiquid.h:13-60): walls from addSurface,
// particles from the caller or generateBox, advance() the whole step
// (DualParticleSphLiquid.cpp:107-116), frames in the reference's formats.
template <int Dim>
class DualParticleSphLiquidGpu : public VirtualParticleGpu<Dim> {
    using Base = VirtualParticleGpu<Dim>;

public:
    explicit DualParticleSphLiquidGpu(const vps_params& params) : Base(params) {}

    // BoundaryParticles::addSurface (SmoothedParticles.cpp:111-155)
    void addSurface(const vps_surface& s, const double* center, const double* halfLengths) {
        check(vps_add_surface(this->handle(), &s, center, halfLengths), this->handle(),
              "vps_add_surface");
    }

    // Shapes::generateBox without Poisson sampling (Shapes.h:32-58)
    std::vector<double> generateBox(const double* center, const double* halfLengths) {
        int64_t n = 0;
        check(vps_generate_box(this->handle(), center, halfLengths, 0, nullptr, &n),
              this->handle(), "vps_generate_box");
        std::vector<double> x((std::size_t)n * Dim);
        check(vps_generate_box(this->handle(), center, halfLengths, n, x.data(), &n),
              this->handle(), "vps_generate_box");
        return x;
    }

    // DualParticleSphLiquid::advance: move + colliders, gravity, viscosity,
    // generateParticles, project
    vps_stats advance(const vps_step_params& step) {
        vps_stats st{};
        check(vps_advance(this->handle(), &step, &st), this->handle(), "vps_advance");
        return st;
    }

    // writeFrame / saveFrame / loadFrame (DualParticleSphLiquid.cpp:41-76;
    // SmthParticleHydrodLiquid.cpp:34-55)
    void writeFrame(const std::string& dir) {
        check(vps_write_frame(this->handle(), dir.c_str()), this->handle(), "vps_write_frame");
    }
    void saveFrame(const std::string& dir) {
        check(vps_save_frame(this->handle(), dir.c_str()), this->handle(), "vps_save_frame");
    }
    void loadFrame(const std::string& dir) {
        check(vps_load_frame(this->handle(), dir.c_str()), this->handle(), "vps_load_frame");
    }
};

// DEMParticleSand<Dim> (Cores/Physics/DEMParticleSand.h:13-58): the grains
// are the context's particles, the walls its boundary.
template <int Dim>
class DEMParticleSandGpu : public VirtualParticleGpu<Dim> {
    using Base = VirtualParticleGpu<Dim>;

public:
    DEMParticleSandGpu(const vps_params& params, const vps_dem_params& dem)
        : Base(params), dem_(dem) {}

    void setParticles(std::size_t n, const double* x, const double* v) {
        check(vps_set_particles(this->handle(), (int64_t)n, x, v), this->handle(),
              "vps_set_particles");
    }

    // DEMParticleSand::advance (DEMParticleSand.cpp:90-152)
    void advance(double dt, double gravity = 9.8) {
        check(vps_dem_step(this->handle(), dt, gravity, &dem_), this->handle(), "vps_dem_step");
    }

    // the grains as moving walls of a liquid context
    void coupleInto(VirtualParticleGpu<Dim>& water) {
        check(vps_couple_grains(water.handle(), this->handle()), water.handle(),
              "vps_couple_grains");
    }

private:
    vps_dem_params dem_;
};

}  // namespace vps

#endif  // VPS_VIRTUAL_PARTICLE_HPP
