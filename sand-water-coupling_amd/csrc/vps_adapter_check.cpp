// vps_adapter_check -- drives the projection through the C++ adapter
// (include/vps_virtual_particle.hpp) in the order of
// DualParticleSphLiquid<Dim>::advance (DualParticleSphLiquid.cpp:112-126):
// moveParticles (x += v dt, then the neighbour rebuild and densities),
// external forces (v.y -= g dt), generateParticles, applyPressureForce.
//
// usage: vps_adapter_check <scene.bin> <out.bin>
// scene.bin: int64 dim, n, nb, steps, sizeof(vps_params); vps_params bytes;
//            double dt, gravity; x[n*dim], v[n*dim], bx, bn, bv [nb*dim]
// out.bin:   int64 n_virtual[steps], cg_iterations[steps]; v[n*dim]
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/vps_virtual_particle.hpp"

template <int Dim>
static int run(FILE* f, FILE* o, int64_t n, int64_t nb, int64_t steps,
               const vps_params& p, double dt, double g) {
    auto rd = [&](std::vector<double>& a, size_t cnt) {
        a.resize(cnt);
        return fread(a.data(), 8, cnt, f) == cnt;
    };
    std::vector<double> x, v, bx, bn, bv;
    if (!rd(x, n * Dim) || !rd(v, n * Dim) || !rd(bx, nb * Dim) ||
        !rd(bn, nb * Dim) || !rd(bv, nb * Dim)) {
        fprintf(stderr, "short scene file\n");
        return 2;
    }
    vps::VirtualParticleGpu<Dim> vp(p);
    vp.setBoundary(nb, bx.data(), bn.data(), bv.data());
    vp.initialize(n, x.data(), v.data());
    std::vector<int64_t> nv(steps), it(steps);
    for (int64_t s = 0; s < steps; s++) {
        for (int64_t i = 0; i < n; i++)
            for (int a = 0; a < Dim; a++) x[i * Dim + a] += v[i * Dim + a] * dt;
        for (int64_t i = 0; i < n; i++) v[i * Dim + 1] -= g * dt;
        vp.moveParticles(x.data(), v.data());
        vp.generateParticles();
        const vps_stats st = vp.project(v.data());
        nv[s] = st.n_virtual;
        it[s] = st.cg_iterations;
        printf("step %lld: n_virtual %lld cg_iterations %lld true_rel_residual %.3e\n",
               (long long)s, (long long)st.n_virtual, (long long)st.cg_iterations,
               st.cg_true_rel_residual);
    }
    fwrite(nv.data(), 8, steps, o);
    fwrite(it.data(), 8, steps, o);
    fwrite(v.data(), 8, v.size(), o);
    return 0;
}

int main(int argc, char** argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s scene.bin out.bin\n", argv[0]);
        return 2;
    }
    FILE* f = fopen(argv[1], "rb");
    if (!f) {
        perror(argv[1]);
        return 2;
    }
    int64_t hdr[5];
    vps_params p;
    double dt, g;
    if (fread(hdr, 8, 5, f) != 5 || hdr[4] != (int64_t)sizeof(vps_params) ||
        fread(&p, sizeof p, 1, f) != 1 || fread(&dt, 8, 1, f) != 1 ||
        fread(&g, 8, 1, f) != 1) {
        fprintf(stderr, "bad scene header\n");
        return 2;
    }
    FILE* o = fopen(argv[2], "wb");
    if (!o) {
        perror(argv[2]);
        return 2;
    }
    int rc = 2;
    try {
        if (hdr[0] == 2) rc = run<2>(f, o, hdr[1], hdr[2], hdr[3], p, dt, g);
        else if (hdr[0] == 3) rc = run<3>(f, o, hdr[1], hdr[2], hdr[3], p, dt, g);
    } catch (const vps::Error& e) {
        fprintf(stderr, "vps error %d: %s\n", e.code, e.what());
        rc = 1;
    }
    fclose(f);
    fclose(o);
    return rc;
}
