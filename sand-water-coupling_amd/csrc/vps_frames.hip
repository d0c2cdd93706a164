// vps_frames.hip -- frame output in the reference's file formats and a
// checkpoint of the projection state (host code).
//
// Formats, as the reference writes them:
//   <frame>/particles.mesh, virtual_particles.mesh, boundary_particles.mesh
//       (DualParticleSphLiquid::writeFrame, DualParticleSphLiquid.cpp:41-76):
//       uint32 count; count x float[Dim] positions; in 3D count x float[3]
//       normals (0, 0, 1); then count floats: |v| for the fluid, V_I for the
//       virtual particles, nothing for the boundary.
//   <frame>/particles.sav, velocities.sav (SmthParticleHydrodLiquid::saveFrame,
//       SmthParticleHydrodLiquid.cpp:34-43; ParticlesAttribute::save,
//       ParticlesAttribute.h:89-90): the raw FP64 Dim-vectors, no header.
//   <out>/description.yaml (Simulator.cpp:33-39 with writeDescription,
//       DualParticleSphLiquid.cpp:8-39) and <out>/end_frame.txt = frame + 1
//       (Simulator.cpp:80-89).
// The reference does not checkpoint the projection: its loadFrame
// (SmthParticleHydrodLiquid.cpp:45-55) recomputes the volumes from the
// loaded positions and restarts the pressures.  <frame>/projection.sav holds
// what a faithful resume needs: the frozen real volumes V_j, rho0, kappa,
// alpha0, the last solve's slot layout and pressures (both warm starts that
// carry pressures read them) and the learned warm start's parameters,
// optimiser moments and node history.  Loading a frame with it continues
// the run exactly; without it the load re-initialises as the reference does.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <string>
#include <vector>

#include "vps_internal.h"

namespace vps {
namespace {

// version 2 adds the 3D CG's measured FP32 drift (it sets the recursion's
// stopping threshold, so a resumed run needs it to stop where the original
// run stops); version 1 files load with the default drift
constexpr char kMagic[8] = {'V', 'P', 'S', 'P', 'R', 'O', 'J', '2'};
constexpr char kMagicV1[8] = {'V', 'P', 'S', 'P', 'R', 'O', 'J', '1'};

struct Blob {
    std::vector<char> b;
    template <typename T>
    void put(const T& v) {
        const char* p = (const char*)&v;
        b.insert(b.end(), p, p + sizeof(T));
    }
    template <typename T>
    void put_array(const T* p, size_t n) {
        const char* q = (const char*)p;
        b.insert(b.end(), q, q + n * sizeof(T));
    }
};

struct Reader {
    const char* p;
    const char* end;
    template <typename T>
    bool get(T& v) {
        if (end - p < (ptrdiff_t)sizeof(T)) return false;
        std::memcpy(&v, p, sizeof(T));
        p += sizeof(T);
        return true;
    }
    template <typename T>
    bool get_array(T* out, size_t n) {
        if ((size_t)(end - p) < n * sizeof(T)) return false;
        std::memcpy(out, p, n * sizeof(T));
        p += n * sizeof(T);
        return true;
    }
};

bool write_bytes(const std::string& path, const void* data, size_t n) {
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    f.write((const char*)data, (std::streamsize)n);
    return (bool)f;
}

bool read_bytes(const std::string& path, std::vector<char>& out) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return false;
    const std::streamsize n = f.tellg();
    if (n < 0) return false;
    out.resize((size_t)n);
    f.seekg(0);
    return n == 0 || (bool)f.read(out.data(), n);
}

// uint count; float pos[Dim] each; 3D: float normal (0,0,1) each; scalar each
bool write_mesh(const std::string& path, int dim, const std::vector<float>& pos,
                const std::vector<float>* scalar) {
    const uint32_t n = (uint32_t)(pos.size() / dim);
    Blob o;
    o.put(n);
    o.put_array(pos.data(), pos.size());
    if (dim == 3) {
        const float unit_z[3] = {0.f, 0.f, 1.f};
        for (uint32_t i = 0; i < n; i++) o.put_array(unit_z, 3);
    }
    if (scalar) o.put_array(scalar->data(), scalar->size());
    return write_bytes(path, o.b.data(), o.b.size());
}

int fail(vps_context* c, int code, const std::string& what) {
    c->err = what;
    return code;
}
int fail_io(vps_context* c, const std::string& what) { return fail(c, VPS_ERR_IO, what); }

template <typename T>
int pull(vps_context* c, std::vector<T>& out, const T* src, size_t n) {
    out.resize(n);
    if (n == 0) return VPS_OK;
    VPS_HIP(c, hipMemcpyAsync(out.data(), src, n * sizeof(T), hipMemcpyDeviceToHost,
                              c->stream));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    return VPS_OK;
}

template <typename T>
int push(vps_context* c, T* dst, const std::vector<T>& in) {
    if (in.empty()) return VPS_OK;
    VPS_HIP(c, hipMemcpyAsync(dst, in.data(), in.size() * sizeof(T), hipMemcpyHostToDevice,
                              c->stream));
    VPS_HIP(c, hipStreamSynchronize(c->stream));
    return VPS_OK;
}

}  // namespace

int write_frame(vps_context* c, const std::string& dir) {
    const int D = c->dim;
    const int64_t n = c->n;
    std::vector<double> x((size_t)(n * D)), v((size_t)(n * D));
    VPS_TRY(vps_get_particles(c, x.data(), v.data()));
    {   // fluid particles: position and |v| (FP64 norm, then FP32)
        std::vector<float> pos((size_t)(n * D)), speed((size_t)n);
        for (int64_t i = 0; i < n; i++) {
            double s = 0;
            for (int a = 0; a < D; a++) {
                pos[(size_t)(D * i + a)] = (float)x[(size_t)(D * i + a)];
                s += v[(size_t)(D * i + a)] * v[(size_t)(D * i + a)];
            }
            speed[(size_t)i] = (float)std::sqrt(s);
        }
        if (!write_mesh(dir + "/particles.mesh", D, pos, &speed))
            return fail_io(c, "cannot write " + dir + "/particles.mesh");
    }
    {   // virtual particles of the last generation: node position and V_I
        const int64_t nv = c->nvirt;
        std::vector<int64_t> node((size_t)nv);
        std::vector<double> vol((size_t)nv);
        if (nv > 0) VPS_TRY(vps_get_virtual(c, node.data(), nullptr, vol.data(), nullptr,
                                            nullptr, nullptr));
        std::vector<float> pos((size_t)(nv * D)), sv((size_t)nv);
        const KConst& k = c->k;
        for (int64_t I = 0; I < nv; I++) {
            const int64_t lin = node[(size_t)I];
            const int64_t cc[3] = {lin % k.ndims[0], (lin / k.ndims[0]) % k.ndims[1],
                                   lin / k.ndims[0] / k.ndims[1]};
            for (int a = 0; a < D; a++)   // Grid::dataPosition: origin + coord * spacing
                pos[(size_t)(D * I + a)] = (float)(k.origin[a] + (double)cc[a] * k.spacing);
            sv[(size_t)I] = (float)vol[(size_t)I];
        }
        if (!write_mesh(dir + "/virtual_particles.mesh", D, pos, &sv))
            return fail_io(c, "cannot write " + dir + "/virtual_particles.mesh");
    }
    {   // boundary particles (walls and any coupled grains), in input order
        const int64_t nb = c->nb;
        std::vector<double4> bp;
        std::vector<int32_t> bid;
        VPS_TRY(pull(c, bp, c->bpos.p, (size_t)nb));
        VPS_TRY(pull(c, bid, c->bid.p, (size_t)nb));
        std::vector<float> pos((size_t)(nb * D));
        for (int64_t k2 = 0; k2 < nb; k2++) {
            const int64_t i = bid[(size_t)k2];
            const double pc[3] = {bp[(size_t)k2].x, bp[(size_t)k2].y, bp[(size_t)k2].z};
            for (int a = 0; a < D; a++) pos[(size_t)(D * i + a)] = (float)pc[a];
        }
        if (!write_mesh(dir + "/boundary_particles.mesh", D, pos, nullptr))
            return fail_io(c, "cannot write " + dir + "/boundary_particles.mesh");
    }
    return VPS_OK;
}

int save_frame(vps_context* c, const std::string& dir) {
    const int D = c->dim;
    const int64_t n = c->n;
    std::vector<double> x((size_t)(n * D)), v((size_t)(n * D));
    VPS_TRY(vps_get_particles(c, x.data(), v.data()));
    if (!write_bytes(dir + "/particles.sav", x.data(), x.size() * 8) ||
        !write_bytes(dir + "/velocities.sav", v.data(), v.size() * 8))
        return fail_io(c, "cannot write " + dir + "/particles.sav / velocities.sav");
    Blob o;
    o.put_array(kMagic, 8);
    o.put((int32_t)D);
    o.put((int32_t)c->p.warm_start);
    o.put((int32_t)(c->initialized ? 1 : 0));
    o.put(n);
    o.put(c->rho0);
    o.put(c->kappa);
    o.put(c->alpha0);
    o.put(c->cg_drift);
    {   // frozen real volumes, caller order
        std::vector<double> vol((size_t)n);
        VPS_TRY(vps_get_real_fields(c, nullptr, vol.data()));
        std::vector<float> vf((size_t)n);
        for (int64_t i = 0; i < n; i++) vf[(size_t)i] = (float)vol[(size_t)i];
        o.put_array(vf.data(), vf.size());
    }
    // the last solve: its slot map and slot-major pressures, available while
    // the current slot map is the one that was solved on
    const int lb = c->last_solve_buf;
    const bool last = c->initialized && lb >= 0 && lb == c->which && c->nslots > 0;
    o.put((int32_t)(last ? 1 : 0));
    if (last) {
        const int64_t ns = c->nslots, nn = ns * c->brick_nodes;
        std::vector<int32_t> sb;
        std::vector<float> xs, pr;
        VPS_TRY(pull(c, sb, c->slot_brick[lb].p, (size_t)ns));
        VPS_TRY(pull(c, xs, c->vx[lb].p, (size_t)nn));
        VPS_TRY(pull(c, pr, c->p_by_rank[lb].p, (size_t)c->last_solve_nvirt));
        o.put(ns);
        o.put(c->last_solve_nvirt);
        o.put((int32_t)c->prev_iters);
        o.put_array(sb.data(), sb.size());
        o.put_array(xs.data(), xs.size());
        o.put_array(pr.data(), pr.size());
    }
    std::vector<char> mlp;
    VPS_TRY(mlp_state_save(c, last ? lb : -1, last ? c->nslots : 0, mlp));
    o.put((int64_t)mlp.size());
    o.put_array(mlp.data(), mlp.size());
    if (!write_bytes(dir + "/projection.sav", o.b.data(), o.b.size()))
        return fail_io(c, "cannot write " + dir + "/projection.sav");
    return VPS_OK;
}

int load_frame(vps_context* c, const std::string& dir) {
    const int D = c->dim;
    std::vector<char> xb, vb;
    if (!read_bytes(dir + "/particles.sav", xb) || !read_bytes(dir + "/velocities.sav", vb))
        return fail_io(c, "cannot read " + dir + "/particles.sav / velocities.sav");
    const size_t row = (size_t)D * 8;
    if (xb.empty() || xb.size() % row != 0 || vb.size() != xb.size())
        return fail_io(c, "particles.sav / velocities.sav: sizes do not match dim " +
                              std::to_string(D));
    const int64_t n = (int64_t)(xb.size() / row);
    std::vector<char> sb;
    const bool have_state = read_bytes(dir + "/projection.sav", sb);
    if (!have_state) {
        // the reference's loadFrame: new positions and velocities, everything
        // else recomputed
        c->initialized = false;
        VPS_TRY(vps_set_particles(c, n, (const double*)xb.data(), (const double*)vb.data()));
        return vps_initialize(c, nullptr);
    }
    Reader r{sb.data(), sb.data() + sb.size()};
    char magic[8];
    int32_t dim = 0, warm = 0, init = 0;
    int64_t ns_state = 0;
    double rho0 = 0, kappa = 0, alpha0 = 0, drift = 1.0;
    if (!r.get_array(magic, 8))
        return fail_io(c, dir + "/projection.sav: not a projection checkpoint");
    const bool v2 = std::memcmp(magic, kMagic, 8) == 0;
    if ((!v2 && std::memcmp(magic, kMagicV1, 8) != 0) || !r.get(dim) || !r.get(warm) ||
        !r.get(init) || !r.get(ns_state) || !r.get(rho0) || !r.get(kappa) ||
        !r.get(alpha0) || (v2 && !r.get(drift)) || !std::isfinite(drift) || !(drift > 0))
        return fail_io(c, dir + "/projection.sav: not a projection checkpoint");
    if (dim != D || ns_state != n)
        return fail_io(c, dir + "/projection.sav: dimension or particle count differs "
                                "from the .sav files");
    std::vector<float> vol((size_t)n);
    int32_t last = 0;
    if (!r.get_array(vol.data(), vol.size()) || !r.get(last))
        return fail_io(c, dir + "/projection.sav: truncated");
    // positions and velocities in caller order (a fresh upload: identity ids)
    c->initialized = false;
    VPS_TRY(vps_set_particles(c, n, (const double*)xb.data(), (const double*)vb.data()));
    {   // frozen volumes into the velocity records' w
        std::vector<float4> vel;
        VPS_TRY(pull(c, vel, c->rvel[c->cur].p, (size_t)n));
        std::vector<int32_t> id;
        VPS_TRY(pull(c, id, c->rid[c->cur].p, (size_t)n));
        for (int64_t k2 = 0; k2 < n; k2++) vel[(size_t)k2].w = vol[(size_t)id[(size_t)k2]];
        VPS_TRY(push(c, c->rvel[c->cur].p, vel));
    }
    c->rho0 = rho0;
    c->kappa = kappa;
    c->alpha0 = alpha0;
    c->cg_drift = drift;   // after vps_set_particles, which resets it
    c->k.rho0 = (float)rho0;
    c->k.kappa = (float)kappa;
    c->last_solve_buf = -1;
    c->last_solve_nvirt = 0;
    c->prev_iters = 0;
    c->nvirt = 0;
    c->nslots = 0;
    c->which = 0;
    int64_t ns = 0;
    if (last) {
        int64_t nvl = 0;
        int32_t piters = 0;
        if (!r.get(ns) || !r.get(nvl) || !r.get(piters) || ns <= 0 || ns > c->nbricks ||
            nvl < 0 || nvl > ns * c->brick_nodes)
            return fail_io(c, dir + "/projection.sav: bad last-solve header");
        std::vector<int32_t> sbk((size_t)ns);
        std::vector<float> xs((size_t)(ns * c->brick_nodes)), pr((size_t)nvl);
        if (!r.get_array(sbk.data(), sbk.size()) || !r.get_array(xs.data(), xs.size()) ||
            !r.get_array(pr.data(), pr.size()))
            return fail_io(c, dir + "/projection.sav: truncated last solve");
        // the solved layout goes into slot map 0; the next generation uses 1
        std::vector<int32_t> map((size_t)c->nbricks, -1);
        for (int64_t s = 0; s < ns; s++) {
            const int32_t b = sbk[(size_t)s];
            if (b < 0 || b >= c->nbricks || map[(size_t)b] >= 0)
                return fail_io(c, dir + "/projection.sav: bad slot map");
            map[(size_t)b] = (int32_t)s;
        }
        VPS_TRY(push(c, c->brick_slot[0].p, map));
        VPS_TRY(push(c, c->slot_brick[0].p, sbk));
        VPS_TRY(push(c, c->vx[0].p, xs));
        VPS_TRY(push(c, c->p_by_rank[0].p, pr));
        c->last_solve_buf = 0;
        c->last_solve_nvirt = nvl;
        c->prev_iters = piters;
        c->nslots = ns;   // grid hint of the next generation
    }
    int64_t mlen = 0;
    if (!r.get(mlen) || mlen < 0 || (int64_t)(r.end - r.p) < mlen)
        return fail_io(c, dir + "/projection.sav: truncated learned state");
    VPS_TRY(mlp_state_load(c, last ? 0 : -1, ns, r.p, (size_t)mlen));
    c->initialized = init != 0;
    c->generated = false;
    return VPS_OK;
}

}  // namespace vps

using namespace vps;

extern "C" {

int vps_write_frame(vps_context* c, const char* frame_dir) {
    if (!c || !frame_dir) return VPS_ERR_INVALID_ARG;
    if (c->n == 0) return fail(c, VPS_ERR_STATE, "write frame: no particles");
    return write_frame(c, frame_dir);
}

int vps_save_frame(vps_context* c, const char* frame_dir) {
    if (!c || !frame_dir) return VPS_ERR_INVALID_ARG;
    if (c->n == 0) return fail(c, VPS_ERR_STATE, "save frame: no particles");
    return save_frame(c, frame_dir);
}

int vps_load_frame(vps_context* c, const char* frame_dir) {
    if (!c || !frame_dir) return VPS_ERR_INVALID_ARG;
    return load_frame(c, frame_dir);
}

int vps_write_description(int32_t dim, const char* output_dir) {
    if ((dim != 2 && dim != 3) || !output_dir) return VPS_ERR_INVALID_ARG;
    // the three point-list objects of DualParticleSphLiquid::writeDescription,
    // albedo (52, 108, 156, 255) / 255
    std::string y = "dimension: " + std::to_string(dim) + "\nobjects:\n";
    const char* names[3] = {"particles", "virtual_particles", "boundary_particles"};
    const bool cmap[3] = {true, true, false};
    char albedo[128];
    std::snprintf(albedo, sizeof(albedo), "[%.9g, %.9g, %.9g, %.9g]", 52.f / 255.f,
                  108.f / 255.f, 156.f / 255.f, 255.f / 255.f);
    for (int i = 0; i < 3; i++) {
        y += std::string("  - name: ") + names[i] + "\n";
        y += "    data_mode: dynamic\n";
        y += "    primitive_type: point_list\n";
        y += std::string("    material:\n      diffuse_albedo: ") + albedo + "\n";
        y += "    indexed: false\n";
        y += std::string("    color_map:\n      enabled: ") + (cmap[i] ? "true" : "false") + "\n";
    }
    std::error_code ec;
    std::filesystem::create_directories(output_dir, ec);
    if (!write_bytes(std::string(output_dir) + "/description.yaml", y.data(), y.size()))
        return VPS_ERR_IO;
    return VPS_OK;
}

int vps_write_end_frame(const char* output_dir, int32_t frame) {
    if (!output_dir || frame < 0) return VPS_ERR_INVALID_ARG;
    const std::string s = std::to_string((int64_t)frame + 1) + "\n";
    if (!write_bytes(std::string(output_dir) + "/end_frame.txt", s.data(), s.size()))
        return VPS_ERR_IO;
    return VPS_OK;
}

int vps_write_and_save_frame(vps_context* c, const char* output_dir, int32_t frame) {
    if (!c || !output_dir || frame < 0) return VPS_ERR_INVALID_ARG;
    const std::string dir = std::string(output_dir) + "/" + std::to_string(frame);
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);
    if (ec) return fail(c, VPS_ERR_IO, "cannot create " + dir + ": " + ec.message());
    VPS_TRY(vps_write_frame(c, dir.c_str()));
    VPS_TRY(vps_save_frame(c, dir.c_str()));
    if (vps_write_end_frame(output_dir, frame) != VPS_OK)
        return fail(c, VPS_ERR_IO, std::string("cannot write ") + output_dir + "/end_frame.txt");
    return VPS_OK;
}

int vps_load_frame_number(vps_context* c, const char* output_dir, int32_t frame) {
    if (!c || !output_dir || frame < 0) return VPS_ERR_INVALID_ARG;
    const std::string dir = std::string(output_dir) + "/" + std::to_string(frame);
    if (!std::filesystem::exists(dir))
        return fail(c, VPS_ERR_IO, "no output directory for frame " + std::to_string(frame));
    return vps_load_frame(c, dir.c_str());
}

}  // extern "C"
