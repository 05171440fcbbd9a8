// Driver for the sanitizer build of the host C++ (tests/test_sanitizers.py):
// the config parser and scene set-up, the STL loader and build_mesh
// (zdx_host.cpp), the output writers and polyline.py helpers
// (zdx_output.cpp), the per-pair geometry and topology that the kernels run,
// compiled for the host (zdx_geom.h / zdx_topo.h through hosttest.cpp), and
// the oracle's time step.  Built with g++ -fsanitize=address,undefined
// -fno-sanitize-recover=all: any out-of-bounds access, use after free, leak or
// undefined behaviour ends the run with a report and a non-zero status.
//
//   sanitize_main OUT_DIR CONFIG... -- STL...
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../oracle/zdx_oracle.h"
#include "../../zedm-x_amd/csrc/zdx_internal.h"
#include "hosttest.cpp"

using namespace zdx;

static int fail(const char* what, const std::string& arg) {
    std::fprintf(stderr, "sanitize_main: %s failed for %s\n", what, arg.c_str());
    return 1;
}

static std::string dir_of(const std::string& p) {
    const size_t k = p.find_last_of('/');
    return k == std::string::npos ? std::string(".") : p.substr(0, k);
}

int main(int argc, char** argv) {
    if (argc < 2) return fail("arguments", "");
    const std::string out = argv[1];
    std::vector<std::string> configs, stls;
    bool after = false;
    for (int i = 2; i < argc; ++i) {
        if (std::strcmp(argv[i], "--") == 0) { after = true; continue; }
        (after ? stls : configs).push_back(argv[i]);
    }
    // configs: parse, build the scene, write both output files of the first
    // step, run the oracle's step on the same file
    int nfile = 0;
    for (const auto& c : configs) {
        SimConfig cfg;
        if (!parse_config(c, cfg)) return fail("parse_config", c);
        ConfigScene sc;
        std::string err;
        if (!build_scene(cfg, dir_of(c), sc, err)) return fail("build_scene", c + ": " + err);
        const int n = (int)sc.inits.size();
        std::vector<double> state(16 * (size_t)n, 0.0), mass(n), radius(n), F(3 * (size_t)n, 0.5), T(3 * (size_t)n, -0.25);
        std::vector<int32_t> counts(n, 1);
        std::vector<const HostMesh*> mp(n);
        for (int i = 0; i < n; ++i) {
            const ParticleInit& p = sc.inits[i];
            const HostMesh& m = sc.meshes[sc.mesh_of[i]];
            ParticleConst pc = particle_const(m, sc.mesh_of[i], p.density, p.young, p.poisson, p.mu, p.rest);
            double* s = &state[16 * (size_t)i];
            for (int k = 0; k < 3; ++k) { s[k] = p.pos[k]; s[7 + k] = p.vel[k]; s[10 + k] = p.omega[k]; }
            for (int k = 0; k < 4; ++k) s[3 + k] = p.quat[k];
            initial_L(pc, p.quat, p.omega, s + 13);
            mass[i] = pc.mass;
            radius[i] = pc.radius;
            mp[i] = &m;
        }
        const std::string base = out + "/s" + std::to_string(nfile++);
        if (!write_vtk_particles(base + ".vtk", n, mp.data(), state.data(), mass.data(), radius.data(), F.data(),
                                 T.data(), counts.data()))
            return fail("write_vtk_particles", c);
        if (!write_particle_state_txt(base + ".txt", n, state.data(), sc.inits.data())) return fail("state txt", c);
        orc_sim* o = orc_sim_from_config(c.c_str(), dir_of(c).c_str());
        if (!o) return fail("orc_sim_from_config", c);
        int64_t ctr[7];
        orc_sim_step(o, 20, ctr);
        orc_sim_destroy(o);
    }
    // meshes: load, build (centred and not), run the pair pipeline on each
    // mesh against itself at a few overlaps, and its sensitivities
    std::vector<std::vector<double>> meshes;
    for (const auto& s : stls) {
        std::vector<double> tris;
        if (!load_stl(s, tris)) return fail("load_stl", s);
        HostMesh a = build_mesh(tris, true);
        HostMesh b = build_mesh(tris, false);
        if (a.tri_idx.size() != b.tri_idx.size()) return fail("build_mesh", s);
        meshes.push_back(a.tris);
        const int T = (int)(a.tris.size() / 9);
        const double qa[4] = {1, 0, 0, 0}, qb[4] = {0.9, 0.2, -0.3, 0.1};
        for (double f : {0.95, 1.6, 1.9}) {
            const double pa[3] = {0, 0, 0}, pb[3] = {f * a.radius, 0.1 * a.radius, -0.05 * a.radius};
            std::vector<double> seg(6 * 20000), rec(28 * 512), vtx(3 * 20000);
            long long cnt[7];
            ht_pair(a.tris.data(), T, a.tris.data(), T, pa, qa, pb, qb, a.mean_edge, a.bbox_diag, 1, seg.data(), 20000,
                    rec.data(), 512, vtx.data(), 20000, cnt);
        }
    }
    // polyline.py helpers and writers
    const double vals[] = {0.0, -0.0, 1.0, 0.1, 1e16, 1e-4, 9.999e-5, 123456789.0, 5e-324, 1.7976931348623157e308};
    for (double v : vals)
        if (py_float_repr(v).empty()) return fail("py_float_repr", "");
    const double sq[12] = {0, 0, 0, 1, 0, 0, 1, 1, 0, 0, 1, 0};
    PolylineContact pc = polyline_contact(sq, 4, 2.0);
    if (!(pc.area > 0.99 && pc.area < 1.01) || !polyline_keeps_loop(sq, 4, 0.1)) return fail("polyline_contact", "");
    double moved[12];
    const double t[3] = {0.5, -1, 2}, q[4] = {0.8, 0.2, 0.1, -0.3};
    polyline_transform_points(sq, 4, t, q, moved);
    if (!meshes.empty() &&
        !write_vtk_polydata_tris(out + "/tris.vtk", meshes[0].data(), (int64_t)(meshes[0].size() / 9), "sanitize"))
        return fail("write_vtk_polydata_tris", "");
    std::vector<std::vector<V3>> loops(1);
    for (int k = 0; k < 4; ++k) loops[0].push_back(v3(moved[3 * k], moved[3 * k + 1], moved[3 * k + 2]));
    std::vector<V3> cp = {v3(0, 0, 0)};
    if (!write_vtk_polydata_polylines(out + "/loops.vtk", loops, cp, std::vector<char>{1}, std::vector<int>{0}))
        return fail("write_vtk_polydata_polylines", "");
    std::printf("sanitize_main ok: %zu configs, %zu meshes\n", configs.size(), stls.size());
    return 0;
}
