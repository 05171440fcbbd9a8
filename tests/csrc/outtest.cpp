// Host build of the product's output writers (zedm-x_amd/csrc/zdx_output.cpp)
// and config set-up (zdx_host.cpp), driven from tests/test_output.py with
// state, forces and contact counts taken from the oracle's time-stepper.
#include <string>
#include <vector>

#include "../../zedm-x_amd/csrc/zdx_internal.h"

using namespace zdx;

extern "C" int ot_write(const char* config, const char* base_dir, int n, const double* state, const double* F,
                        const double* T, const int* counts, const char* vtk_path, const char* txt_path) {
    SimConfig cfg;
    if (!parse_config(config, cfg)) return 1;
    ConfigScene sc;
    std::string err;
    if (!build_scene(cfg, base_dir ? base_dir : "", sc, err)) return 2;
    if ((int)sc.inits.size() != n) return 3;
    std::vector<const HostMesh*> mesh(n);
    std::vector<double> mass(n), radius(n);
    for (int i = 0; i < n; ++i) {
        const ParticleInit& in = sc.inits[i];
        const HostMesh& m = sc.meshes[sc.mesh_of[i]];
        ParticleConst pc = particle_const(m, sc.mesh_of[i], in.density, in.young, in.poisson, in.mu, in.rest);
        mesh[i] = &m;
        mass[i] = pc.mass;
        radius[i] = pc.radius;
    }
    if (!write_vtk_particles(vtk_path, n, mesh.data(), state, mass.data(), radius.data(), F, T, counts)) return 4;
    if (!write_particle_state_txt(txt_path, n, state, sc.inits.data())) return 5;
    return 0;
}
