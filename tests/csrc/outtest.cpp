// Host build of the product's output writers (zedm-x_amd/csrc/zdx_output.cpp)
// and config set-up (zdx_host.cpp), driven from tests/test_output.py with
// state, forces and contact counts taken from the oracle's time-stepper.
#include <string>
#include <vector>

#define ZDX_HD
#include "../../zedm-x_amd/csrc/zdx_geom.h"
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

// polyline.py's float text, transform, per-loop contact and VTK writers
// (zedm-x_amd/csrc/zdx_output.cpp), driven from tests/test_polyline_cli.py.
extern "C" int ot_py_repr(double x, char* buf, int cap) {
    const std::string s = py_float_repr(x);
    if ((int)s.size() + 1 > cap) return -1;
    for (size_t k = 0; k < s.size(); ++k) buf[k] = s[k];
    buf[s.size()] = 0;
    return (int)s.size();
}

extern "C" void ot_transform(const double* in, long long n, const double* t, const double* q, double* out) {
    polyline_transform_points(in, n, t, q, out);
}

// out: Sn(3) Gn(3) area nA(3) xc0(3) Fn FA(3) has_xc0 = 18 doubles
extern "C" void ot_contact(const double* pts, long long m, double k, double* out) {
    const PolylineContact c = polyline_contact(pts, m, k);
    double* o = out;
    for (int d = 0; d < 3; ++d) *o++ = c.Sn[d];
    for (int d = 0; d < 3; ++d) *o++ = c.Gn[d];
    *o++ = c.area;
    for (int d = 0; d < 3; ++d) *o++ = c.nA[d];
    for (int d = 0; d < 3; ++d) *o++ = c.xc0[d];
    *o++ = c.Fn;
    for (int d = 0; d < 3; ++d) *o++ = c.FA[d];
    *o++ = c.has_xc0 ? 1.0 : 0.0;
}

extern "C" int ot_write_tris(const char* path, const double* tris, long long T, const char* comment) {
    return write_vtk_polydata_tris(path, tris, T, comment) ? 0 : 1;
}

// loops: n_loops loops, loop l has points [off[l], off[l+1]) of pts (3 doubles each);
// n_cp contact points with has[k] = 0 for None
extern "C" int ot_write_loops(const char* path, int n_loops, const int* off, const double* pts, int n_cp,
                              const double* cp, const int* has, const int* ids, int n_ids) {
    std::vector<std::vector<V3>> loops(n_loops);
    for (int l = 0; l < n_loops; ++l)
        for (int v = off[l]; v < off[l + 1]; ++v) loops[l].push_back(v3(pts[3 * v], pts[3 * v + 1], pts[3 * v + 2]));
    std::vector<V3> c;
    std::vector<char> h;
    for (int k = 0; k < n_cp; ++k) {
        c.push_back(v3(cp[3 * k], cp[3 * k + 1], cp[3 * k + 2]));
        h.push_back(has[k] ? 1 : 0);
    }
    return write_vtk_polydata_polylines(path, loops, c, h, std::vector<int>(ids, ids + n_ids)) ? 0 : 1;
}

extern "C" int ot_keeps_loop(const double* pts, long long m, double tol) { return polyline_keeps_loop(pts, m, tol); }
